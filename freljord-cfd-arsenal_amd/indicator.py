"""Training and evaluation of the learned troubled-cell indicator on the GPU.

    python freljord-cfd-arsenal_amd/indicator.py --out <weights.npz> [--n 256] [--ics 64]

1. Seeded randomised initial conditions: MHD rotors of the init4.txt
   family (disk radius r0 in U[0.08, 0.12], r1 = r0 + 0.015, spin in
   U[0.75, 1.25], rho 10 in / 1 out, p = 0.5, Bx = 2.4/sqrt(4 pi)),
   four-quadrant Riemann problems (rho in U[0.5, 2], p in U[0.1, 2],
   velocities in U[-0.75, 0.75], uniform B) and perturbed Orszag-Tang states
   (random phases and amplitude).  The stronger rotor (spin 2, p = 1,
   Bx = 5/sqrt(4 pi)) is not used: the scheme itself (the reference and its
   CPU restatement alike) loses positivity on it and blows up after a few
   hundredths of a time unit at 128^2, from 1e-13 perturbations of the
   initial data.  Snapshots whose pressure minimum falls below -1 are
   skipped for the same reason.
2. Each is advanced with the plain FP64 method (TVB gated by KXRCF,
   positivity on).  At several snapshots the state is copied into a probe
   context, advanced by one unlimited RK stage (the limiter's input), and its
   cells are labelled with the plain method's decision on that input: KXRCF
   flags the cell AND TVB changes it.  Every positive and a random fraction
   of the negatives go into the device dataset (mhd2d_trainer_collect).
3. FP32 training on the GPU (mhd2d_trainer_fit: Adam over random batches).
4. Held-out initial conditions (including the benchmark flow through its
   shocks at 1024^2, t = 0.6, and a 2048^2 Orszag-Tang state) give the two
   thresholds.  mlp_lo: the cells whose 5-cell stencil maximum of p reaches
   it are handed to KXRCF (the one-cell dilation of the screen), and it is
   the largest value that keeps the missed labels within the budget.
   mlp_hi: cells with p >= mlp_hi are limited on the network's decision,
   without KXRCF; it is the smallest value at which the held-out cells above
   it are labelled in at least --hi-precision of the cases (1.0, never
   deciding, if none is).  The held-out statistics (mhd2d_indicator_stats)
   give the missed-shock and over-flag rates.
   --evaluate <weights.npz> repeats step 4's statistics for saved weights.
Resolutions: rotors and Riemann problems at --n; perturbed Orszag-Tang
states also at 2x, 4x and 8x that (the benchmark runs the flow at 2048^2),
with the negative sampling rate scaled as 256 / N.  The TVB constant M is
drawn per initial condition (random_ic), which covers the TVB thresholds of
finer grids than the training ones.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _pkg():
    name = "freljord_cfd_arsenal_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def init_params(seed=0):
    """He-style initialisation (same layout as MHD2D_MLP_NPARAM)"""
    rng = np.random.default_rng(seed)
    nin, nh = 60, 64
    W1 = rng.standard_normal((nin, nh)) * np.sqrt(2.0 / nin)
    W2 = rng.standard_normal((nh, nh)) * np.sqrt(2.0 / nh)
    w3 = rng.standard_normal(nh) * np.sqrt(1.0 / nh)
    return np.concatenate([W1.ravel(), np.zeros(nh), W2.ravel(), np.zeros(nh), w3, [0.0]]).astype(np.float32)


def random_ic(rng, kind, N=256, tmax=None):
    """(ic, extra, domain, bc, tend, M) of one seeded randomised initial
    condition.  The TVB constant M varies: what TVB sees is M h^2 against the
    local modes, so a coarse grid with a small M stands for a fine grid with
    the benchmark's M = 50 (M h^2 at 4096^2 with M = 50 is M h^2 at 512^2
    with M = 0.8)."""
    P = _pkg()
    if kind == "rotor":
        r0 = rng.uniform(0.08, 0.12)
        spin = rng.uniform(0.75, 1.25)
        xc, yc = rng.uniform(0.45, 0.55, 2)
        extra = [r0, r0 + 0.015, spin, 0.5, 2.4 / np.sqrt(4 * np.pi), 10.0, xc, yc]
        return P.IC_ROTOR, extra, (0.0, 1.0, 0.0, 1.0), 1, rng.uniform(0.05, 0.15), float(rng.choice([2.0, 10.0, 50.0]))
    if kind == "riemann":
        x0, y0 = rng.uniform(0.35, 0.65, 2)
        bx, by = rng.uniform(-0.5, 0.5, 2)
        extra = [x0, y0]
        for _ in range(4):
            extra += [rng.uniform(0.5, 2.0), rng.uniform(-0.75, 0.75), rng.uniform(-0.75, 0.75),
                      rng.uniform(0.1, 2.0), bx, by]
        return P.IC_RIEMANN4, extra, (0.0, 1.0, 0.0, 1.0), 2, rng.uniform(0.05, 0.2), float(rng.choice([2.0, 10.0, 50.0]))
    amp = rng.uniform(0.0, 0.05)
    seed = int(rng.integers(1, 2**31))
    # the finer grids are run to earlier times (cost grows as N^3): the whole
    # shock formation up to 512^2, t <= 0.6 at 1024^2, t <= 0.3 at 2048^2
    if tmax is None:
        tmax = 1.2 * min(1.0, 512.0 / N)
    M = float(rng.choice([1.0, 3.0, 10.0, 50.0])) if N <= 512 else 50.0
    return P.IC_OT_PERTURBED, [amp, float(seed)], (0.0, 2 * np.pi, 0.0, 2 * np.pi), 1, rng.uniform(0.3, 1.0) * tmax, M


def snapshots(ic, N, nsnap, on_state, cfl=0.2, band=(0.2, 0.8)):
    """advance one initial condition with the plain method and call
    on_state(probe) with the limiter input of the first stage at nsnap times"""
    P = _pkg()
    kind, extra, dom, bc, tend, M = ic
    kw = dict(flux_type=P.HLLD, bc=bc, domain=dom, M=M, CFL=cfl, pp=1, indicator=P.IND_KXRCF, precision=P.FP64,
              Ck=1.0, mlp_band=band)
    s = P.Solver(N, N, **kw)
    probe = P.Solver(N, N, **kw)
    s.init_data(kind, extra)
    s.rk3_prologue()
    times = np.linspace(0.0, tend, nsnap + 1)[1:]
    for tk in times:
        t = s.get_time()[0]
        while t < tk:
            t = s.step(tk, sync=True)
        d = s.diagnostics()
        if not (np.isfinite(d["umax"]) and np.isfinite(d["pmin"]) and d["pmin"] > -1.0):
            print(f"[indicator] ic {kind} stopped at t={t:.4f}: pmin {d['pmin']:.3e}", flush=True)
            break
        uh, Bx, By = s.download()
        probe.upload(uh, Bx, By)
        probe.set_bc()
        probe.Lh()
        probe.LhEz()
        dt = probe.calculate_dt()
        probe.rk_stage(1, dt)        # UI = S + dt L(S): what the stage-1 limiter sees
        on_state(probe)
    s.close()
    probe.close()


def limiter_input_stats(solver):
    """indicator statistics on a limiter input: advances `solver` by one
    unlimited stage (set_bc, Lh, LhEz, dt, UI = S + dt L(S)), the state the
    stage-1 limiter sees, and evaluates the screen and the plain method
    there.  The solver's state is left at that stage value."""
    solver.set_bc()
    solver.Lh()
    solver.LhEz()
    solver.rk_stage(1, solver.calculate_dt())
    return solver.indicator_stats()


def build_dataset(trainer, ics, nsnap, neg_keep, seed, log=print):
    """ics: list of (N, ic); the negative sampling rate is neg_keep at 256^2
    and scales as 256 / N, so the finer grids add more rows per snapshot
    (their smooth cells are the ones the screen must learn to skip)"""
    t0 = time.time()
    for k, (N, ic) in enumerate(ics):
        keep = neg_keep * (256.0 / N)

        def on_state(probe, k=k, keep=keep):
            trainer.collect(probe, keep, seed * 100003 + k * 101 + int(1e3 * probe.get_time()[0]))
        snapshots(ic, N, nsnap, on_state)
        n, npos = trainer.size()
        log(f"[indicator] data: ic {k + 1}/{len(ics)} (kind {ic[0]}, {N}^2): {n} rows, {npos} positive, "
            f"{time.time() - t0:.1f} s")
    return trainer.size()


def stencil_max(p, periodic=True):
    """max of p over each cell's 5-cell stencil (the screen's dilation)"""
    m = p.copy()
    for ax in (0, 1):
        for sh in (1, -1):
            r = np.roll(p, sh, axis=ax)
            if not periodic:
                if ax == 0:
                    r[0 if sh == 1 else -1, :] = 0.0
                else:
                    r[:, 0 if sh == 1 else -1] = 0.0
            m = np.maximum(m, r)
    return m


def evaluate(params, ics, nsnap, lo, hi=1.0, with_raw=False):
    """indicator statistics summed over the snapshots of held-out ICs with the
    band [lo, hi), and (stencil max of p, label[, p]) of every cell for
    choosing the thresholds"""
    P = _pkg()
    tot = dict.fromkeys(P.STATS_KEYS, 0)
    ps, labels = [], []

    praw = []
    for N, ic in ics:
        tr = P.Trainer(capacity=N * N, max_batch=64)

        def on_state(probe, tr=tr, periodic=ic[3] == 1):
            probe.mlp_set_weights(params)
            st = probe.indicator_stats()
            for key in tot:
                tot[key] += st[key]
            p = probe.mlp_probability()
            pm = stencil_max(p, periodic)
            tr.clear()
            tr.collect(probe, 1.0, 0)
            _, Y, idx = tr.get_data()
            ps.append(pm.ravel(order="F")[idx])
            praw.append(p.ravel(order="F")[idx])
            labels.append(Y)

        snapshots(ic, N, nsnap, on_state, band=(lo, hi))
        tr.close()
    if with_raw:
        return tot, np.concatenate(ps), np.concatenate(labels), np.concatenate(praw)
    return tot, np.concatenate(ps), np.concatenate(labels)


def choose_lo(pm, y, miss_budget=1e-3):
    """the largest threshold that leaves at most miss_budget of the positive
    cells' stencil maxima below it"""
    pos = np.sort(pm[y > 0.5])
    if not len(pos):
        return 0.5
    return float(min(pos[int(miss_budget * len(pos))], 0.5))


def choose_hi(p, y, lo, precision=0.99, min_cells=200):
    """the smallest threshold at which the network decides cells itself: the
    cells with p >= hi must be labelled (the plain method limits them) in at
    least `precision` of the cases, over at least min_cells held-out cells;
    1.0 (never decide) when no threshold reaches that"""
    order = np.argsort(-p)
    ps, ys = p[order], y[order] > 0.5
    tp = np.cumsum(ys)
    n = np.arange(1, len(ps) + 1)
    ok = (n >= min_cells) & (tp >= precision * n) & (ps >= max(lo, 0.5))
    if not ok.any():
        return 1.0
    k = np.nonzero(ok)[0].max()
    return float(ps[k])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", help="where the trained weights go (required unless --evaluate)")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--ics", type=int, default=64)
    ap.add_argument("--snap", type=int, default=6)
    ap.add_argument("--neg-keep", type=float, default=0.06)
    ap.add_argument("--steps", type=int, default=120000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--pos-weight", type=float, default=1.0)
    ap.add_argument("--hi-precision", type=float, default=0.99,
                    help="label precision the cells the network decides by itself must reach on the held-out set")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--holdout", type=int, default=6)
    ap.add_argument("--evaluate", metavar="WEIGHTS", help="only re-evaluate saved weights on the held-out set")
    args = ap.parse_args()
    if not args.out and not args.evaluate:
        ap.error("--out is required")
    P = _pkg()
    rng = np.random.default_rng(args.seed)
    kinds = ["rotor", "riemann", "ot"]
    N = args.n
    ics = [(N, random_ic(rng, kinds[k % 3], N)) for k in range(args.ics)]
    for m, cnt in ((2, 8), (4, 3), (8, 2)):   # perturbed Orszag-Tang on finer grids
        ics += [(N * m, random_ic(rng, "ot", N * m)) for _ in range(cnt)]
    held = [(N, random_ic(rng, kinds[k % 3], N)) for k in range(args.holdout)]
    # the benchmark flow through its shocks at 1024^2, and at 2048^2
    P_ = _pkg()
    held += [(N * 4, (P_.IC_OT_PERTURBED, [1e-2, 20261203.0], (0.0, 2 * np.pi, 0.0, 2 * np.pi), 1, 0.6, 50.0))]
    held += [(N * 8, random_ic(rng, "ot", N * 8, tmax=0.3))]
    if args.evaluate:
        with np.load(args.evaluate, allow_pickle=False) as z:
            params, lo, hi = np.asarray(z["params"], dtype=np.float32), float(z["mlp_lo"]), float(z["mlp_hi"])
        stats, _, _ = evaluate(params, held, 3, lo, hi)
        rep = {"mlp_lo": lo, "mlp_hi": hi, "heldout": stats, "missed_rate": stats["missed"] / max(stats["label"], 1),
               "missed_hydro_rate": stats["missed_hydro"] / max(stats["label"], 1),
               "candidate_fraction": stats["candidates"] / max(stats["cells"], 1)}
        print("[indicator] " + json.dumps(rep), flush=True)
        return
    tr = P.Trainer(capacity=8_000_000, max_batch=args.batch)
    n, npos = build_dataset(tr, ics, args.snap, args.neg_keep, args.seed)
    print(f"[indicator] dataset: {n} rows, {npos} positive", flush=True)
    params = init_params(args.seed)
    t0 = time.time()
    losses = []
    chunk = 1000
    for s0 in range(0, args.steps, chunk):
        lr = args.lr * 0.5 * (1.0 + np.cos(np.pi * s0 / args.steps))     # cosine decay per chunk
        params, l = tr.fit(params, min(chunk, args.steps - s0), args.batch, lr, args.pos_weight,
                           seed=args.seed * 7919 + s0)
        losses.append(float(l.mean()))
        print(f"[indicator] steps {s0 + len(l)}: mean loss {losses[-1]:.5f} ({time.time() - t0:.1f} s)", flush=True)
    train_s = time.time() - t0
    _, p_all, y_all, p_raw = evaluate(params, held, 3, 0.5, with_raw=True)
    lo = choose_lo(p_all, y_all)
    hi = choose_hi(p_raw, y_all, lo, precision=args.hi_precision)
    sure = p_raw >= hi
    hi_stats = {"heldout_rows_decided": int(sure.sum()), "decided_labelled": int((sure & (y_all > 0.5)).sum())}
    stats, _, _ = evaluate(params, held, 3, lo, hi)
    rep = {"rows": n, "positives": npos, "train_seconds": train_s, "losses": losses, "mlp_lo": lo, "mlp_hi": hi,
           "hi_precision": args.hi_precision, "hi_rows": hi_stats, "heldout": stats,
           "missed_rate": stats["missed"] / max(stats["label"], 1),
           "missed_hydro_rate": stats["missed_hydro"] / max(stats["label"], 1),
           "over_flag_rate": stats["over"] / max(stats["candidates"], 1),
           "candidate_fraction": stats["candidates"] / max(stats["cells"], 1)}
    print("[indicator] " + json.dumps(rep), flush=True)
    np.savez(args.out, params=params, mlp_lo=lo, mlp_hi=hi, report=json.dumps(rep))
    tr.close()


if __name__ == "__main__":
    main()
