"""Where the mixed form leaves the plain FP64 form on the benchmark flow.

    python freljord-cfd-arsenal_amd/drift_study.py [--n 1024] [--out drift.json]

The seeded perturbed Orszag-Tang flow of the benchmark (16 velocity modes of
amplitude 1e-2, seed 20261203; HLLD, TVB M = 50 gated by KXRCF with Ck = 1,
positivity, CFL 0.2) is advanced at N^2 in three pairs of runs, each pair
from the same initial data (RK3.f90:58-174 as graph-replayed steps):
- plain FP64 against the mixed form: the relative L2 of the density means at
  t = 0.1 ... 0.5, the share of the squared error held by the worst 0.01 %,
  0.1 %, 1 % and 10 % of the cells, and at t = 0.3, 0.4, 0.5 the worst cells
  with their state and their distance to the fast sonic points
  (|u| - c_f,x) / c_f,x and (|v| - c_f,y) / c_f,y (eigenvalueMm.f90's fast speed);
- plain FP64 against plain FP64 with relative noise of 1e-7 put once on
  every mode of the six non-magnetic variables after the prologue (the
  mixed form's storage rounding, applied once instead of every stage);
- the limiter decisions (KXRCF flags, and KXRCF flags AND TVB changes, the
  plain method's decision, TVB_Limiter_P2.f90:36-103) of the plain and the
  mixed method on the FP64 state, on the same state with its higher modes
  rounded to FP32 (the mixed storage), and on the mixed state, at t = 0.4,
  0.45 and 0.5;
- the residual du = Lh(u) of both forms on one input (the FP64 state at
  t = 0.3 with the mixed storage's rounding), per mode, relative to the
  largest |du| of that mode: the mixed form's FP32 quadratures on their own.
"""
import argparse
import importlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from precision_study import _pkg  # noqa: E402

GAMMA = 5.0 / 3.0
TIMES = (0.1, 0.2, 0.3, 0.4, 0.45, 0.5)
NONMAG = (0, 1, 2, 3, 4, 7)          # rho, rho u, rho v, rho w, E, Bz


def _solver(pkg, N, prec):
    return pkg.Solver(N, N, flux_type=pkg.HLLD, domain=(0.0, 2 * np.pi, 0.0, 2 * np.pi), M=50.0, CFL=0.2, pp=1,
                      indicator=pkg.IND_KXRCF, precision=prec, Ck=1.0)


def _start(pkg, s):
    s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 20261203])
    s.rk3_prologue()


def _advance(s, tend):
    t = s.get_time()[0]
    while t < tend:
        s.step_graph(16, tend)
        t = s.get_time()[0]
    return t


def _rho_rel(a, b):
    a, b = a[1:-1, 1:-1, 0, 0], b[1:-1, 1:-1, 0, 0]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _fast(rho, p, bn, b2):
    """fast magnetosonic speed along the direction of bn (eigenvalueMm.f90)"""
    c2, a2, an2 = GAMMA * p / rho, b2 / rho, bn * bn / rho
    return np.sqrt(0.5 * (c2 + a2 + np.sqrt(np.maximum((c2 + a2) ** 2 - 4 * c2 * an2, 0.0))))


def _worst(F, e, k=6):
    """the k cells of largest |e| with their state and fast-sonic distances"""
    m = F[1:-1, 1:-1, 0]
    rho, u, v, w = m[..., 0], m[..., 1] / m[..., 0], m[..., 2] / m[..., 0], m[..., 3] / m[..., 0]
    bx, by, bz = m[..., 5], m[..., 6], m[..., 7]
    b2 = bx * bx + by * by + bz * bz
    p = (GAMMA - 1) * (m[..., 4] - 0.5 * rho * (u * u + v * v + w * w) - 0.5 * b2)
    out = []
    for idx in np.argsort(np.abs(e).ravel())[::-1][:k]:
        i, j = np.unravel_index(idx, e.shape)
        cx, cy = _fast(rho[i, j], p[i, j], bx[i, j], b2[i, j]), _fast(rho[i, j], p[i, j], by[i, j], b2[i, j])
        out.append({"cell": [int(i) + 1, int(j) + 1], "err": float(e[i, j]), "rho": float(rho[i, j]),
                    "u": float(u[i, j]), "v": float(v[i, j]), "Bx": float(bx[i, j]), "By": float(by[i, j]),
                    "p": float(p[i, j]), "x_sonic": float((abs(u[i, j]) - cx) / cx),
                    "y_sonic": float((abs(v[i, j]) - cy) / cy)})
    return out


def _mixed_storage(uh):
    u = uh.copy()
    for n in NONMAG:
        u[:, :, 1:6, n] = u[:, :, 1:6, n].astype(np.float32).astype(np.float64)
    return u


def _decisions(pkg, N, prec, state, t, weights):
    s = _solver(pkg, N, prec)
    s.mlp_set_weights(weights)
    s.upload(*state)
    s.set_time(t, 0)
    r = importlib.import_module(pkg.__name__ + ".indicator").limiter_input_stats(s)
    s.close()
    return {"kxrcf": r["kxrcf"], "label": r["label"]}


def _du(pkg, N, prec, state):
    s = _solver(pkg, N, prec)
    s.upload(*state)
    s.set_bc()
    s.Lh()
    d = s.debug_get(pkg.DU)[1:-1, 1:-1]
    s.close()
    return d


def run(N, log=print):
    pkg = _pkg()
    out = {"N": N}
    # 1. plain FP64 against the mixed form; the decisions along the way
    sF, sM = _solver(pkg, N, pkg.FP64), _solver(pkg, N, pkg.MIXED)
    for s in (sF, sM):
        _start(pkg, s)
    W = pkg.load_indicator_weights()
    rows, dec = [], []
    for tend in TIMES:
        _advance(sF, tend)
        _advance(sM, tend)
        F, Mx = sF.download(), sM.download()
        e = Mx[0][1:-1, 1:-1, 0, 0] - F[0][1:-1, 1:-1, 0, 0]
        e2 = np.sort((e * e).ravel())[::-1]
        share = [float(e2[:max(1, int(f * e2.size))].sum() / e2.sum()) for f in (1e-4, 1e-3, 1e-2, 1e-1)]
        row = {"t": tend, "rho_rel_l2": _rho_rel(Mx[0], F[0]), "error_share_top": share}
        if tend in (0.3, 0.4, 0.5):
            row["worst"] = _worst(F[0], e)
        rows.append(row)
        log(f"[drift] t={tend}: mixed vs FP64 rho {row['rho_rel_l2']:.2e}, share of the worst 0.1 % cells "
            f"{share[1]:.2f}")
        if tend in (0.4, 0.45, 0.5):
            Fr = (_mixed_storage(F[0]), F[1], F[2])
            d = {"t": tend}
            for name, prec, st in (("plain_on_fp64", pkg.FP64, F), ("plain_on_fp64_rounded", pkg.FP64, Fr),
                                   ("mixed_on_fp64", pkg.MIXED, F), ("plain_on_mixed", pkg.FP64, Mx),
                                   ("mixed_on_mixed", pkg.MIXED, Mx)):
                d[name] = _decisions(pkg, N, prec, st, tend, W)
            dec.append(d)
            log(f"[drift] t={tend}: decisions {d}")
        if tend == 0.3:
            Fr = (_mixed_storage(F[0]), F[1], F[2])
            d64, dmx = _du(pkg, N, pkg.FP64, Fr), _du(pkg, N, pkg.MIXED, Fr)
            res = {}
            for n, nm in ((0, "rho"), (1, "rhou"), (4, "E")):
                for m in (0, 1, 2):
                    a, b = dmx[:, :, m, n], d64[:, :, m, n]
                    r = np.abs(a - b) / np.abs(b).max()
                    res[f"{nm}_mode{m + 1}"] = {"median": float(np.median(r)), "q999": float(np.quantile(r, 0.999)),
                                                "max": float(r.max())}
            out["du_mixed_vs_fp64_t0.3"] = res
    out["mixed_vs_fp64"] = rows
    out["decisions"] = dec
    sF.close()
    sM.close()
    # 2. FP64 against FP64 with one-time relative noise of 1e-7
    rng = np.random.default_rng(7)
    sA, sB = _solver(pkg, N, pkg.FP64), _solver(pkg, N, pkg.FP64)
    for s in (sA, sB):
        _start(pkg, s)
    t, _, n = sB.get_time()
    uh, Bx, By = sB.download()
    for v in NONMAG:
        blk = uh[1:-1, 1:-1, 0:6, v]
        uh[1:-1, 1:-1, 0:6, v] = blk * (1.0 + 1e-7 * rng.uniform(-1.0, 1.0, size=blk.shape))
    sB.upload(uh, Bx, By)
    sB.set_time(t, n)
    noise = []
    for tend in TIMES:
        _advance(sA, tend)
        _advance(sB, tend)
        noise.append({"t": tend, "rho_rel_l2": _rho_rel(sB.download()[0], sA.download()[0])})
    log("[drift] FP64 + one-time 1e-7 noise vs FP64: " + ", ".join(f"t={r['t']} {r['rho_rel_l2']:.1e}" for r in noise))
    out["fp64_noise_vs_fp64"] = noise
    sA.close()
    sB.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    res = run(a.n, log=lambda s: print(s, file=sys.stderr, flush=True))
    txt = json.dumps(res, indent=1)
    if a.out:
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        with open(a.out, "w") as f:
            f.write(txt)
    print(txt)


if __name__ == "__main__":
    main()
