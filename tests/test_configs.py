"""The workload configurations besides the benchmark line, on the GPU.

- CPU-runnable case: Orszag-Tang 128^2, HLL, TVB (M = 50) gated by KXRCF
  (Ck = 1), CFL 0.2, FP64 to t = 0.5 (234 steps): the fused GPU run against
  the oracle (itself bit-identical to the reference on the fixtures)
- plain method at 1024^2: seeded perturbed Orszag-Tang, HLLD, KXRCF-gated
  TVB, positivity, 500 steps: div B at roundoff, mass conserved, positive,
  and the mixed form with the learned screen within 1e-5 of it
- precision: the smooth vortex (init1.txt) at 64^2, 128^2, 256^2 to t = 0.1:
  P2's third order in the L2 error against the exact solution in FP64; the
  mixed form within 1e-5 of FP64 (relative L2, density).  The mixed form's
  own error against the exact solution carries an FP32 rounding floor of
  about 4e-6 (L2 over [-5, 5]^2, i.e. 4e-7 per unit area): it follows the
  FP64 error to 0.2 % at 64^2 and converges at third order while the
  truncation error dominates, and flattens once that error falls below the
  floor (256^2 here)
- the learned-indicator configuration at its size: the init4.txt rotor at
  2048^2, plain FP64 method against mixed + learned indicator over 60 steps
  (rates, density difference and step times recorded)
"""
import importlib
import numpy as np
import pytest

from conftest import interior, rel_l2

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


def test_cpu_case_vs_oracle(pkg, oracle_mod):
    O = oracle_mod
    N, dom = 128, (0.0, TWO_PI, 0.0, TWO_PI)
    o = O.Oracle(N, N, flux_type=O.HLL, domain=dom, M=50.0, CFL=0.2, pp=0, indicator=O.IND_KXRCF, Ck=1.0)
    o.init_data(O.IC_OT, [])
    s = pkg.Solver(N, N, flux_type=pkg.HLL, domain=dom, M=50.0, CFL=0.2, pp=0, indicator=pkg.IND_KXRCF, Ck=1.0)
    s.upload(o.uh, o.Bx, o.By)
    o.rk3_prologue()
    while o.rk3_step(0.5):
        pass
    t, n = s.RK3(0.5)
    sc = o.scalars()
    uh, Bx, By = s.download()
    err = rel_l2(interior(uh), interior(o.uh))
    print(f"[configs] OT 128^2 HLL KXRCF t=0.5: {n} steps, rel L2 vs oracle {err:.2e}")
    assert n == int(sc["count"]) and t == sc["t"] == 0.5
    assert err < 1e-12
    assert rel_l2(Bx, o.Bx) < 1e-12 and rel_l2(By, o.By) < 1e-12
    assert s.diagnostics()["totaldiv"] < 1e-12


def test_plain_1024_invariants_and_mixed(pkg):
    N, dom = 1024, (0.0, TWO_PI, 0.0, TWO_PI)
    z = np.load(pkg.WEIGHTS_PATH, allow_pickle=False)
    band = (float(z["mlp_lo"]), float(z["mlp_hi"]))
    res = {}
    for name, prec, ind in (("fp64", pkg.FP64, pkg.IND_KXRCF), ("mixed_nn", pkg.MIXED, pkg.IND_MLP)):
        s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=ind, precision=prec,
                       Ck=1.0, mlp_band=band)
        s.mlp_set_weights(pkg.load_indicator_weights())
        s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 20261203])
        s.rk3_prologue()
        m0 = interior(s.download()[0])[:, :, 0, 0].sum()
        for _ in range(500):
            s.step()
        uh = s.download()[0]
        res[name] = (uh, s.diagnostics(), s.get_time(), m0)
        s.close()
    uh, d, (t, _, n), m0 = res["fp64"]
    assert n == 500 and np.isfinite(uh).all()
    # the discrete divergence of the div-free reconstruction is roundoff,
    # eps |B| / h per quadrature point, and grows as a random walk over the
    # steps: 2e-12 after 500 steps here (the same in the oracle, which this
    # path matches per stage); without constrained transport it is O(1e-2)
    assert d["totaldiv"] < 1e-11 and d["pmin"] > 0
    assert abs(interior(uh)[:, :, 0, 0].sum() - m0) < 1e-12 * abs(m0)
    um, dm, tm, _ = res["mixed_nn"]
    err = rel_l2(interior(um)[:, :, :, 0], interior(uh)[:, :, :, 0])
    print(f"[configs] perturbed OT 1024^2, 500 steps (t = {t:.4f}): mixed + screen vs FP64 density rel L2 {err:.2e}, "
          f"div B {dm['totaldiv']:.1e}")
    assert abs(tm[0] - t) < 1e-6 * t
    assert err < 1e-5 and dm["totaldiv"] < 1e-11


def test_vortex_convergence(pkg):
    ps = importlib.import_module("freljord_cfd_arsenal_amd.precision_study")
    Ns = (64, 128, 256)
    rows = ps.study(Ns, 0.1)
    e64 = np.array([r["L2_fp64"][0] for r in rows])
    emx = np.array([r["L2_mixed"][0] for r in rows])
    dmx = [r["mixed_vs_fp64"]["rho"] for r in rows]
    print(f"[configs] vortex density L2: FP64 {e64}, orders {rows[1]['order_fp64'][0]:.2f} "
          f"{rows[2]['order_fp64'][0]:.2f}; mixed {emx}; mixed vs FP64 {dmx}")
    assert rows[1]["order_fp64"][0] > 2.7 and rows[2]["order_fp64"][0] > 2.7
    assert rows[1]["order_mixed"][0] > 2.5
    assert abs(emx[0] - e64[0]) < 1e-2 * e64[0]
    assert max(dmx) < 1e-5
    assert all(r["divB_fp64"] < 1e-12 and r["divB_mixed"] < 1e-12 for r in rows)


def test_vortex_precision_config(pkg):
    """the precision configuration at its own domain and end time: the
    smooth vortex on [-5, 5]^2 to t = 1.0 at 256^2, 512^2 and 1024^2 (the
    2048^2 row, 2.5 more minutes, is in the committed precision profile).
    FP64 converges at third order in density; the mixed form stays within
    1e-5 of FP64 in every field (measured 1.5e-6 in density and 5.5e-6 in
    Bx, By at 1024^2) and keeps div B at roundoff.  On this domain the
    vortex's tail is cut by the periodic boundary, which fixes the momentum
    and energy errors of both forms at 1.7-2.4e-4, and the mixed form's own
    density error sits on its FP32 floor (7e-6 to 1.5e-5), so neither order
    is asserted there."""
    ps = importlib.import_module("freljord_cfd_arsenal_amd.precision_study")
    rows = ps.study((256, 512, 1024), 1.0, domain=(-5.0, 5.0, -5.0, 5.0))
    for r in rows:
        print(f"[configs] vortex [-5,5]^2 t=1 N={r['N']}: L2 rho FP64 {r['L2_fp64'][0]:.3e} mixed "
              f"{r['L2_mixed'][0]:.3e}; mixed vs FP64 max {max(r['mixed_vs_fp64'].values()):.2e}; "
              f"div B {r['divB_fp64']:.1e} / {r['divB_mixed']:.1e}")
    assert rows[1]["order_fp64"][0] > 2.9 and rows[2]["order_fp64"][0] > 2.9
    assert all(max(r["mixed_vs_fp64"].values()) < 1e-5 for r in rows)
    assert all(r["divB_fp64"] < 1e-12 and r["divB_mixed"] < 1e-12 for r in rows)


def test_plain_1024_vs_oracle(pkg, oracle_mod):
    """the plain method's own configuration at its size: perturbed
    Orszag-Tang 1024^2, HLLD, TVB (M = 50) gated by KXRCF, positivity, CFL
    0.2, three fused GPU steps against three oracle steps from the same
    projected input (the oracle is bit-identical to the reference on every
    fixture): 1e-12 relative L2 in the state and the faces, equal t"""
    O = oracle_mod
    N, dom, steps = 1024, (0.0, TWO_PI, 0.0, TWO_PI), 3
    o = O.Oracle(N, N, flux_type=O.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=O.IND_KXRCF, Ck=1.0,
                 nthreads=16)
    o.init_data(O.IC_OT_PERTURBED, [1e-2, 20261203])
    s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=pkg.IND_KXRCF, Ck=1.0)
    s.upload(o.uh, o.Bx, o.By)
    o.rk3_prologue()
    s.rk3_prologue()
    for k in range(steps):
        print(f"[configs] 1024^2 oracle step {k + 1} of {steps}", flush=True)
        o.rk3_step(1e9)
        s.step()
    uh, Bx, By = s.download()
    err = rel_l2(interior(uh), interior(o.uh))
    print(f"[configs] perturbed OT 1024^2 HLLD+pp+KXRCF, {steps} steps: rel L2 vs oracle {err:.2e}")
    assert s.get_time()[0] == o.scalars()["t"]
    assert err < 1e-12 and rel_l2(Bx, o.Bx) < 1e-12 and rel_l2(By, o.By) < 1e-12
    assert np.array_equal(s.debug_get(pkg.TROUBLE) != 0, o.trouble != 0)


def test_combined_4096(pkg):
    """the combined workload: seeded perturbed Orszag-Tang 4096^2 (805 M DOF),
    HLLD, mixed precision with the learned screen (KXRCF on its candidates),
    200 graph-replayed RK3 steps, against the plain FP64 method over the same
    200 steps: relative L2 <= 1e-5 in every field, div B <= 1e-12 in the
    mixed form, mass conserved, finite state, and the screen's decisions no
    more than the plain method's plus the cells the network decides alone"""
    N, dom, steps = 4096, (0.0, TWO_PI, 0.0, TWO_PI), 200
    z = np.load(pkg.WEIGHTS_PATH, allow_pickle=False)
    band = (float(z["mlp_lo"]), float(z["mlp_hi"]))
    res = {}
    for name, prec, ind in (("fp64", pkg.FP64, pkg.IND_KXRCF), ("mixed_nn", pkg.MIXED, pkg.IND_MLP)):
        s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=ind, precision=prec,
                       Ck=1.0, mlp_band=band)
        s.mlp_set_weights(pkg.load_indicator_weights())
        s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 20261203])
        s.rk3_prologue()
        m0 = float(interior(s.download()[0])[:, :, 0, 0].sum())
        s.step_graph(0)                      # capture
        for k in range(0, steps, 50):
            s.step_graph(50)
            print(f"[configs] 4096^2 {name}: {k + 50} of {steps} steps", flush=True)
        uh = interior(s.download()[0])[:, :, :6, :].copy()
        d = s.diagnostics()
        res[name] = dict(uh=uh, d=d, t=s.get_time(), m0=m0, jump=s.face_jump(),
                         stats=s.indicator_stats() if name == "mixed_nn" else None,
                         decided=int((s.mlp_probability() >= band[1]).sum()) if name == "mixed_nn" else 0)
        s.close()
    a, b = res["mixed_nn"], res["fp64"]
    assert a["t"][2] == b["t"][2] == steps and abs(a["t"][0] - b["t"][0]) < 1e-6 * b["t"][0]
    w = np.sqrt(np.array([1.0, 1 / 3, 1 / 3, 4 / 45, 1 / 9, 4 / 45]))[None, None, :]
    errs = {}
    for n, nm in ((0, "rho"), (1, "rhou"), (2, "rhov"), (4, "E"), (5, "Bx"), (6, "By")):
        ref = b["uh"][:, :, :, n] * w
        errs[nm] = float(np.linalg.norm(a["uh"][:, :, :, n] * w - ref) / np.linalg.norm(ref))
    st = a["stats"]
    print(f"[configs] combined 4096^2, {steps} steps: rel L2 vs FP64 {errs}, div B {a['d']['totaldiv']:.2e} "
          f"(FP64 {b['d']['totaldiv']:.2e}), face jump {a['jump']:.1e}, screen {st}")
    assert max(errs.values()) <= 1e-5
    assert a["d"]["totaldiv"] <= 1e-12 and a["jump"] <= 1e-13
    for r in (a, b):
        assert np.isfinite(r["uh"]).all() and r["d"]["pmin"] > 0
        assert abs(r["uh"][:, :, 0, 0].sum() - r["m0"]) <= 1e-12 * abs(r["m0"])
    # the screen hands its band to KXRCF and limits the cells it is sure of
    # (p >= mlp_hi) itself: its decisions are the candidates KXRCF flags plus
    # those cells
    assert st["decision"] <= st["kxrcf"] + a["decided"] and st["decision"] <= st["candidates"]


def test_learned_indicator_rotor_2048(pkg):
    """the learned-indicator configuration at its size: the MHD rotor (the
    reference's init4: disk radius 0.1 tapered to 0.115, spin 2, density 10
    inside and 1 outside, p = 1, Bx = 5/sqrt(4 pi)) at 2048^2, HLLD,
    positivity, TVB (M = 10), CFL 0.15.  The plain FP64 method (KXRCF gating
    TVB) and the mixed form with the learned indicator run the same 60 steps:
    both stay finite and positive, the mixed + learned step is the faster,
    and the screen's rates on the plain run's next limiter input and the
    density difference are recorded with bars at about three times the
    measured values.  This is where the screen falls short: on that grid
    TVB's threshold M h^2 is 2.4e-6 and it changes 55 k cells of the disk's
    edge that the network, trained on 256^2 rotors, reads as smooth (about
    27 % of the plain method's decisions missed; density within 8e-5 of the
    plain method after 60 steps, where the benchmark flow stays within
    2e-7)"""
    import time
    ind = importlib.import_module("freljord_cfd_arsenal_amd.indicator")
    z = np.load(pkg.WEIGHTS_PATH, allow_pickle=False)
    band = (float(z["mlp_lo"]), float(z["mlp_hi"]))
    N, steps = 2048, 60
    res = {}
    for name, prec, indicator in (("plain", pkg.FP64, pkg.IND_KXRCF), ("mixed_nn", pkg.MIXED, pkg.IND_MLP)):
        s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=(0.0, 1.0, 0.0, 1.0), M=10.0, CFL=0.15, pp=1,
                       indicator=indicator, precision=prec, Ck=1.0, mlp_band=band)
        s.mlp_set_weights(pkg.load_indicator_weights())
        s.init_data(pkg.IC_ROTOR, [])
        s.rk3_prologue()
        s.step()
        s.sync()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            s.step()
        s.sync()
        ms = 1e3 * (time.perf_counter() - t0) / (steps - 1)
        uh = interior(s.download()[0])[:, :, :6, 0].copy()
        d = s.diagnostics()
        res[name] = dict(uh=uh, d=d, t=s.get_time(), ms=ms,
                         stats=ind.limiter_input_stats(s) if name == "plain" else None)
        print(f"[configs] rotor 2048^2 {name}: {steps} steps to t = {res[name]['t'][0]:.6f}, {ms:.2f} ms/step, "
              f"pmin {d['pmin']:.3e}", flush=True)
        s.close()
    a, b = res["mixed_nn"], res["plain"]
    w = np.sqrt(np.array([1.0, 1 / 3, 1 / 3, 4 / 45, 1 / 9, 4 / 45]))[None, None, :]
    err = float(np.linalg.norm((a["uh"] - b["uh"]) * w) / np.linalg.norm(b["uh"] * w))
    st = b["stats"]
    print(f"[configs] rotor 2048^2: density rel L2 mixed + learned vs plain {err:.3e}; screen on the plain "
          f"method's limiter input {st}; speed-up {b['ms'] / a['ms']:.2f}")
    for r in (a, b):
        assert np.isfinite(r["uh"]).all() and r["d"]["pmin"] > 0
    assert st["candidates"] <= 0.1 * st["cells"]
    assert st["missed"] <= 0.6 * st["label"] and st["missed_hydro"] <= st["missed"]
    assert err <= 2.5e-4
    assert a["ms"] < b["ms"]
