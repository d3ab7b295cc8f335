"""HIP path (libmhd2d.so through its C ABI) against the reference fixtures and
the CPU oracle.

Bar (plain FP64 form): relative L2 <= 1e-12 per reference subroutine and per
RK stage on identical inputs; whole multi-step runs <= 1e-12 as well.  The
GPU contracts multiply-adds into FMAs and divides through reciprocals, so it
is not bit-identical to the reference, only within rounding.
"""
import numpy as np
import pytest

from conftest import golden, interior, rel_l2, unpack

pytestmark = pytest.mark.gpu

TOL = 1e-12


def solver_from(pkg, g, uh_key="uh", ft=4, bc=None, M=None, pp=0, **kw):
    uh = unpack(g[uh_key])
    bc = int(g["bc"]) if bc is None and "bc" in g else (1 if bc is None else bc)
    M = float(g["M"]) if M is None and "M" in g else (10.0 if M is None else M)
    s = pkg.Solver(uh.shape[0] - 2, uh.shape[1] - 2, flux_type=ft, bc=bc, domain=tuple(g["domain"]),
                   M=M, pp=pp, **kw)
    return s, uh


@pytest.mark.parametrize("tag,ic", [("ot", 3), ("vortex", 1), ("rotor", 4)])
def test_init_data(pkg, tag, ic):
    g = golden(f"init_{tag}_16")
    s = pkg.Solver(16, 16, domain=tuple(g["domain"]))
    s.init_data(ic)
    uh, Bx, By = s.download()
    assert rel_l2(interior(uh), interior(unpack(g["uh"]))) < 1e-13
    assert rel_l2(Bx, g["Bx"]) < 1e-13 and rel_l2(By, g["By"]) < 1e-13


@pytest.mark.parametrize("name,ft", [("lh_f1_16", 1), ("lh_f2_16", 2), ("lh_f3_16", 3), ("lh_f4_16", 4),
                                     ("lh_f4_12x8_outflow", 4)])
def test_lh_lhez(pkg, name, ft):
    g = golden(name)
    s, uh = solver_from(pkg, g, ft=ft)
    s.upload(uh, g["Bx"], g["By"])
    s.set_bc()
    if "uh_bc" in g:
        u2, _, _ = s.download()
        assert np.array_equal(u2, unpack(g["uh_bc"]))
    s.Lh()
    s.LhEz()
    assert rel_l2(s.debug_get(pkg.DU), unpack(g["du"])) < TOL
    assert rel_l2(s.debug_get(pkg.DEZ1), g["dEz1"]) < TOL
    assert rel_l2(s.debug_get(pkg.DEZ2), g["dEz2"]) < TOL
    assert rel_l2(s.debug_get(pkg.EZVERTEX), g["EzVertex"]) < TOL


def test_lhez_requires_lh(pkg):
    g = golden("lh_f4_16")
    s, uh = solver_from(pkg, g)
    s.upload(uh, g["Bx"], g["By"])
    with pytest.raises(pkg.MHD2DError):
        s.LhEz()


@pytest.mark.parametrize("M", [0, 10])
def test_tvb_stale_halos(pkg, M):
    g = golden(f"tvb_M{M}_16")
    s, uh = solver_from(pkg, g)
    s.upload(uh, g["Bx"], g["By"])
    s.TVB_Limiter_P2()
    out, _, _ = s.download()
    ref = unpack(g["out"])
    assert rel_l2(out, ref) < TOL
    # the limiter's zeroed modes are zero on the GPU too
    assert np.array_equal(out[1:-1, 1:-1][ref[1:-1, 1:-1] == 0], ref[1:-1, 1:-1][ref[1:-1, 1:-1] == 0])


def test_pp_limiter(pkg):
    g = golden("pp_16")
    s, uh = solver_from(pkg, g)
    s.upload(uh, g["Bx"], g["By"])
    s.pp_Limiter()
    out, _, _ = s.download()
    assert rel_l2(out, unpack(g["out"])) < TOL


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_pp_limiter_near_bounds(pkg, oracle_mod, form):
    """random cells whose higher modes span 1e-4 to 2 times their averages,
    so that some cells are far inside the positivity limits, some close to
    them and some beyond: every cell against the oracle's pp_Limiter.  The
    kernel settles a cell by a bound on its point values before evaluating
    the points; a cell settled wrongly would keep modes the oracle scales by
    at most 1/2.  The mixed form stores the higher modes of the non-magnetic
    variables in FP32: both sides start from those rounded values, and the
    limited modes are stored rounded again (bar 1e-6 per cell)."""
    from conftest import ACTIVE
    N, gam = 32, 5.0 / 3.0
    rng = np.random.default_rng(20261203)
    uh = np.zeros((N + 2, N + 2, 10, 8), order="F")
    rho = rng.uniform(0.2, 2.0, (N + 2, N + 2))
    vel = rng.uniform(-1.0, 1.0, (N + 2, N + 2, 3))
    B = rng.uniform(-1.0, 1.0, (N + 2, N + 2, 3))
    p = rng.uniform(0.05, 1.0, (N + 2, N + 2))
    avg = np.zeros((N + 2, N + 2, 8))
    avg[..., 0] = rho
    avg[..., 1:4] = rho[..., None] * vel
    avg[..., 4] = p / (gam - 1) + 0.5 * rho * (vel ** 2).sum(-1) + 0.5 * (B ** 2).sum(-1)
    avg[..., 5:8] = B
    scale = 10.0 ** rng.uniform(-4.0, 0.3, (N + 2, N + 2))
    for d, n in ACTIVE:
        if d == 0:
            uh[:, :, 0, n] = avg[..., n]
        else:
            uh[:, :, d, n] = rng.normal(size=(N + 2, N + 2)) * scale * (np.abs(avg[..., n]) + 0.1)
    if form == "mixed":
        for n in (0, 1, 2, 3, 4, 7):
            uh[:, :, 1:6, n] = uh[:, :, 1:6, n].astype(np.float32).astype(np.float64)
    Bx = np.zeros((N + 1, N, 3))
    By = np.zeros((N, N + 1, 3))
    s = pkg.Solver(N, N, flux_type=4, bc=1, precision=pkg.MIXED if form == "mixed" else pkg.FP64)
    o = oracle_mod.Oracle(N, N, flux_type=4, bc=1)
    s.upload(uh, Bx, By)
    o.load(uh, Bx, By)
    s.pp_Limiter()
    o.pp_Limiter()
    out = interior(s.download()[0])
    ref = interior(o.uh)
    changed = np.abs(ref - interior(uh)).reshape(N, N, -1).max(-1) > 0
    err = np.abs(out - ref).reshape(N, N, -1).max(-1) / np.abs(ref).reshape(N, N, -1).max(-1)
    print(f"[pp] {changed.sum()} of {N * N} cells limited by the oracle; worst cell difference {err.max():.2e}")
    assert 0.1 * N * N < changed.sum() < 0.9 * N * N
    assert err.max() < (1e-10 if form == "fp64" else 1e-6)


def test_div_free(pkg):
    g = golden("divfree_16")
    s, uh = solver_from(pkg, g)
    s.upload(uh, g["Bx"], g["By"])
    s.div_free_Balsara()
    out, _, _ = s.download()
    assert rel_l2(out, unpack(g["out"])) < 1e-14


def test_dt_and_diagnostics(pkg):
    g = golden("dt_diag_16")
    s, uh = solver_from(pkg, g)
    s.upload(uh, g["Bx"], g["By"])
    dt = s.calculate_dt()
    assert abs(dt - float(g["dt"])) < 1e-14 * float(g["dt"])
    d = s.diagnostics()
    assert d["umax"] == float(g["umax"])
    assert abs(d["pmin"] - float(g["pmin"])) < 1e-13 * abs(float(g["pmin"]))
    assert abs(d["totaldiv"] - float(g["totaldiv"])) < 1e-12 * float(g["totaldiv"])


RUNS = [("run_f4_16", 4, 0, 1), ("run_f2_16", 2, 0, 1), ("run_f1_16", 1, 0, 1), ("run_f3_16", 3, 0, 1),
        ("run_f4_pp_16_outflow", 4, 1, 1), ("run_f4_pp_32", 4, 1, 1), ("run_f2_32", 2, 0, 1),
        ("run_rotor_f4_pp_32", 4, 1, 1),
        # TVB gated by KXRCF_Detector (Ck = 1)
        ("run_f2_kx_32", 2, 0, 2), ("run_rotor_f4_pp_kx_32", 4, 1, 2),
        # shock problems from P0 data (init6, init8b), outflow boundaries
        ("run_cloud_f4_pp_16", 4, 1, 1), ("run_briowu_y_f4_pp_16", 4, 1, 1)]


# The rotor carries B = (B0, 0) in most of the domain: B is parallel to the x
# normal, where the reference's eigenvector normalisation alpha_f =
# sqrt|a^2 - c_s^2| / sqrt|c_f^2 - c_s^2| (eigenmatrix.f90:58) is the square
# root of a rounding residue.  Over several steps the GPU's FMA-contracted
# residual differs from the reference by ulps, which that square root turns
# into ~1e-8 relative changes of the limiter's eigenvectors.  Its whole-run bar
# is therefore 1e-9; on identical inputs every RK stage of the same run meets
# 1e-12 (test_per_stage_identical_inputs).
RUN_TOL = {"run_rotor_f4_pp_32": 1e-9, "run_rotor_f4_pp_kx_32": 1e-9,
           # 87 steps; GPU vs oracle grows by ~1e-17 per step (measured), past 1e-12 at the end
           "run_briowu_y_f4_pp_16": 5e-12}
# The MHD blast (init5: B = 100/sqrt(4 pi), p = 0.1 outside the disk, plasma
# beta ~ 1e-3) is not a whole-run case.  Its x-symmetric faces have a star
# velocity S* that is zero up to rounding, and the reference's HLLC/HLLD flux
# is not continuous across S* = 0 (its total pressure is gamma1 (E - rho
# |u|^2 / 2), HLLD_Flux.f90:61-68): the star branch, hence an O(1) part of the
# face flux, is picked by the sign of a rounding residue.  From the second
# stage on (once the first limiter pass has broken exact mirror symmetry) the
# GPU and the reference pick differently on a few faces; after a few steps
# the reference's own state has negative pressure.  Its parity is checked on
# the prologue and the first stage (test_per_stage_identical_inputs).


@pytest.mark.parametrize("name", ["kxrcf_16", "kxrcf_rotor_32"])
def test_kxrcf(pkg, name):
    """KXRCF_Detector flags, cell by cell, against the reference (P6 build)"""
    g = golden(name)
    s, uh = solver_from(pkg, g, Ck=float(g["Ck"]))
    s.upload(uh, g["Bx"], g["By"])
    n = s.KXRCF_Detector()
    flags = s.debug_get(pkg.TROUBLE)
    assert n == int(g["trouble"].sum())
    assert np.array_equal(flags.astype(np.int32), g["trouble"])


@pytest.mark.parametrize("name,ft,pp,ind", RUNS)
def test_rk3_fused(pkg, name, ft, pp, ind):
    """mhd2d_step (fused kernels, device-side dt) against the reference run"""
    g = golden(name)
    s, uh0 = solver_from(pkg, g, uh_key="uh0", ft=ft, pp=pp, indicator=ind, Ck=float(g.get("Ck", 1.0)))
    s.upload(uh0, g["Bx0"], g["By0"])
    t, n = s.RK3(float(g["tend"]))
    assert n == int(g["count"])
    assert abs(t - float(g["t"])) < 1e-15
    uh, Bx, By = s.download()
    tol = RUN_TOL.get(name, TOL)
    assert rel_l2(interior(uh), interior(unpack(g["uh"]))) < tol
    assert rel_l2(Bx, g["Bx"]) < tol and rel_l2(By, g["By"]) < tol
    d = s.diagnostics()
    assert d["totaldiv"] < 1e-12
    if "trouble" in g:      # the last KXRCF call of the run flags the same cells
        assert np.array_equal(s.debug_get(pkg.TROUBLE).astype(np.int32), g["trouble"])


@pytest.mark.parametrize("name,ft,pp", [("run_f4_pp_32", 4, 1), ("run_f4_pp_16_outflow", 4, 1)])
def test_rk3_per_subroutine(pkg, oracle_mod, name, ft, pp):
    """the RK3.f90 loop driven through the per-subroutine entry points,
    checked against the oracle after every stage"""
    O = oracle_mod
    g = golden(name)
    s, uh0 = solver_from(pkg, g, uh_key="uh0", ft=ft, pp=pp)
    o = O.Oracle(uh0.shape[0] - 2, uh0.shape[1] - 2, flux_type=ft, bc=int(g["bc"]), domain=tuple(g["domain"]),
                 M=float(g["M"]), pp=pp)
    s.upload(uh0, g["Bx0"], g["By0"])
    o.load(uh0, g["Bx0"], g["By0"])
    for x in (s, o):
        x.set_bc(); x.TVB_Limiter_P2(); x.div_free_Balsara()

    def check(tag):
        uh, Bx, By = s.download()
        assert rel_l2(interior(uh), interior(o.uh)) < TOL, tag
        assert rel_l2(Bx, o.Bx) < TOL and rel_l2(By, o.By) < TOL, tag

    check("prologue")
    for step in range(3):
        o.calculate_dt()
        dt = o.scalars()["dt"]
        assert abs(s.calculate_dt() - dt) < 1e-14 * dt
        for stage in (1, 2, 3):
            if stage != 2:
                s.set_bc(); o.set_bc()
            s.Lh(); s.LhEz(); s.rk_stage(stage, dt)
            o.Lh(); o.LhEz()
            # the oracle's combination for this stage (RK3.f90:80-135)
            _oracle_stage(o, stage, dt)
            check(f"step {step} stage {stage} combine")
            s.TVB_Limiter_P2(); s.div_free_Balsara()
            o.TVB_Limiter_P2(); o.div_free_Balsara()
            if stage == 1:
                s.set_bc(); o.set_bc()
            check(f"step {step} stage {stage} limited")


@pytest.mark.parametrize("name,steps", [("run_rotor_f4_pp_32", 3), ("run_blast_f4_pp_16", 1),
                                        ("run_briowu_y_f4_pp_16", 30)])
def test_per_stage_identical_inputs(pkg, oracle_mod, name, steps):
    """every RK stage of a run on identical inputs: the GPU gets the oracle's
    state before each reference subroutine and must agree to 1e-12.  For the
    blast only the prologue and the first stage are compared (see the note
    above RUN_TOL)."""
    O = oracle_mod
    g = golden(name)
    s, uh0 = solver_from(pkg, g, uh_key="uh0", ft=4, pp=1)
    N = uh0.shape[0] - 2
    o = O.Oracle(N, N, flux_type=4, bc=int(g["bc"]), domain=tuple(g["domain"]), M=float(g["M"]), pp=1)
    o.load(uh0, g["Bx0"], g["By0"])

    def same_input():
        s.upload(o.uh, o.Bx, o.By)

    def cmp(tag):
        uh, Bx, By = s.download()
        assert rel_l2(interior(uh), interior(o.uh)) < TOL, tag

    o.set_bc()
    same_input(); s.TVB_Limiter_P2(); o.TVB_Limiter_P2(); cmp("prologue TVB")
    same_input(); s.div_free_Balsara(); o.div_free_Balsara(); cmp("prologue div_free")
    compared = 0
    for step in range(steps):
        o.calculate_dt()
        dt = o.scalars()["dt"]
        for stage in (1, 2, 3):
            if stage != 2:
                o.set_bc()
            if name.startswith("run_blast") and compared == 1:
                return
            compared += 1
            same_input(); s.Lh(); s.LhEz()
            o.Lh(); o.LhEz()
            assert rel_l2(s.debug_get(pkg.DU), o.du) < TOL, (step, stage, "Lh")
            assert rel_l2(s.debug_get(pkg.DEZ1), o.dEz1) < TOL, (step, stage, "LhEz")
            assert rel_l2(s.debug_get(pkg.DEZ2), o.dEz2) < TOL, (step, stage, "LhEz")
            _oracle_stage(o, stage, dt)
            same_input(); s.TVB_Limiter_P2(); o.TVB_Limiter_P2(); cmp((step, stage, "TVB+pp"))
            same_input(); s.div_free_Balsara(); o.div_free_Balsara(); cmp((step, stage, "div_free"))
            if stage == 1:
                o.set_bc()


def _oracle_stage(o, stage, dt):
    st = o.__dict__.setdefault("_rk", {})
    uh, Bx, By, du = o.uh, o.Bx, o.By, o.du
    if stage == 1:
        st["uI"] = uh + dt * du
        st["BxI"] = Bx + dt * o.dEz1
        st["ByI"] = By + dt * o.dEz2
        st["uh0"], st["Bx0"], st["By0"] = uh.copy(), Bx.copy(), By.copy()
        uh[...] = st["uI"]; Bx[...] = st["BxI"]; By[...] = st["ByI"]
    elif stage == 2:
        uh[...] = st["uI"] + dt * du
        Bx[...] = st["BxI"] + dt * o.dEz1
        By[...] = st["ByI"] + dt * o.dEz2
        st["uII"] = 0.75 * st["uh0"] + 0.25 * uh
        st["BxII"] = 0.75 * st["Bx0"] + 0.25 * Bx
        st["ByII"] = 0.75 * st["By0"] + 0.25 * By
        uh[...] = st["uII"]; Bx[...] = st["BxII"]; By[...] = st["ByII"]
    else:
        Bx[...] = st["BxII"] + dt * o.dEz1
        By[...] = st["ByII"] + dt * o.dEz2
        uh[...] = st["uII"] + dt * du
        uh[...] = (1.0 / 3.0) * st["uh0"] + (2.0 / 3.0) * uh
        Bx[...] = (1.0 / 3.0) * st["Bx0"] + (2.0 / 3.0) * Bx
        By[...] = (1.0 / 3.0) * st["By0"] + (2.0 / 3.0) * By


@pytest.mark.parametrize("N,ft,pp,steps", [(128, 4, 1, 3), (96, 2, 0, 3)])
def test_scale_vs_oracle(pkg, oracle_mod, N, ft, pp, steps):
    """perturbed Orszag-Tang at larger N: GPU vs oracle, same seeded input"""
    O = oracle_mod
    dom = (0.0, 2 * np.pi, 0.0, 2 * np.pi)
    o = O.Oracle(N, N, flux_type=ft, domain=dom, M=50.0, CFL=0.2, pp=pp)
    o.init_data(O.IC_OT_PERTURBED, [1e-2, 20261203])
    s = pkg.Solver(N, N, flux_type=ft, domain=dom, M=50.0, CFL=0.2, pp=pp)
    s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 20261203])
    uh, Bx, By = s.download()
    assert rel_l2(interior(uh), interior(o.uh)) < 1e-13
    s.upload(o.uh, o.Bx, o.By)      # identical inputs from here on
    o.rk3_prologue()
    s.rk3_prologue()
    for _ in range(steps):
        o.rk3_step(1e9)
        s.step(1e9)
    uh, Bx, By = s.download()
    assert rel_l2(interior(uh), interior(o.uh)) < TOL
    assert rel_l2(Bx, o.Bx) < TOL and rel_l2(By, o.By) < TOL
    t, dt, n = s.get_time()
    sc = o.scalars()
    assert n == steps and abs(t - sc["t"]) < 1e-14


def test_invariants_2048(pkg):
    """the metric's shape: div B stays at roundoff, the cell-average mass is
    conserved (periodic), and the state stays finite"""
    N = 2048
    dom = (0.0, 2 * np.pi, 0.0, 2 * np.pi)
    s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1)
    s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 20261203])
    s.rk3_prologue()
    uh, _, _ = s.download()
    m0 = interior(uh)[:, :, 0, 0].sum()
    for _ in range(3):
        s.step()
    d = s.diagnostics()
    uh, Bx, By = s.download()
    assert np.isfinite(uh).all()
    assert d["totaldiv"] < 1e-12
    assert abs(interior(uh)[:, :, 0, 0].sum() - m0) < 1e-12 * abs(m0)
    assert d["pmin"] > 0


@pytest.mark.parametrize("N,steps,ft", [(256, 20, 4), (128, 20, 1), (128, 20, 2), (128, 20, 3)],
                         ids=["HLLD", "LF", "HLL", "HLLC"])
def test_mixed_vs_fp64(pkg, N, steps, ft):
    """mixed form (FP32 face fluxes, quadratures and higher-mode storage;
    FP64 averages, magnetic modes, CT, div-free, limiters) against the plain FP64 form on the same input:
    relative L2 of every field <= 1e-6 (measured: 1.9-2.4e-7 for the four
    flux types), div B at roundoff.  Every flux type: the mixed face kernels
    and the vertex EMF are instantiated per flux type (the plain form's flux
    types are pinned to the oracle by test_rk3_fused and test_scale_vs_oracle)."""
    dom = (0.0, 2 * np.pi, 0.0, 2 * np.pi)
    out = {}
    for prec in (pkg.FP64, pkg.MIXED):
        s = pkg.Solver(N, N, flux_type=ft, domain=dom, M=50.0, CFL=0.2, pp=1, precision=prec)
        s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 20261203])
        s.rk3_prologue()
        for _ in range(steps):
            s.step()
        out[prec] = (s.download(), s.diagnostics(), s.get_time())
    (u64, _, _), d64, t64 = out[pkg.FP64]
    (u32, _, _), d32, t32 = out[pkg.MIXED]
    assert abs(t64[0] - t32[0]) < 1e-6 * t64[0]
    errs = [rel_l2(interior(u32)[:, :, :, n], interior(u64)[:, :, :, n]) for n in range(8)
            if np.linalg.norm(interior(u64)[:, :, :, n]) > 0]
    print(f"[mixed] flux type {ft}, {N}^2, {steps} steps: largest field rel L2 mixed vs FP64 {max(errs):.2e}, "
          f"totaldiv {d32['totaldiv']:.1e}")
    assert max(errs) < 1e-6
    assert d32["totaldiv"] < 1e-12


def test_mixed_storage_roundtrip(pkg):
    """mixed-form state storage: cell averages and every magnetic-field mode
    stay FP64 (exact round trip), the higher modes of rho, rho u, rho v,
    rho w, E and Bz are stored in FP32 (round to nearest)"""
    N = 24
    rng = np.random.default_rng(5)
    s64 = pkg.Solver(N, N, flux_type=pkg.HLLD)
    uh = rng.standard_normal((N + 2, N + 2, 10, 8))
    _, sx, sy = s64.shapes()
    Bx, By = rng.standard_normal(sx), rng.standard_normal(sy)
    s64.upload(uh, Bx, By)
    ref = s64.download()[0]                 # inactive modes read back as zero
    assert np.count_nonzero(np.any(ref != 0, axis=(0, 1))) == 52
    sm = pkg.Solver(N, N, flux_type=pkg.HLLD, precision=pkg.MIXED)
    sm.upload(ref, Bx, By)
    got, gx, gy = sm.download()
    assert np.array_equal(gx, Bx) and np.array_equal(gy, By)
    for n in range(8):
        for b in range(10):
            a, r = got[:, :, b, n], ref[:, :, b, n]
            if b == 0 or n in (5, 6):
                assert np.array_equal(a, r), (b, n)
            else:
                assert np.array_equal(a, r.astype(np.float32).astype(np.float64)), (b, n)


def test_mixed_shock_rotor(pkg):
    """the mixed form on a shock flow (the reference's rotor, init4.txt, with
    positivity and KXRCF-gated TVB): stays finite and divergence-free and
    close to the plain form (measured 6.1e-6 in density after 20 steps at
    128^2; the bar is about three times that).  FP32 rounding can move a
    limiter decision at a discontinuity, so the bar is looser than the smooth
    benchmark flow's 1e-6 (test_mixed_vs_fp64).  The mixed form's list
    kernels limit the cells TVB changes in place and keep their unlimited
    value apart for the next stage's RK combination; combining the limited
    value instead would move every such cell by a share of the limiter's
    change, far past the bar."""
    N = 128
    out = {}
    for prec in (pkg.FP64, pkg.MIXED):
        s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=(0.0, 1.0, 0.0, 1.0), M=10.0, CFL=0.15, pp=1,
                       indicator=pkg.IND_KXRCF, precision=prec)
        s.init_data(pkg.IC_ROTOR)
        s.rk3_prologue()
        for _ in range(20):
            s.step()
        out[prec] = (s.download()[0], s.diagnostics())
    (u64, d64), (u32, d32) = out[pkg.FP64], out[pkg.MIXED]
    err = rel_l2(interior(u32)[:, :, :, 0], interior(u64)[:, :, :, 0])
    print(f"[mixed] rotor {N}^2, 20 steps: density rel L2 mixed vs FP64 {err:.3e}, "
          f"totaldiv {d32['totaldiv']:.2e}, pmin {d32['pmin']:.3e} / {d64['pmin']:.3e}")
    assert np.isfinite(u32).all()
    assert d32["totaldiv"] < 1e-12
    assert err < 2e-5


@pytest.mark.parametrize("ind", ["kxrcf", "none"])
def test_odd_grid_mixed_boundaries(pkg, oracle_mod, ind):
    """37 x 23 cells (rows not a multiple of the 16-lane padding, non-square
    cells), outflow in x and periodic in y: the plain form against the oracle
    after 10 steps, and the mixed form (its non-square cell-kernel variant and
    the double-double CT with hx != hy) within 1e-5 of the plain form; with
    KXRCF-gated TVB and with no TVB (the plain form's fused limiter pass with
    and without the list of TVB-changed cells)"""
    O = oracle_mod
    Nx, Ny, dom, bc = 37, 23, (0.0, 2 * np.pi, 0.0, 2 * np.pi), (2, 2, 1, 1)
    oi, pi = (O.IND_KXRCF, pkg.IND_KXRCF) if ind == "kxrcf" else (O.IND_NONE, pkg.IND_NONE)
    o = O.Oracle(Nx, Ny, flux_type=O.HLLD, bc=bc, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=oi, Ck=1.0)
    o.init_data(O.IC_OT_PERTURBED, [1e-2, 5])
    runs = {}
    for name, prec in (("fp64", pkg.FP64), ("mixed", pkg.MIXED)):
        s = pkg.Solver(Nx, Ny, flux_type=pkg.HLLD, bc=bc, domain=dom, M=50.0, CFL=0.2, pp=1,
                       indicator=pi, precision=prec, Ck=1.0)
        s.upload(o.uh, o.Bx, o.By)
        s.rk3_prologue()
        for _ in range(10):
            s.step()
        runs[name] = (s.download(), s.diagnostics(), s.get_time())
    o.rk3_prologue()
    for _ in range(10):
        o.rk3_step(1e9)
    (uh, Bx, By), d, (t, _, n) = runs["fp64"]
    err = rel_l2(interior(uh), interior(o.uh))
    (um, mx, my), dm, tm = runs["mixed"]
    errm = rel_l2(interior(um)[:, :, :, 0], interior(uh)[:, :, :, 0])
    print(f"[odd grid] fp64 vs oracle {err:.2e}; mixed vs fp64 density {errm:.2e}; div B {d['totaldiv']:.1e} / "
          f"{dm['totaldiv']:.1e}")
    assert n == 10 and abs(t - o.scalars()["t"]) < 1e-14
    assert err < TOL and rel_l2(Bx, o.Bx) < TOL and rel_l2(By, o.By) < TOL
    assert errm < 1e-5 and rel_l2(mx, Bx) < 1e-5 and rel_l2(my, By) < 1e-5
    assert dm["totaldiv"] < 1e-12 and d["totaldiv"] < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_face_jump(pkg, prec):
    """the face-jump half of the divergence report: after the limiter pass
    every in-cell normal-B trace equals its face polynomial up to rounding,
    and a face changed from outside shows up as exactly that change"""
    P = pkg.FP64 if prec == "fp64" else pkg.MIXED
    s = pkg.Solver(48, 48, flux_type=pkg.HLLD, M=50.0, CFL=0.2, pp=1, indicator=pkg.IND_KXRCF, precision=P, Ck=1.0)
    s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 7])
    s.rk3_prologue()
    for _ in range(5):
        s.step()
    j = s.face_jump()
    print(f"[face jump] {prec}: {j:.2e}")
    assert j < 1e-13
    uh, Bx, By = s.download()
    Bx[10, 10, 0] += 1e-6            # face i = 10 of row 11: interior cells on both sides
    s.upload(uh, Bx, By)
    assert abs(s.face_jump() - 1e-6) < 1e-12


def test_grid_minimum(pkg):
    """fewer than 2 cells in a direction is rejected at creation"""
    for Nx, Ny in ((1, 1), (5, 1), (1, 4)):
        with pytest.raises(pkg.MHD2DError):
            pkg.Solver(Nx, Ny)


@pytest.mark.parametrize("Nx,Ny", [(2, 2), (2, 3), (5, 2), (65, 2)])
def test_tiny_grids(pkg, oracle_mod, Nx, Ny):
    """grids smaller than a wave or a tile (2 x 2 cells, two rows, 65 cells
    per row): the plain form against the oracle after 3 steps, and the mixed
    form with the learned screen stays finite (periodic)"""
    O = oracle_mod
    dom = (0.0, 2 * np.pi, 0.0, 2 * np.pi)
    o = O.Oracle(Nx, Ny, flux_type=O.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=O.IND_KXRCF, Ck=1.0)
    o.init_data(O.IC_OT_PERTURBED, [1e-2, 9])
    s = pkg.Solver(Nx, Ny, flux_type=pkg.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=pkg.IND_KXRCF, Ck=1.0)
    s.upload(o.uh, o.Bx, o.By)
    o.rk3_prologue()
    s.rk3_prologue()
    for _ in range(3):
        o.rk3_step(1e9)
        s.step()
    uh, Bx, By = s.download()
    assert rel_l2(interior(uh), interior(o.uh)) < TOL
    assert rel_l2(Bx, o.Bx) < TOL and rel_l2(By, o.By) < TOL
    m = pkg.Solver(Nx, Ny, flux_type=pkg.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=pkg.IND_MLP,
                   precision=pkg.MIXED, Ck=1.0)
    m.mlp_set_weights(pkg.load_indicator_weights())
    m.upload(o.uh, o.Bx, o.By)
    m.rk3_prologue()
    m.step_graph(3)
    d = m.diagnostics()
    assert d["flags"] == 0 and np.isfinite(m.download()[0]).all()
