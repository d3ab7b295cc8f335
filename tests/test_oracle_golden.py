"""The CPU restatement (oracle/mhd2d_oracle.c) against the golden fixtures
written by the reference Fortran (tests/golden/make_golden.py).

These pin the oracle: every reference subroutine on the hot path is run on
seeded inputs by the reference itself and compared here.  The restatement
follows the reference's operation order and the reference is built with
DOUBLE PRECISION kept at 8 bytes (oracle/build_ref.sh), so the two agree
bit for bit: the tolerances are zero.
"""
import numpy as np
import pytest

from conftest import golden, interior, rel_l2, unpack

TOL_FN = 0.0      # bit-identical to the reference (see module docstring)
TOL_RUN = 0.0


@pytest.mark.parametrize("tag,ic", [("ot", 3), ("vortex", 1), ("rotor", 4)])
def test_init_data(oracle_mod, tag, ic):
    O = oracle_mod
    g = golden(f"init_{tag}_16")
    o = O.Oracle(16, 16, domain=tuple(g["domain"]))
    o.init_data(ic)
    ref = unpack(g["uh"])
    assert rel_l2(interior(o.uh), interior(ref)) <= TOL_FN
    assert rel_l2(o.Bx, g["Bx"]) <= TOL_FN
    assert rel_l2(o.By, g["By"]) <= TOL_FN


@pytest.mark.parametrize("name,ft,bc", [("lh_f1_16", 1, 1), ("lh_f2_16", 2, 1), ("lh_f3_16", 3, 1),
                                        ("lh_f4_16", 4, 1), ("lh_f4_12x8_outflow", 4, 2)])
def test_lh_lhez(oracle_mod, name, ft, bc):
    O = oracle_mod
    g = golden(name)
    uh = unpack(g["uh"])
    o = O.Oracle(uh.shape[0] - 2, uh.shape[1] - 2, flux_type=ft, bc=bc, domain=tuple(g["domain"]))
    o.load(uh, g["Bx"], g["By"])
    o.set_bc()
    if "uh_bc" in g:   # set_bc with outflow boundaries, corners included
        assert np.array_equal(o.uh, unpack(g["uh_bc"]))
    o.Lh()
    o.LhEz()
    assert rel_l2(o.du, unpack(g["du"])) <= TOL_FN
    assert rel_l2(o.dEz1, g["dEz1"]) <= TOL_FN
    assert rel_l2(o.dEz2, g["dEz2"]) <= TOL_FN
    assert rel_l2(o.EzVertex, g["EzVertex"]) <= TOL_FN


@pytest.mark.parametrize("M", [0, 10])
def test_tvb_stale_halos(oracle_mod, M):
    O = oracle_mod
    g = golden(f"tvb_M{M}_16")
    o = O.Oracle(16, 16, domain=tuple(g["domain"]), M=float(g["M"]))
    o.load(unpack(g["uh"]), g["Bx"], g["By"])
    o.TVB_Limiter_P2()
    ref = unpack(g["out"])
    assert int(g["cells_changed"]) > 0
    assert rel_l2(o.uh, ref) <= TOL_FN
    # mode slots the limiter zeroes stay exactly zero
    assert np.array_equal(o.uh == 0, ref == 0)


def test_pp_limiter(oracle_mod):
    O = oracle_mod
    g = golden("pp_16")
    o = O.Oracle(16, 16, domain=tuple(g["domain"]))
    o.load(unpack(g["uh"]), g["Bx"], g["By"])
    o.pp_Limiter()
    assert int(g["cells_changed"]) > 20
    assert rel_l2(o.uh, unpack(g["out"])) <= TOL_FN


def test_div_free(oracle_mod):
    O = oracle_mod
    g = golden("divfree_16")
    o = O.Oracle(16, 16, domain=tuple(g["domain"]))
    o.load(unpack(g["uh"]), g["Bx"], g["By"])
    o.div_free_Balsara()
    assert rel_l2(o.uh, unpack(g["out"])) <= TOL_FN


def test_dt_and_diagnostics(oracle_mod):
    O = oracle_mod
    g = golden("dt_diag_16")
    o = O.Oracle(16, 16, domain=tuple(g["domain"]))
    o.load(unpack(g["uh"]), g["Bx"], g["By"])
    o.calculate_dt()
    o.calculate_umax()
    o.calculate_totaldiv()
    o.calculate_pmin()
    s = o.scalars()
    for k in ("dt", "alphax", "alphay", "umax", "totaldiv", "pmin"):
        assert s[k] == float(g[k]), k


@pytest.mark.parametrize("name", ["kxrcf_16", "kxrcf_rotor_32"])
def test_kxrcf(oracle_mod, name):
    O = oracle_mod
    g = golden(name)
    uh = unpack(g["uh"])
    o = O.Oracle(uh.shape[0] - 2, uh.shape[1] - 2, domain=tuple(g["domain"]), Ck=float(g["Ck"]))
    o.load(uh, g["Bx"], g["By"])
    o.KXRCF_Detector()
    assert 0 < int(g["trouble"].sum()) < g["trouble"].size
    assert np.array_equal(o.trouble, g["trouble"])


def test_kxrcf_closed_form(oracle_mod):
    """Piecewise-constant states: a uniform flow has no inflow jump (no
    flags); a density step advected in +x flags exactly the column of cells
    whose left (inflow) face carries the jump, KXRCF_Detector.f90:62-68."""
    O = oracle_mod
    N = 8
    o = O.Oracle(N, N, domain=(0.0, 1.0, 0.0, 1.0), Ck=1.0)
    uh = np.zeros((N + 2, N + 2, 10, 8), order="F")
    uh[:, :, 0, 0] = 1.0
    uh[:, :, 0, 1] = 0.5          # rho u > 0: inflow through every left face
    uh[:, :, 0, 4] = 2.0
    Bx, By = np.zeros((N + 1, N, 3), order="F"), np.zeros((N, N + 1, 3), order="F")
    o.load(uh, Bx, By)
    o.KXRCF_Detector()
    assert o.trouble.sum() == 0
    uh[5:, :, 0, 0] = 2.0         # cells i >= 5 (1-based) are denser: jump on the left face of i = 5
    uh[5:, :, 0, 1] = 1.0         # same velocity
    o.load(uh, Bx, By)
    o.KXRCF_Detector()
    expect = np.zeros((N, N), dtype=np.int32)
    expect[4, :] = 1
    assert np.array_equal(o.trouble, expect)


@pytest.mark.parametrize("name,ft,pp,ind", [("run_f4_16", 4, 0, 1), ("run_f2_16", 2, 0, 1), ("run_f1_16", 1, 0, 1),
                                            ("run_f3_16", 3, 0, 1), ("run_f4_pp_16_outflow", 4, 1, 1),
                                            ("run_f4_pp_32", 4, 1, 1), ("run_f2_32", 2, 0, 1),
                                            ("run_rotor_f4_pp_32", 4, 1, 1),
                                            ("run_f2_kx_32", 2, 0, 2), ("run_rotor_f4_pp_kx_32", 4, 1, 2)])
def test_rk3_run(oracle_mod, name, ft, pp, ind):
    O = oracle_mod
    g = golden(name)
    uh0 = unpack(g["uh0"])
    o = O.Oracle(uh0.shape[0] - 2, uh0.shape[1] - 2, flux_type=ft, bc=int(g["bc"]),
                 domain=tuple(g["domain"]), M=float(g["M"]), pp=pp, indicator=ind)
    o.load(uh0, g["Bx0"], g["By0"])
    o.rk3_prologue()
    while o.rk3_step(float(g["tend"])):
        pass
    s = o.scalars()
    assert s["count"] == float(g["count"])
    assert s["t"] == float(g["t"])
    assert rel_l2(interior(o.uh), interior(unpack(g["uh"]))) <= TOL_RUN
    assert rel_l2(o.Bx, g["Bx"]) <= TOL_RUN
    assert rel_l2(o.By, g["By"]) <= TOL_RUN
    assert s["pmin"] == float(g["pmin"])
    assert s["totaldiv"] < 1e-12 and float(g["totaldiv"]) < 1e-12
