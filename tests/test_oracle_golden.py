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


IC_LIBRARY = [("blast", 5), ("cloud", 6), ("cloud2", 61), ("loop", 7), ("briowu", 8), ("briowu_y", 81),
              ("ryujones", 9), ("ryujones_y", 91)]


@pytest.mark.parametrize("tag,ic", IC_LIBRARY)
def test_init_library(oracle_mod, tag, ic):
    """init_data.f90 with init5-9b.txt (incl. the P0 path of shock = 1)"""
    O = oracle_mod
    g = golden(f"init_{tag}_16")
    o = O.Oracle(16, 16, domain=tuple(g["domain"]), bc=int(g["bc"]))
    o.init_data(ic)
    assert np.array_equal(interior(o.uh), interior(unpack(g["uh"])))
    assert np.array_equal(o.Bx, g["Bx"]) and np.array_equal(o.By, g["By"])


def test_init_library_shock_is_p0():
    """shock = 1 ICs hold cell-centre values and no higher modes; the
    smooth-projection ICs do not"""
    for tag, ic in IC_LIBRARY:
        uh = unpack(golden(f"init_{tag}_16")["uh"])
        hi = np.abs(interior(uh)[:, :, 1:, :]).max()
        if tag in ("blast", "cloud", "cloud2", "briowu_y"):
            assert hi == 0.0, tag
        else:
            assert hi > 0.0, tag


def test_calculate_az(oracle_mod):
    O = oracle_mod
    g = golden("az_loop_16")
    o = O.Oracle(16, 16, domain=tuple(g["domain"]))
    o.load(unpack(g["uh"]), g["Bx0"], g["By0"])
    o.calculate_Az()
    assert np.array_equal(o.Bx, g["Bx"]) and np.array_equal(o.By, g["By"])
    # closed form: face averages of curl(Az) have zero discrete divergence
    hx, hy = (g["domain"][1] - g["domain"][0]) / 16, (g["domain"][3] - g["domain"][2]) / 16
    div = (o.Bx[1:, :, 0] - o.Bx[:-1, :, 0]) / hx + (o.By[:, 1:, 0] - o.By[:, :-1, 0]) / hy
    assert np.abs(div).max() < 1e-15


def test_calculate_l2error(oracle_mod):
    O = oracle_mod
    g = golden("l2err_vortex_16")
    o = O.Oracle(16, 16, domain=tuple(g["domain"]))
    o.load(unpack(g["uh"]), g["Bx"], g["By"])
    for t20, ref in zip(g["t20"], g["L2"]):
        assert np.array_equal(o.calculate_L2error(float(t20)), ref)
    # the exact projection of a constant has zero error; w and Bz are zero
    assert g["L2"][0][3] == 0.0 and g["L2"][0][7] == 0.0


@pytest.mark.parametrize("name,ft,pp,ind", [("run_f4_16", 4, 0, 1), ("run_f2_16", 2, 0, 1), ("run_f1_16", 1, 0, 1),
                                            ("run_f3_16", 3, 0, 1), ("run_f4_pp_16_outflow", 4, 1, 1),
                                            ("run_blast_f4_pp_16", 4, 1, 1), ("run_cloud_f4_pp_16", 4, 1, 1),
                                            ("run_briowu_y_f4_pp_16", 4, 1, 1),
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
