"""Why three whole runs of the GPU plain form sit above the 1e-12 bar.

The GPU compiler contracts products into sums (fused multiply-adds); the
reference under -fdefault-real-8 on x86-64 does not.  The CPU restatement is
built twice from one source: exactly as the reference rounds (tolerance 0
against every reference fixture, test_oracle_golden.py) and with every
contraction gcc finds (-mfma -ffp-contract=fast, oracle/Makefile).  Run on the
reference's whole-run fixtures, the contracted build:
- stays within 1e-13 of the reference on nine of the thirteen runs;
- drifts on exactly the cases the GPU parity tests document: the rotor
  (1e-11 here; B along x, eigenmatrix.f90:58 takes sqrt|a^2 - c_s^2| of a
  rounding residue), y-Brio-Wu (1e-12 over 87 steps) and the MHD blast
  (O(1e-2): its symmetric faces pick the HLLD star branch by the sign of a
  rounding residue, HLLD_Flux.f90:61-68);
- and parts from it, on identical inputs, by the same mechanisms: slow
  accumulation (y-Brio-Wu), a rounding-discontinuous first TVB call (rotor)
  and a star-branch flip in the first step (blast).
So the whole-run exceptions are the reference algorithm's own sensitivity to
rounding, reproduced on the CPU by contraction alone, not a defect of the
GPU path.  Needs an FMA-capable host (x86-64 with FMA3).
"""
import numpy as np
import pytest

from conftest import golden, interior, rel_l2, unpack

RUNS = [("run_f4_16", 4, 0, 1), ("run_f2_16", 2, 0, 1), ("run_f1_16", 1, 0, 1), ("run_f3_16", 3, 0, 1),
        ("run_f4_pp_16_outflow", 4, 1, 1), ("run_cloud_f4_pp_16", 4, 1, 1), ("run_f4_pp_32", 4, 1, 1),
        ("run_f2_32", 2, 0, 1), ("run_f2_kx_32", 2, 0, 2), ("run_blast_f4_pp_16", 4, 1, 1),
        ("run_briowu_y_f4_pp_16", 4, 1, 1), ("run_rotor_f4_pp_32", 4, 1, 1), ("run_rotor_f4_pp_kx_32", 4, 1, 2)]
SENSITIVE = {"run_blast_f4_pp_16", "run_briowu_y_f4_pp_16", "run_rotor_f4_pp_32", "run_rotor_f4_pp_kx_32"}


def fma_host():
    try:
        return " fma " in " " + open("/proc/cpuinfo").read().replace("\n", " ") + " "
    except OSError:
        return False


pytestmark = pytest.mark.skipif(not fma_host(), reason="the contracted build needs FMA3")


def oracle(O, g, ft, pp, ind, variant):
    uh0 = unpack(g["uh0"])
    o = O.Oracle(uh0.shape[0] - 2, uh0.shape[1] - 2, flux_type=ft, bc=int(g["bc"]), domain=tuple(g["domain"]),
                 M=float(g["M"]), pp=pp, indicator=ind, variant=variant)
    o.load(uh0, g["Bx0"], g["By0"])
    return o


@pytest.mark.parametrize("name,ft,pp,ind", RUNS)
def test_contracted_whole_runs(oracle_mod, name, ft, pp, ind):
    O = oracle_mod
    g = golden(name)
    o = oracle(O, g, ft, pp, ind, "fma")
    o.rk3_prologue()
    while o.rk3_step(float(g["tend"])):
        pass
    err = rel_l2(interior(o.uh), interior(unpack(g["uh"])))
    print(f"[fma] {name}: contracted build vs reference after {int(o.scalars()['count'])} steps: {err:.2e}")
    if name not in SENSITIVE:
        assert err <= 1e-13
    elif name.startswith("run_blast"):
        assert err > 1e-6          # star-branch flips: an O(1) part of some face fluxes
    elif name.startswith("run_rotor"):
        assert 1e-12 < err <= 1e-9  # the GPU parity test's documented bar for the rotor
    else:
        assert 1e-13 < err <= 5e-12  # y-Brio-Wu, the GPU bar


@pytest.mark.parametrize("name,ft,pp,ind", [r for r in RUNS if r[0] in SENSITIVE])
def test_contracted_mechanisms(oracle_mod, name, ft, pp, ind):
    """where the two builds part, from identical inputs: y-Brio-Wu agrees to
    1e-15 per step (its drift is slow accumulation over 87 steps); the
    rotor's very first TVB call already differs (eigenmatrix.f90:54-60: its
    eigenvectors depend on cf^2 - cs^2, a rounding residue where B is along
    x); the blast's first step differs at O(1e-4) (the HLLD star branch of
    its symmetric faces, HLLD_Flux.f90:61-68)"""
    O = oracle_mod
    g = golden(name)
    a, b = oracle(O, g, ft, pp, ind, ""), oracle(O, g, ft, pp, ind, "fma")
    for o in (a, b):
        o.rk3_prologue()
    pro = rel_l2(interior(b.uh), interior(a.uh))
    for o in (a, b):
        o.rk3_step(float(g["tend"]))
    one = rel_l2(interior(b.uh), interior(a.uh))
    print(f"[fma] {name}: prologue {pro:.2e}, first step {one:.2e}")
    if name.startswith("run_briowu"):
        assert pro == 0.0 and one <= 1e-15
    elif name.startswith("run_rotor"):
        assert pro > 1e-12 and one < 1e-9
    else:
        assert pro == 0.0 and one > 1e-6
