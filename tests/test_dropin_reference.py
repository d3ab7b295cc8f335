"""The drop-in inside the reference program itself.

oracle/build_ref.sh builds the reference's own main.f90 and RK3.f90 (not
edited: GDF_Balsara_P2_0316/main.f90:1-23, RK3.f90:1-175) twice, with the
same configuration (Orszag-Tang 16 x 16, HLLD, frame output on with
frameMAX = 10, tend = 0.5):
- ref_main_r16_f4: with the reference's own hot-path sources (CPU);
- dropin_r16_f4: with fortran/mhd2d_com_shims.f90 in their place, so that
  RK3's calls of set_bc, Lh, LhEz, TVB_Limiter_P2, div_free_Balsara,
  calculate_dt and the diagnostics go through COMMON (com.txt:19-54) and the
  C ABI to the GPU, while RK3's whole-array statements stay on the host.
Both write what the reference writes: the frames Q1..Q8flash.txt with
T.txt (RK3.f90:29-53,154-172), Q1..Q8.txt (save_solution.f90) and
Xc/Yc/lambda/weight.txt (save_XY.f90).  Bar: every frame and the final
state within 1e-12 relative L2 per variable, the frame times equal to
1e-14.  (The reference opens its frame files on units 1-9, so its stdout,
unit 6, lands in Q6flash.txt: the per-step lines are skipped by reading
only one-number lines.)
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu

REF_DIR = os.path.join(ROOT, "oracle", "_ref")
N = 16


def numbers(path):
    """the one-number lines of a list-directed Fortran output file"""
    out = []
    with open(path, errors="replace") as f:
        for line in f:
            tok = line.split()
            if len(tok) == 1:
                try:
                    out.append(float(tok[0]))
                except ValueError:
                    pass
    return np.array(out)


def run_program(name, cwd):
    exe = os.path.join(REF_DIR, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (oracle/build_ref.sh, where the reference sources are mounted)")
    print(f"[drop-in] running {name}", flush=True)
    r = subprocess.run([exe], cwd=cwd, capture_output=True, text=True, errors="replace", timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def frames(d, n):
    """Qnflash.txt as (frames, 10, N, N) in the reference's d -> j -> i order"""
    v = numbers(os.path.join(d, f"Q{n}flash.txt"))
    per = 10 * N * N
    assert v.size % per == 0, (n, v.size)
    return v.reshape(-1, 10, N, N)


@pytest.fixture(scope="module")
def runs(tmp_path_factory):
    a = tmp_path_factory.mktemp("ref_main")
    b = tmp_path_factory.mktemp("dropin")
    run_program("ref_main_r16_f4", a)
    run_program("dropin_r16_f4", b)
    return str(a), str(b)


def test_dropin_frames_and_times(runs):
    a, b = runs
    ta, tb = numbers(os.path.join(a, "T.txt")), numbers(os.path.join(b, "T.txt"))
    assert ta.size == tb.size and ta.size >= 5
    assert ta[0] == 0.0 and np.allclose(tb, ta, rtol=1e-14, atol=0)
    for n in range(1, 9):
        fa, fb = frames(a, n), frames(b, n)
        assert fa.shape == fb.shape and fa.shape[0] == ta.size
        for k in range(fa.shape[0]):
            err = rel_l2(fb[k], fa[k])
            assert err <= 1e-12, (n, k, err)


def test_dropin_final_state(runs):
    a, b = runs
    for n in range(1, 9):
        qa, qb = numbers(os.path.join(a, f"Q{n}.txt")), numbers(os.path.join(b, f"Q{n}.txt"))
        assert qa.size == qb.size == 10 * N * N
        assert rel_l2(qb, qa) <= 1e-12, n
    for f in ("Xc.txt", "Yc.txt", "lambda.txt", "weight.txt"):
        assert np.array_equal(numbers(os.path.join(a, f)), numbers(os.path.join(b, f)))
