"""mhd2d_diagnostics_noalloc, the allocation-free entry point, through the
C ABI against mhd2d_diagnostics (calculate_umax/_totaldiv/_pmin.f90) on the
same state: bit-identical umax, totaldiv and pmin, on two calls in a row
(the second on a smaller grid, so stale scratch from the first would show),
on the null stream and on a created one, and the scratch maximum rejected
before anything runs.  The test's own device buffers and stream come from
the HIP runtime the library links (libamdhip64.so.7, through ctypes)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


def state(pkg, N, seed):
    s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=(0.0, TWO_PI, 0.0, TWO_PI), M=50.0, CFL=0.2, pp=1,
                   indicator=pkg.IND_KXRCF, Ck=1.0)
    s.init_data(pkg.IC_OT_PERTURBED, [5e-2, seed])
    s.rk3_prologue()
    for _ in range(2):
        s.step()
    d = s.diagnostics()
    uh = s.download()[0]
    s.close()
    return uh, (d["umax"], d["totaldiv"], d["pmin"])


class Hip:
    """the few HIP runtime calls the test needs (the library's own runtime)"""

    def __init__(self, pkg):
        pkg.load_library()
        self.h = C.CDLL("libamdhip64.so.7")

    def chk(self, rc, what):
        assert rc == 0, f"{what}: hipError {rc}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.chk(self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)), "hipMalloc")
        return p

    def free(self, p):
        self.chk(self.h.hipFree(p), "hipFree")

    def h2d(self, dst, a):
        self.chk(self.h.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1), "hipMemcpy H2D")

    def d2h(self, a, src):
        self.chk(self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, C.c_size_t(a.nbytes), 2), "hipMemcpy D2H")

    def stream(self):
        st = C.c_void_p()
        self.chk(self.h.hipStreamCreate(C.byref(st)), "hipStreamCreate")
        return st

    def sync(self, st):
        self.chk(self.h.hipStreamSynchronize(st), "hipStreamSynchronize")


def noalloc(pkg, hip, uh, stream=None):
    N = uh.shape[0] - 2
    flat = np.ascontiguousarray(uh.ravel(order="F"))
    out = np.full(3, np.nan)
    d_uh, d_out = hip.malloc(flat.nbytes), hip.malloc(out.nbytes)
    try:
        hip.h2d(d_uh, flat)
        hip.h2d(d_out, out)
        print(f"[no-alloc] mhd2d_diagnostics_noalloc at {N}x{N}, stream {stream}", flush=True)
        rc = pkg.load_library().mhd2d_diagnostics_noalloc(N, N, 0.0, TWO_PI, 0.0, TWO_PI, d_uh, d_out, stream)
        hip.sync(stream)
        hip.d2h(out, d_out)
    finally:
        hip.free(d_uh)
        hip.free(d_out)
    return rc, tuple(float(v) for v in out)


def test_noalloc_matches_diagnostics(pkg):
    hip = Hip(pkg)
    uh64, ref64 = state(pkg, 64, 11)
    uh32, ref32 = state(pkg, 32, 12)
    rc, got = noalloc(pkg, hip, uh64)
    assert rc == 0 and got == ref64, (got, ref64)
    rc, got = noalloc(pkg, hip, uh32)       # second call, smaller input
    assert rc == 0 and got == ref32, (got, ref32)
    assert ref32 != ref64


def test_noalloc_on_a_created_stream(pkg):
    hip = Hip(pkg)
    uh, ref = state(pkg, 48, 13)
    st = hip.stream()
    try:
        rc, got = noalloc(pkg, hip, uh, st)
    finally:
        hip.chk(hip.h.hipStreamDestroy(st), "hipStreamDestroy")
    assert rc == 0 and got == ref


def test_noalloc_rejects_beyond_its_scratch(pkg):
    L = pkg.load_library()
    one = C.c_void_p(16)     # never dereferenced: the size check comes first
    assert L.mhd2d_diagnostics_noalloc(4097, 4096, 0.0, 1.0, 0.0, 1.0, one, one, None) == -3
    assert L.mhd2d_diagnostics_noalloc(8, 8, 0.0, 1.0, 0.0, 1.0, None, one, None) == -1
    assert L.mhd2d_diagnostics_noalloc(1, 8, 0.0, 1.0, 0.0, 1.0, one, one, None) == -1
