"""Python access to the CPU restatement (``mhd2d_oracle.c``) and to the
reference binaries built by ``build_ref.sh``.

TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's
cpu_baseline leg use it as the checker.  The product path never imports it.

Arrays are numpy views in the reference's Fortran layouts
(GDF_Balsara_P2_0316/com.txt:22,48):
  uh (Nx+2, Ny+2, 10, 8), Bx (Nx+1, Ny, 3), By (Nx, Ny+1, 3), order='F'.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmhd2d_oracle.so")
# the FMA-contracted build of the same source (see Makefile)
FMA_LIB_PATH = os.path.join(HERE, "_build", "libmhd2d_oracle_fma.so")
REF_DIR = os.path.join(HERE, "_ref")

LF, HLL, HLLC, HLLD = 1, 2, 3, 4
IND_NONE, IND_ALL, IND_KXRCF = 0, 1, 2
IC_OT, IC_VORTEX, IC_ROTOR, IC_OT_PERTURBED, IC_RIEMANN4 = 3, 1, 4, 103, 104
IC_BLAST, IC_CLOUD_SHOCK, IC_CLOUD_SHOCK_2, IC_MAGNETIC_LOOP = 5, 6, 61, 7
IC_BRIO_WU, IC_BRIO_WU_Y, IC_RYU_JONES, IC_RYU_JONES_Y = 8, 81, 9, 91


class OrcParams(C.Structure):
    _fields_ = [
        ("Nx", C.c_int), ("Ny", C.c_int), ("flux_type", C.c_int),
        ("bcR", C.c_int), ("bcL", C.c_int), ("bcU", C.c_int), ("bcD", C.c_int),
        ("pp", C.c_int), ("indicator", C.c_int), ("nthreads", C.c_int),
        ("xa", C.c_double), ("xb", C.c_double), ("ya", C.c_double), ("yb", C.c_double),
        ("M", C.c_double), ("CFL", C.c_double), ("Ck", C.c_double),
    ]


_lib = None
_libs = {}


def build():
    """Compile the restatement (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(variant=""):
    """the restatement's library; variant "fma" is the FMA-contracted build"""
    global _lib
    if variant:
        if variant not in _libs:
            path = {"fma": FMA_LIB_PATH}[variant]
            if not os.path.exists(path):
                build()
            _libs[variant] = _declare(C.CDLL(path))
        return _libs[variant]
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _declare(C.CDLL(LIB_PATH))
    return _lib


def _declare(L):
    if True:
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [P]
        for name in ("orc_uh", "orc_Bx", "orc_By", "orc_du", "orc_dEz1", "orc_dEz2", "orc_EzVertex"):
            getattr(L, name).restype = dp
            getattr(L, name).argtypes = [P]
        L.orc_trouble.restype = C.POINTER(C.c_int32)
        L.orc_trouble.argtypes = [P]
        L.orc_scalars.argtypes = [P, dp]
        L.orc_set_time.argtypes = [P, C.c_double, C.c_double]
        L.orc_init_data.argtypes = [P, C.c_int, dp, C.c_int]
        for name in ("orc_set_bc", "orc_Lh", "orc_LhEz", "orc_TVB_Limiter_P2", "orc_pp_Limiter",
                     "orc_KXRCF_Detector", "orc_div_free_Balsara", "orc_calculate_dt",
                     "orc_calculate_umax", "orc_calculate_pmin", "orc_calculate_totaldiv",
                     "orc_rk3_prologue", "orc_calculate_Az"):
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = None
        L.orc_calculate_L2error.argtypes = [P, C.c_double, dp]
        L.orc_calculate_L2error.restype = None
        L.orc_rk3_step.argtypes = [P, C.c_double]
        L.orc_rk3_step.restype = C.c_int
        L.orc_rk3_stages.argtypes = [P, C.c_double]
        L.orc_eigenvalueMm.argtypes = [dp, C.c_int, C.c_int, dp, dp]
        L.orc_riemann_1d.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, dp, dp, dp, dp, dp, dp]
        L.orc_vertex_emf.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_vertex_emf.restype = C.c_double
        L.orc_eigenmatrix.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double]
        L.orc_mhd_flux.argtypes = [dp, dp, dp]
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One restatement state (the reference's COMMON blocks)."""

    def __init__(self, Nx, Ny=None, flux_type=HLLD, bc=1, domain=(0.0, 2 * np.pi, 0.0, 2 * np.pi),
                 M=10.0, CFL=0.15, pp=0, indicator=IND_ALL, Ck=1.0, nthreads=0, variant=""):
        Ny = Nx if Ny is None else Ny
        self.Nx, self.Ny = Nx, Ny
        bcs = (bc, bc, bc, bc) if np.isscalar(bc) else tuple(bc)
        self.params = OrcParams(Nx, Ny, flux_type, *bcs, pp, indicator, nthreads,
                                *domain, M, CFL, Ck)
        self.L = lib(variant)
        self.h = self.L.orc_create(C.byref(self.params))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def _view(self, fn, shape):
        n = int(np.prod(shape))
        ptr = getattr(self.L, fn)(self.h)
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(shape, order="F")

    @property
    def uh(self):
        return self._view("orc_uh", (self.Nx + 2, self.Ny + 2, 10, 8))

    @property
    def Bx(self):
        return self._view("orc_Bx", (self.Nx + 1, self.Ny, 3))

    @property
    def By(self):
        return self._view("orc_By", (self.Nx, self.Ny + 1, 3))

    @property
    def du(self):
        return self._view("orc_du", (self.Nx + 2, self.Ny + 2, 10, 8))

    @property
    def dEz1(self):
        return self._view("orc_dEz1", (self.Nx + 1, self.Ny, 3))

    @property
    def dEz2(self):
        return self._view("orc_dEz2", (self.Nx, self.Ny + 1, 3))

    @property
    def EzVertex(self):
        return self._view("orc_EzVertex", (self.Nx + 1, self.Ny + 1))

    @property
    def trouble(self):
        ptr = self.L.orc_trouble(self.h)
        return np.ctypeslib.as_array(ptr, shape=(self.Nx * self.Ny,)).reshape((self.Nx, self.Ny), order="F")

    def scalars(self):
        out = np.zeros(8)
        self.L.orc_scalars(self.h, _dp(out))
        keys = ("t", "dt", "umax", "totaldiv", "pmin", "alphax", "alphay", "count")
        return dict(zip(keys, out))

    def set_time(self, t, dt):
        self.L.orc_set_time(self.h, t, dt)

    def init_data(self, ic, extra=()):
        ex = np.asarray(extra, dtype=np.float64)
        if ex.size == 0:
            ex = np.zeros(1)
            n = 0
        else:
            n = ex.size
        self.L.orc_init_data(self.h, ic, _dp(ex), n)

    def calculate_L2error(self, t20):
        out = np.zeros(8)
        self.L.orc_calculate_L2error(self.h, t20, _dp(out))
        return out

    def load(self, uh, Bx, By):
        self.uh[...] = uh
        self.Bx[...] = Bx
        self.By[...] = By

    def __getattr__(self, name):
        # orc_<name> subroutine wrappers: o.Lh(), o.TVB_Limiter_P2(), ...
        L = self.__dict__.get("L")
        if L is not None and hasattr(L, "orc_" + name):
            fn = getattr(L, "orc_" + name)
            return lambda *a: fn(self.h, *a)
        raise AttributeError(name)


# -------------------------------------------------------------------------
# reference binaries (oracle/_ref/<variant>; see build_ref.sh)

def ref_available(variant):
    return os.path.exists(os.path.join(REF_DIR, variant))


def run_ref(variant, mode, uh, Bx, By, domain=(0.0, 2 * np.pi, 0.0, 2 * np.pi), bc=1, M=10.0,
            tend=0.0, t=0.0, dt=0.0, nsteps=0, t20=0.0, timeout=600):
    """Run one mode of the reference driver; returns a dict of outputs."""
    exe = os.path.join(REF_DIR, variant)
    Nx, Ny = uh.shape[0] - 2, uh.shape[1] - 2
    bcs = (bc, bc, bc, bc) if np.isscalar(bc) else tuple(bc)
    hdr = np.zeros(16)
    hdr[0:4] = domain
    hdr[4:8] = bcs
    hdr[8], hdr[9], hdr[10], hdr[11], hdr[12], hdr[13] = M, tend, t, dt, nsteps, t20
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(hdr.tobytes())
            f.write(np.asfortranarray(uh, dtype=np.float64).tobytes(order="F"))
            f.write(np.asfortranarray(Bx, dtype=np.float64).tobytes(order="F"))
            f.write(np.asfortranarray(By, dtype=np.float64).tobytes(order="F"))
        subprocess.run([exe, mode, fin, fout], cwd=td, check=True, timeout=timeout,
                       stdout=subprocess.DEVNULL)
        raw = np.fromfile(fout, dtype=np.float64)
        l2 = np.fromfile(fout + ".l2", dtype=np.float64) if mode == "l2err" else None
    o = {"L2": l2}
    k = 0

    def take(shape):
        nonlocal k
        n = int(np.prod(shape))
        a = raw[k:k + n].reshape(shape, order="F").copy()
        k += n
        return a
    sc = take((8,))
    o.update(dict(zip(("t", "dt", "umax", "totaldiv", "pmin", "alphax", "alphay", "count"), sc)))
    mesh = take((8,))
    o.update(dict(zip(("xa", "xb", "ya", "yb", "hx", "hy", "M", "tend"), mesh)))
    o["uh"] = take((Nx + 2, Ny + 2, 10, 8))
    o["Bx"] = take((Nx + 1, Ny, 3))
    o["By"] = take((Nx, Ny + 1, 3))
    o["du"] = take((Nx + 2, Ny + 2, 10, 8))
    o["dEz1"] = take((Nx + 1, Ny, 3))
    o["dEz2"] = take((Nx, Ny + 1, 3))
    o["EzVertex"] = take((Nx + 1, Ny + 1))
    o["trouble"] = take((Nx, Ny)).astype(np.int32)
    return o


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    return num / den if den > 0 else num
