"""Python binding of the gfx950 MHD2D library (C ABI: include/mhd2d.h).

The class :class:`Solver` keeps the names and meaning of the reference's
Fortran subroutines (GDF_Balsara_P2_0316: ``set_bc``, ``Lh``, ``LhEz``,
``TVB_Limiter_P2``, ``pp_Limiter``, ``KXRCF_Detector``, ``div_free_Balsara``,
``calculate_dt``, the diagnostics and ``RK3``) over one device context.
Host arrays use the reference layouts (column-major):
``uh (Nx+2, Ny+2, 10, 8)``, ``Bx (Nx+1, Ny, 3)``, ``By (Nx, Ny+1, 3)``.

There is no CPU fallback: if ``libmhd2d.so`` is missing or no GPU is
visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmhd2d.so")

LF, HLL, HLLC, HLLD = 1, 2, 3, 4
IND_NONE, IND_ALL, IND_KXRCF, IND_MLP = 0, 1, 2, 3
FP64, MIXED = 0, 1
IC_VORTEX, IC_OT, IC_ROTOR, IC_OT_PERTURBED, IC_RIEMANN4 = 1, 3, 4, 103, 104
FLAG_NONFINITE, FLAG_DT_SMALL, FLAG_NEG_PRESSURE = 1, 2, 4   # MHD2D_FLAG_* (positive statuses)
IC_BLAST, IC_CLOUD_SHOCK, IC_CLOUD_SHOCK_2, IC_MAGNETIC_LOOP = 5, 6, 61, 7
IC_BRIO_WU, IC_BRIO_WU_Y, IC_RYU_JONES, IC_RYU_JONES_Y = 8, 81, 9, 91
DU, DEZ1, DEZ2, EZVERTEX, TROUBLE = 0, 1, 2, 3, 4

# exported symbols of include/mhd2d.h
EXPORTS = (
    "mhd2d_create", "mhd2d_destroy", "mhd2d_last_error", "mhd2d_sync", "mhd2d_upload",
    "mhd2d_download", "mhd2d_init_data", "mhd2d_set_bc", "mhd2d_lh", "mhd2d_lhez",
    "mhd2d_rk_stage", "mhd2d_tvb_limiter_p2", "mhd2d_pp_limiter", "mhd2d_kxrcf_detector",
    "mhd2d_div_free_balsara", "mhd2d_calculate_dt", "mhd2d_diagnostics", "mhd2d_rk3_prologue",
    "mhd2d_step", "mhd2d_step_graph", "mhd2d_face_jump", "mhd2d_get_time", "mhd2d_set_time", "mhd2d_debug_get",
    "mhd2d_profile_enable", "mhd2d_profile_get",
    "mhd2d_mlp_set_weights", "mhd2d_mlp_probability", "mhd2d_indicator_stats",
    "mhd2d_trainer_create", "mhd2d_trainer_destroy", "mhd2d_trainer_last_error", "mhd2d_trainer_collect",
    "mhd2d_trainer_size", "mhd2d_trainer_clear", "mhd2d_trainer_get_data", "mhd2d_trainer_set_data",
    "mhd2d_trainer_grad", "mhd2d_trainer_fit",
    "mhd2d_checkpoint_save", "mhd2d_checkpoint_load", "mhd2d_ic_defaults", "mhd2d_calculate_az",
    "mhd2d_calculate_l2error", "mhd2d_calculate_ez", "mhd2d_stream_triad", "mhd2d_diagnostics_noalloc",
)
MLP_NPARAM = 8129        # W1[60][64], b1[64], W2[64][64], b2[64], w3[64], b3
MLP_NIN = 60
WEIGHTS_PATH = os.path.join(HERE, "indicator_weights.npz")
STATS_KEYS = ("cells", "label", "kxrcf", "decision", "candidates", "missed", "over", "missed_hydro")
KERNEL_CLASSES = ("xface", "yface", "vertex", "cell", "ct", "limit", "setbc", "dt", "kxrcf", "diag")


class Config(C.Structure):
    """mhd2d_config (include/mhd2d.h)."""
    _fields_ = [
        ("Nx", C.c_int32), ("Ny", C.c_int32), ("flux_type", C.c_int32),
        ("bcR", C.c_int32), ("bcL", C.c_int32), ("bcU", C.c_int32), ("bcD", C.c_int32),
        ("pp", C.c_int32), ("indicator", C.c_int32), ("precision", C.c_int32), ("reserved", C.c_int32),
        ("xa", C.c_double), ("xb", C.c_double), ("ya", C.c_double), ("yb", C.c_double),
        ("M", C.c_double), ("CFL", C.c_double), ("Ck", C.c_double),
        ("mlp_lo", C.c_double), ("mlp_hi", C.c_double),
    ]


class ICInfo(C.Structure):
    """mhd2d_ic_info (include/mhd2d.h)."""
    _fields_ = [
        ("xa", C.c_double), ("xb", C.c_double), ("ya", C.c_double), ("yb", C.c_double),
        ("tend", C.c_double), ("M", C.c_double),
        ("bcR", C.c_int32), ("bcL", C.c_int32), ("bcU", C.c_int32), ("bcD", C.c_int32),
        ("shock", C.c_int32), ("pad", C.c_int32),
    ]


_lib = None


class MHD2DError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """dlopen libmhd2d.so and declare the C ABI; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MHD2DError(f"{path} is not built (run __graft_entry__.build() or make -C {HERE})")
    L = C.CDLL(path)
    P, dp, i32, i64p = C.c_void_p, C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_int64)
    fp = C.POINTER(C.c_float)
    sig = {
        "mhd2d_create": ([C.POINTER(Config), C.POINTER(C.c_void_p)], C.c_int),
        "mhd2d_destroy": ([P], C.c_int),
        "mhd2d_last_error": ([P], C.c_char_p),
        "mhd2d_sync": ([P], C.c_int),
        "mhd2d_upload": ([P, dp, dp, dp], C.c_int),
        "mhd2d_download": ([P, dp, dp, dp], C.c_int),
        "mhd2d_init_data": ([P, i32, dp, i32], C.c_int),
        "mhd2d_set_bc": ([P], C.c_int),
        "mhd2d_lh": ([P], C.c_int),
        "mhd2d_lhez": ([P], C.c_int),
        "mhd2d_rk_stage": ([P, i32, C.c_double], C.c_int),
        "mhd2d_tvb_limiter_p2": ([P], C.c_int),
        "mhd2d_pp_limiter": ([P], C.c_int),
        "mhd2d_kxrcf_detector": ([P, i64p], C.c_int),
        "mhd2d_div_free_balsara": ([P], C.c_int),
        "mhd2d_calculate_dt": ([P, dp], C.c_int),
        "mhd2d_diagnostics": ([P, dp, dp, dp], C.c_int),
        "mhd2d_rk3_prologue": ([P], C.c_int),
        "mhd2d_step": ([P, C.c_double, dp], C.c_int),
        "mhd2d_step_graph": ([P, C.c_double, C.c_int64, dp], C.c_int),
        "mhd2d_face_jump": ([P, dp], C.c_int),
        "mhd2d_get_time": ([P, dp, dp, i64p], C.c_int),
        "mhd2d_set_time": ([P, C.c_double, C.c_int64], C.c_int),
        "mhd2d_debug_get": ([P, i32, dp], C.c_int),
        "mhd2d_profile_enable": ([P, i32], C.c_int),
        "mhd2d_profile_get": ([P, dp, i64p, i32], C.c_int),
        "mhd2d_mlp_set_weights": ([P, fp, i32], C.c_int),
        "mhd2d_mlp_probability": ([P, fp], C.c_int),
        "mhd2d_indicator_stats": ([P, i64p], C.c_int),
        "mhd2d_trainer_create": ([C.c_int64, i32, C.POINTER(C.c_void_p)], C.c_int),
        "mhd2d_trainer_destroy": ([P], C.c_int),
        "mhd2d_trainer_last_error": ([P], C.c_char_p),
        "mhd2d_trainer_collect": ([P, P, C.c_double, C.c_uint64, i64p, i64p], C.c_int),
        "mhd2d_trainer_size": ([P, i64p, i64p], C.c_int),
        "mhd2d_trainer_clear": ([P], C.c_int),
        "mhd2d_trainer_get_data": ([P, C.c_int64, C.c_int64, fp, fp, i64p], C.c_int),
        "mhd2d_trainer_set_data": ([P, fp, fp, C.c_int64], C.c_int),
        "mhd2d_trainer_grad": ([P, fp, C.c_int64, i32, C.c_float, fp, fp], C.c_int),
        "mhd2d_trainer_fit": ([P, fp, i32, i32, C.c_float, C.c_float, C.c_uint64, fp], C.c_int),
        "mhd2d_checkpoint_save": ([P, C.c_char_p], C.c_int),
        "mhd2d_checkpoint_load": ([P, C.c_char_p], C.c_int),
        "mhd2d_ic_defaults": ([i32, C.POINTER(ICInfo)], C.c_int),
        "mhd2d_calculate_az": ([P], C.c_int),
        "mhd2d_calculate_l2error": ([P, C.c_double, dp], C.c_int),
        "mhd2d_calculate_ez": ([C.c_int64, dp, dp, dp, dp, dp], C.c_int),
        "mhd2d_stream_triad": ([C.c_int64, i32, i32, dp, dp], C.c_int),
        "mhd2d_diagnostics_noalloc": ([i32, i32, C.c_double, C.c_double, C.c_double, C.c_double, P, P, P],
                                      C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _i64p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int64))


def ic_defaults(ic):
    """what initN.txt sets for one initial condition: domain (xa, xb, ya, yb),
    boundary types (bcR, bcL, bcU, bcD), tend, the TVB constant M and the P0
    flag shock"""
    info = ICInfo()
    if load_library().mhd2d_ic_defaults(int(ic), C.byref(info)) != 0:
        raise MHD2DError(f"unknown initial condition {ic}")
    return {"domain": (info.xa, info.xb, info.ya, info.yb), "bc": (info.bcR, info.bcL, info.bcU, info.bcD),
            "tend": info.tend, "M": info.M, "shock": info.shock}


def load_indicator_weights(path=WEIGHTS_PATH):
    """the trained indicator parameters shipped with the package (FP32)"""
    with np.load(path, allow_pickle=False) as z:
        p = np.ascontiguousarray(z["params"], dtype=np.float32)
    if p.size != MLP_NPARAM:
        raise MHD2DError(f"{path}: {p.size} parameters, expected {MLP_NPARAM}")
    return p


def calculate_Ez(u1, u2, B1, B2):
    """calculate_Ez.f90:1-7, elementwise: Ez = u2*B1 - u1*B2 (mhd2d_calculate_ez)"""
    L = load_library()
    a, b, c, d = (np.ascontiguousarray(np.broadcast_to(x, np.broadcast(u1, u2, B1, B2).shape), dtype=np.float64)
                  for x in (u1, u2, B1, B2))
    out = np.empty(a.shape)
    rc = L.mhd2d_calculate_ez(a.size, _dp(a), _dp(b), _dp(c), _dp(d), _dp(out))
    if rc != 0:
        raise MHD2DError(f"mhd2d_calculate_ez failed with status {rc}")
    return out


def stream_triad(n=1 << 28, unroll=4, reps=10):
    """measured HBM ceiling: (GB/s, ms per launch) of a triad over three
    arrays of n doubles (mhd2d_stream_triad)"""
    L = load_library()
    g, ms = C.c_double(), C.c_double()
    rc = L.mhd2d_stream_triad(int(n), int(unroll), int(reps), C.byref(g), C.byref(ms))
    if rc != 0:
        raise MHD2DError(f"mhd2d_stream_triad failed with status {rc}")
    return g.value, ms.value


class Solver:
    """One device context: the reference's COMMON state on the GPU."""

    def __init__(self, Nx, Ny=None, flux_type=HLLD, bc=1, domain=(0.0, 2 * np.pi, 0.0, 2 * np.pi),
                 M=10.0, CFL=0.15, pp=0, indicator=IND_ALL, precision=FP64, Ck=1.0,
                 mlp_band=(0.2, 0.8)):
        self.L = load_library()
        Ny = Nx if Ny is None else Ny
        self.Nx, self.Ny = int(Nx), int(Ny)
        bcs = (bc, bc, bc, bc) if np.isscalar(bc) else tuple(bc)
        self.cfg = Config(self.Nx, self.Ny, flux_type, *bcs, pp, indicator, precision, 0,
                          *domain, M, CFL, Ck, *mlp_band)
        h = C.c_void_p()
        rc = self.L.mhd2d_create(C.byref(self.cfg), C.byref(h))
        if rc != 0:
            raise MHD2DError(f"mhd2d_create failed with status {rc}")
        self.h = h
        self.flags = 0          # the last numerical flags a call returned (cleared with the time)

    def close(self):
        if getattr(self, "h", None):
            self.L.mhd2d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        """negative: a HIP or usage error (raises); positive: the numerical
        flags (FLAG_*) the time loop's scalars carry, kept in self.flags"""
        if rc < 0:
            msg = self.L.mhd2d_last_error(self.h).decode(errors="replace")
            raise MHD2DError(f"{what}: status {rc}: {msg}")
        if rc > 0:
            self.flags = rc
        return rc

    # --- state transfer -------------------------------------------------
    def shapes(self):
        return ((self.Nx + 2, self.Ny + 2, 10, 8), (self.Nx + 1, self.Ny, 3), (self.Nx, self.Ny + 1, 3))

    def upload(self, uh, Bx, By):
        uh = np.asfortranarray(uh, dtype=np.float64)
        Bx = np.asfortranarray(Bx, dtype=np.float64)
        By = np.asfortranarray(By, dtype=np.float64)
        su, sx, sy = self.shapes()
        if uh.shape != su or Bx.shape != sx or By.shape != sy:
            raise ValueError(f"shapes {uh.shape} {Bx.shape} {By.shape} != {su} {sx} {sy}")
        self._chk(self.L.mhd2d_upload(self.h, _dp(uh), _dp(Bx), _dp(By)), "upload")

    def download(self):
        su, sx, sy = self.shapes()
        uh = np.zeros(su, order="F")
        Bx = np.zeros(sx, order="F")
        By = np.zeros(sy, order="F")
        self._chk(self.L.mhd2d_download(self.h, _dp(uh), _dp(Bx), _dp(By)), "download")
        return uh, Bx, By

    def debug_get(self, which):
        su, sx, sy = self.shapes()
        shape = {DU: su, DEZ1: sx, DEZ2: sy, EZVERTEX: (self.Nx + 1, self.Ny + 1), TROUBLE: (self.Nx, self.Ny)}[which]
        a = np.zeros(shape, order="F")
        self._chk(self.L.mhd2d_debug_get(self.h, which, _dp(a)), "debug_get")
        return a

    # --- the reference's subroutines ------------------------------------
    def init_data(self, ic, extra=()):
        ex = np.ascontiguousarray(extra, dtype=np.float64)
        buf = ex if ex.size else np.zeros(1)
        self._chk(self.L.mhd2d_init_data(self.h, ic, _dp(buf), int(ex.size)), "init_data")

    def set_bc(self):
        self._chk(self.L.mhd2d_set_bc(self.h), "set_bc")

    def calculate_Az(self):
        self._chk(self.L.mhd2d_calculate_az(self.h), "calculate_Az")

    def calculate_L2error(self, t20):
        out = np.zeros(8)
        self._chk(self.L.mhd2d_calculate_l2error(self.h, float(t20), _dp(out)), "calculate_L2error")
        return out

    def save_checkpoint(self, path):
        self._chk(self.L.mhd2d_checkpoint_save(self.h, os.fsencode(path)), "checkpoint_save")

    def load_checkpoint(self, path):
        self._chk(self.L.mhd2d_checkpoint_load(self.h, os.fsencode(path)), "checkpoint_load")

    def Lh(self):
        self._chk(self.L.mhd2d_lh(self.h), "Lh")

    def LhEz(self):
        self._chk(self.L.mhd2d_lhez(self.h), "LhEz")

    def rk_stage(self, stage, dt):
        self._chk(self.L.mhd2d_rk_stage(self.h, stage, dt), "rk_stage")

    def TVB_Limiter_P2(self):
        self._chk(self.L.mhd2d_tvb_limiter_p2(self.h), "TVB_Limiter_P2")

    def pp_Limiter(self):
        self._chk(self.L.mhd2d_pp_limiter(self.h), "pp_Limiter")

    def KXRCF_Detector(self):
        n = C.c_int64()
        self._chk(self.L.mhd2d_kxrcf_detector(self.h, C.byref(n)), "KXRCF_Detector")
        return n.value

    def div_free_Balsara(self):
        self._chk(self.L.mhd2d_div_free_balsara(self.h), "div_free_Balsara")

    def calculate_dt(self):
        d = C.c_double()
        self._chk(self.L.mhd2d_calculate_dt(self.h, C.byref(d)), "calculate_dt")
        return d.value

    def diagnostics(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        f = self._chk(self.L.mhd2d_diagnostics(self.h, C.byref(a), C.byref(b), C.byref(c)), "diagnostics")
        return {"umax": a.value, "totaldiv": b.value, "pmin": c.value, "flags": f}

    def face_jump(self):
        """largest mismatch of normal B between a face and the in-cell traces (mhd2d_face_jump)"""
        v = C.c_double()
        self._chk(self.L.mhd2d_face_jump(self.h, C.byref(v)), "face_jump")
        return v.value

    def rk3_prologue(self):
        self.flags = 0
        self._chk(self.L.mhd2d_rk3_prologue(self.h), "rk3_prologue")

    def step(self, tend=float("inf"), sync=False):
        t = C.c_double()
        self._chk(self.L.mhd2d_step(self.h, tend, C.byref(t) if sync else None), "step")
        return t.value if sync else None

    def step_graph(self, nsteps, tend=float("inf"), sync=False):
        """nsteps steps replayed as HIP graph launches (mhd2d_step_graph)"""
        t = C.c_double()
        self._chk(self.L.mhd2d_step_graph(self.h, tend, int(nsteps), C.byref(t) if sync else None), "step_graph")
        return t.value if sync else None

    def get_time(self):
        t, dt, n = C.c_double(), C.c_double(), C.c_int64()
        self._chk(self.L.mhd2d_get_time(self.h, C.byref(t), C.byref(dt), C.byref(n)), "get_time")
        return t.value, dt.value, n.value

    def set_time(self, t, count=0):
        self.flags = 0
        self._chk(self.L.mhd2d_set_time(self.h, t, count), "set_time")

    def sync(self):
        self._chk(self.L.mhd2d_sync(self.h), "sync")

    def profile(self, on=True):
        self._chk(self.L.mhd2d_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_read(self):
        n = len(KERNEL_CLASSES)
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        self._chk(self.L.mhd2d_profile_get(self.h, _dp(ms), cnt.ctypes.data_as(C.POINTER(C.c_int64)), n),
                  "profile_get")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}

    # --- learned indicator ----------------------------------------------
    def mlp_set_weights(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._chk(self.L.mhd2d_mlp_set_weights(self.h, _fp(p), int(p.size)), "mlp_set_weights")

    def mlp_probability(self):
        p = np.zeros((self.Nx, self.Ny), dtype=np.float32, order="F")
        self._chk(self.L.mhd2d_mlp_probability(self.h, _fp(p)), "mlp_probability")
        return p

    def indicator_stats(self):
        out = np.zeros(8, dtype=np.int64)
        self._chk(self.L.mhd2d_indicator_stats(self.h, _i64p(out)), "indicator_stats")
        return dict(zip(STATS_KEYS, (int(v) for v in out)))

    def RK3(self, tend, max_steps=None, chunk=16):
        """RK3.f90:5-174 without frame output: prologue, then steps while
        t < tend, as graph launches in chunks of `chunk` steps (the steps of a
        chunk that start at t >= tend change nothing); returns (t, steps)."""
        self.rk3_prologue()
        t, n = 0.0, 0
        while t < tend and (max_steps is None or n < max_steps):
            k = chunk if max_steps is None else min(chunk, max_steps - n)
            self.step_graph(k, tend)
            t, _, n = self.get_time()
        return t, n


class Trainer:
    """Device dataset + FP32 training of the indicator MLP (mhd2d_trainer_*)."""

    def __init__(self, capacity, max_batch=4096):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.mhd2d_trainer_create(int(capacity), int(max_batch), C.byref(h))
        if rc != 0:
            raise MHD2DError(f"mhd2d_trainer_create failed with status {rc}")
        self.h = h
        self.max_batch = int(max_batch)

    def close(self):
        if getattr(self, "h", None):
            self.L.mhd2d_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.mhd2d_trainer_last_error(self.h).decode(errors="replace")
            raise MHD2DError(f"{what}: status {rc}: {msg}")

    def collect(self, solver, neg_keep, seed):
        n, npos = C.c_int64(), C.c_int64()
        self._chk(self.L.mhd2d_trainer_collect(self.h, solver.h, float(neg_keep), int(seed), C.byref(n),
                                               C.byref(npos)), "collect")
        return n.value, npos.value

    def size(self):
        n, npos = C.c_int64(), C.c_int64()
        self._chk(self.L.mhd2d_trainer_size(self.h, C.byref(n), C.byref(npos)), "size")
        return n.value, npos.value

    def clear(self):
        self._chk(self.L.mhd2d_trainer_clear(self.h), "clear")

    def get_data(self, offset=0, count=None):
        n, _ = self.size()
        count = n - offset if count is None else count
        X = np.zeros((count, MLP_NIN), dtype=np.float32)
        Y = np.zeros(count, dtype=np.float32)
        idx = np.zeros(count, dtype=np.int64)
        self._chk(self.L.mhd2d_trainer_get_data(self.h, offset, count, _fp(X), _fp(Y), _i64p(idx)), "get_data")
        return X, Y, idx

    def set_data(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        self._chk(self.L.mhd2d_trainer_set_data(self.h, _fp(X), _fp(Y), int(Y.size)), "set_data")

    def grad(self, params, offset, count, pos_weight=1.0):
        p = np.ascontiguousarray(params, dtype=np.float32)
        g = np.zeros(MLP_NPARAM, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._chk(self.L.mhd2d_trainer_grad(self.h, _fp(p), int(offset), int(count), float(pos_weight), _fp(g),
                                            _fp(loss)), "grad")
        return float(loss[0]), g

    def fit(self, params, steps, batch, lr, pos_weight=1.0, seed=0):
        p = np.array(params, dtype=np.float32, copy=True)
        losses = np.zeros(steps, dtype=np.float32)
        self._chk(self.L.mhd2d_trainer_fit(self.h, _fp(p), int(steps), int(batch), float(lr), float(pos_weight),
                                           int(seed), _fp(losses)), "fit")
        return p, losses
