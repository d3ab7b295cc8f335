#!/usr/bin/env python3
"""Throughput of the 2D ideal-MHD P2 DG time step on one MI355X.

Workload: seeded perturbed Orszag-Tang (16 velocity modes, amplitude 1e-2,
seed 20261203) on [0, 2pi]^2 periodic at 2048 x 2048 cells, P2 (48 DOF per
cell), HLLD fluxes and corner EMF, TVB limiter (M = 50) gated by the KXRCF
detector (Ck = 1), positivity limiter on, Balsara divergence-free
reconstruction, SSP-RK3, CFL 0.2.  A step is one pass of the reference's
RK3 loop body: the dt reduction plus three stages.  Inputs are resident in
HBM before timing.

Runs the plain FP64 form (the correctness form), the plain form with TVB on
every cell (the reference as shipped, for information) and the headline
form; prints ONE JSON line.  ``value`` is DOF-updates/s
(48 * N^2 * steps / seconds) of the headline form.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "freljord-cfd-arsenal_amd")
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DOF_PER_CELL = 48

# Algorithmic bytes of the step (the roofline's numerator).  The independent
# state is 42 values per cell: 36 cell modes (6 variables x 6 modes; the
# in-cell Bx, By modes are derived from the faces by div_free_Balsara) and
# 6 face-B modes.  SSP-RK3 keeping U^n moves 8 state-sized transfers per
# step (stage 1: read U^n, write U1; stage 2: read U1 and U^n, write U2;
# stage 3: read U2 and U^n, write U^n+1).  Plain FP64: 42 x 8 B x 8 =
# 2688 B per cell per step.  The mixed storage holds the 6 cell averages
# and the 6 face modes in FP64 and the 30 higher modes in FP32: 216 B per
# cell, 1728 B per cell per step.  A stage is a third of a step.
STEP_ALG_BYTES = {"fp64": 42 * 8 * 8, "mixed": (12 * 8 + 30 * 4) * 8}
STAGE_ALG_BYTES = {k: v / 3.0 for k, v in STEP_ALG_BYTES.items()}


# Bytes each kernel class reads and writes per interior cell and launch as
# this implementation lays them out (not the roofline's numerator: the face
# moments' round trip, halo re-reads and the separate passes are
# implementation traffic).  State: 52 FP64 planes (416 B/cell) in the plain
# form, 296 B/cell in the mixed storage; face B: 6 FP64 planes (48 B/cell).
# Face fluxes (40 values per face), face EMFs (5 per face) are FP64 in the
# plain form and FP32 in the mixed form (bytes B below).  One x face and one
# y face belong to each cell.
def io_bytes(form):
    B = 8 if form == "fp64" else 4
    # one state: 52 FP64 planes; mixed storage: 22 FP64 planes (cell
    # averages, magnetic modes) + 30 FP32 planes (higher modes of the rest)
    S = 416 if form == "fp64" else 22 * 8 + 30 * 4
    # values per face: fluxes and EMFs at the 5 Gauss points (plain form) or
    # their Legendre moments (mixed form: 25 flux and 2 EMF moments)
    NF, NE = (40, 5) if form == "fp64" else (25, 2)
    # face B per cell: 3 modes on one x and one y face; the mixed form adds
    # the double-double low parts of mode 1
    FB = 48 if form == "fp64" else 64
    return {
        "xface": S + (NF + NE) * B,   # state in; face fluxes and EMFs out
        "yface": S + (NF + NE) * B,
        "vertex": S + 8,              # state in; one FP64 vertex EMF out
        # state + x/y fluxes in, stage value out; the RK inputs average
        # (0 + 2 + 2) / 3 state reads (stages 2 and 3 read U0 and the
        # unlimited previous stage value); the plain form's cell kernel also
        # runs the limiter pass without TVB (new faces in, limited value out)
        "cell": S + 2 * NF * B + S + 4 * S / 3.0 + (S + 48 if form == "fp64" else 0),
        "ct": 2 * NE * B + 8 + FB + FB + 4 * FB / 3.0,   # face + vertex EMFs, faces in/out, RK inputs
        # limiter pass: the detector/TVB kernel reads the state (KXRCF
        # traces), the apply pass reads state + faces and writes the state
        # (plain form: only the detector; the TVB-changed cells are sparse)
        "limit": S + 8 if form == "fp64" else S + S + 48 + S + 8,
        "setbc": 0.0,
        "dt": 64,                       # 8 cell averages
        "kxrcf": S,
        "diag": S,
    }


# kernel of each class in the traffic summary (the limiter class also runs
# the detector and the sparse TVB pass, whose traffic is listed on its own)
KERNEL_SYMBOL = {"xface": "k_xface", "yface": "k_yface", "vertex": "k_vertex", "cell": "k_cell", "ct": "k_ct",
                 "limit": "k_limit", "dt": "k_dt", "kxrcf": "k_kxrcf", "diag": "k_diag", "setbc": "k_setbc"}


def measured_traffic(kernel, form, N):
    """HBM bytes per launch of one kernel class, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE runs, scaled by the
    same counters on a plain streaming kernel); None if not measured."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_ot{N}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) >= 5 and cols[0] == KERNEL_SYMBOL.get(kernel) and cols[1] == form:
            try:
                return float(cols[4])
            except ValueError:
                return None
    return None


def indicator_mod():
    import importlib
    return importlib.import_module("freljord_cfd_arsenal_amd.indicator")


def import_pkg():
    name = "freljord_cfd_arsenal_amd"
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(N=512, steps=1):
    """the reference Fortran (OpenMP build in oracle/_ref) on a bounded
    sample of the same workload and method (TVB gated by KXRCF, positivity,
    HLLD): N x N cells, `steps` RK3 steps"""
    from oracle import oracle as O
    variant = f"r{N}_f4_pp_kx_omp"
    if not O.ref_available(variant):
        return None
    # the box's CPU share for one GPU is 16 threads; Lh.f90:9 asks for 16
    threads = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    o = O.Oracle(N, N, flux_type=O.HLLD, domain=(0.0, 2 * np.pi, 0.0, 2 * np.pi), M=50.0, CFL=0.2, pp=1,
                 nthreads=1)
    o.init_data(O.IC_OT_PERTURBED, [1e-2, 20261203])
    env_backup = dict(os.environ)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    try:
        r = O.run_ref(variant, "bench", o.uh, o.Bx, o.By, M=50.0, nsteps=steps, timeout=900)
    finally:
        os.environ.clear()
        os.environ.update(env_backup)
    secs = float(r["count"])      # the bench mode returns the timed seconds in this slot
    return {"value": DOF_PER_CELL * N * N * steps / secs, "unit": "DOF-updates/s", "cores": threads,
            "kind": "reference", "host_logical_cpus": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": f"reference Fortran (flang -O3 -fopenmp) on the same workload and method (perturbed "
                      f"Orszag-Tang, HLLD, TVB M=50 gated by KXRCF Ck=1, positivity) at {N}x{N}, {steps} RK3 "
                      f"steps, {secs:.1f} s, OMP_NUM_THREADS={threads}; only Lh is OpenMP-parallel "
                      f"(Lh.f90:9-300), LhEz, the limiters, the RK statements and the reductions run on one "
                      f"thread; per-step cost does not depend on the CFL (0.15 as shipped)"}


def counters_run(pkg, N, dom):
    """Calibration dispatches plus two steps of each benchmarked form (plain
    FP64, mixed, mixed + learned indicator), for rocprofv3 --pmc.

    The calibration kernel is k_axpy of the per-subroutine stage update
    (mhd2d_rk_stage 1 with dt = 0): over the 52 state planes it reads two
    arrays and writes one, n = 52 * plane doubles each, so its FETCH_SIZE
    and WRITE_SIZE must read 16 n and 8 n bytes; the ratio of measured to
    known bytes scales the counters of the other kernels."""
    s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1)
    s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 20261203])
    s.rk_stage(1, 0.0)
    s.sync()
    P = (N + 2 + 15 + 15) // 16 * 16
    n = 52 * (N + 2) * P
    print(f"[counters] k_axpy over {n} doubles: reads {16 * n} B, writes {8 * n} B", file=sys.stderr, flush=True)
    s.close()
    # a second calibration kernel: the triad over three arrays of 2^28
    # doubles (one warm-up and one timed launch, each reading 4 GiB and
    # writing 2 GiB)
    pkg.stream_triad(1 << 28, 1, 1)
    print(f"[counters] k_triad<1> over {1 << 28} doubles: reads {16 << 28} B, writes {8 << 28} B", file=sys.stderr,
          flush=True)
    z = np.load(pkg.WEIGHTS_PATH, allow_pickle=False)
    params, band = pkg.load_indicator_weights(), (float(z["mlp_lo"]), float(z["mlp_hi"]))
    for name, prec, ind in (("fp64", pkg.FP64, pkg.IND_KXRCF), ("mixed", pkg.MIXED, pkg.IND_KXRCF),
                            ("mixed_nn", pkg.MIXED, pkg.IND_MLP)):
        s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=ind,
                       precision=prec, Ck=1.0, mlp_band=band)
        s.mlp_set_weights(params)
        s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 20261203])
        s.rk3_prologue()
        s.step()
        s.step()
        s.sync()
        s.close()
        print(f"[counters] {name}: two steps at {N}^2 done", file=sys.stderr, flush=True)


FIELDS = ("rho", "rhou", "rhov", "rhow", "E", "Bx", "By", "Bz")


# diagonal mass matrix of the ten modes (get_basis.f90:60-159)
MODE_NORM = np.array([1.0, 1 / 3, 1 / 3, 4 / 45, 1 / 9, 4 / 45, 4 / 175, 4 / 135, 4 / 135, 4 / 175])


def rel_l2_fields(a, b):
    """per-field relative L2 over the interior cells: the modal coefficients
    weighted by the mass matrix, which for the orthogonal basis equals the
    5 x 5 Gauss quadrature of the squared polynomials (calculate_L2error.f90)"""
    a, b = a[1:-1, 1:-1], b[1:-1, 1:-1]
    w = np.sqrt(MODE_NORM[:a.shape[2]])[None, None, :]
    out = {}
    for n, nm in enumerate(FIELDS):
        bn = b[:, :, :, n] * w
        den = np.linalg.norm(bn)
        if den > 0:
            out[nm] = float(np.linalg.norm(a[:, :, :, n] * w - bn) / den)
    return out


def triad_ceiling(pkg, n=1 << 28, reps=10):
    """the measured HBM ceiling on this GPU: the in-tree triad kernel
    (mhd2d_stream_triad) over three arrays of n doubles (6 GiB per launch),
    best over its unroll variants"""
    best = None
    for u in (1, 2, 4, 8):
        gbs, ms = pkg.stream_triad(n, u, reps)
        if best is None or gbs > best["gbs"]:
            best = {"gbs": gbs, "unroll": u, "ms_per_launch": ms, "bytes_per_launch": 24 * n}
    return best


def error_run(pkg, N, tend, dom, nn):
    """the error statement past the onset of limiting: the plain FP64 method
    and the headline form advanced from the same input to t = tend at N^2
    (queued steps until t reaches tend), per-field relative L2 of the mixed
    state against the FP64 one, both forms' div B, and the indicator
    statistics of the FP64 run's last limiter input"""
    out = {"N": N, "t_end": tend}
    states = {}
    for name, prec, ind in (("fp64", pkg.FP64, pkg.IND_KXRCF), ("mixed_nn", pkg.MIXED, pkg.IND_MLP)):
        if name == "mixed_nn" and nn is None:
            continue
        band = (nn[1], nn[2]) if nn else (0.2, 0.8)
        s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=ind,
                       precision=prec, Ck=1.0, mlp_band=band)
        if nn:
            s.mlp_set_weights(nn[0])
        s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 20261203])
        t0 = time.perf_counter()
        t, steps = s.RK3(tend)
        secs = time.perf_counter() - t0
        d = s.diagnostics()
        states[name] = s.download()[0]
        out[name] = {"steps": steps, "t": t, "seconds": secs, "totaldiv": d["totaldiv"], "pmin": d["pmin"],
                     "face_jump": s.face_jump()}
        if name == "mixed_nn":
            out[name]["indicator_stats"] = indicator_mod().limiter_input_stats(s)
        s.close()
        print(f"[bench] error run {name}: {steps} steps to t={t:.4f} at {N}^2 in {secs:.1f} s",
              file=sys.stderr, flush=True)
    if "mixed_nn" in states:
        out["rel_l2_vs_fp64"] = rel_l2_fields(states["mixed_nn"], states["fp64"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--form", choices=["fp64", "mixed", "mixed_nn"], default="mixed_nn",
                    help="the form the headline value is measured on (the plain FP64 form always runs too); "
                         "mixed_nn = mixed precision + the learned troubled-cell indicator")
    ap.add_argument("--all-cells", type=int, default=1,
                    help="also time the plain form with TVB on every cell (the reference as shipped)")
    ap.add_argument("--error", type=int, default=1, help="report mixed-vs-FP64 L2 errors (downloads both states)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--error-n", type=int, default=1024,
                    help="grid of the longer error run (both forms to --error-t, through the shocks); 0 = off")
    ap.add_argument("--error-t", type=float, default=0.5)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--graph", action="store_true",
                    help="time HIP-graph replays of the step (mhd2d_step_graph, one stream) instead of queued "
                         "launches on two streams")
    ap.add_argument("--one-stream", action="store_true",
                    help="run every step eagerly on one stream, as the profiled steps are (for a kernel trace "
                         "whose per-kernel durations are not shared with the overlapped second stream)")
    ap.add_argument("--counters", action="store_true",
                    help="no timing: the dispatches a rocprofv3 --pmc pass reads (a streaming kernel of known "
                         "size, then two steps of each form)")
    args = ap.parse_args()
    if args.profile_steps < 1 and not args.counters:
        ap.error("--profile-steps must be at least 1: the roofline reads the event-timed kernel classes")
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (no sharding, no collectives); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    pkg = import_pkg()
    N = args.n
    dom = (0.0, 2 * np.pi, 0.0, 2 * np.pi)
    if args.counters:
        counters_run(pkg, N, dom)
        return

    nn = None
    if args.form == "mixed_nn":
        z = np.load(pkg.WEIGHTS_PATH, allow_pickle=False)
        nn = (pkg.load_indicator_weights(), float(z["mlp_lo"]), float(z["mlp_hi"]))

    def run(prec, ind, steps, warmup, profile_steps):
        band = (nn[1], nn[2]) if nn else (0.2, 0.8)
        s = pkg.Solver(N, N, flux_type=pkg.HLLD, domain=dom, M=50.0, CFL=0.2, pp=1, indicator=ind,
                       precision=prec, Ck=1.0, mlp_band=band)
        if nn:
            s.mlp_set_weights(nn[0])
        s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 20261203])
        s.rk3_prologue()
        if args.one_stream:        # every step eager on one stream (the per-kernel timing mode)
            s.profile(True)
        if args.graph:
            s.step_graph(0)        # capture the step's HIP graphs (six buffer-role phases) ahead
            s.step_graph(warmup)
        else:
            for _ in range(warmup):
                s.step()
        s.sync()
        t0 = time.perf_counter()
        if args.graph:
            s.step_graph(steps)    # one graph launch per step (one stream), no host round trip
        else:
            for _ in range(steps):
                s.step()           # queued launches on two streams, no host round trip
        s.sync()
        secs = time.perf_counter() - t0
        state = s.download() if args.error else None
        diag = s.diagnostics()
        diag["face_jump"] = s.face_jump()
        tnow = s.get_time()[0]
        prof = None
        if profile_steps:
            s.profile(True)
            for _ in range(profile_steps):
                s.step()
            prof = s.profile_read()
            s.profile(False)
        # screen vs plain method on the next stage-1 limiter input (changes the state: last)
        istats = indicator_mod().limiter_input_stats(s) if (nn and ind == pkg.IND_MLP) else None
        s.close()
        return dict(secs=secs, state=state, diag=diag, t=tnow, prof=prof, ms=1e3 * secs / steps,
                    value=DOF_PER_CELL * N * N * steps / secs, istats=istats)

    # the plain method of the workload: FP64, TVB gated by KXRCF (Ck = 1)
    plan = [("fp64", pkg.FP64, pkg.IND_KXRCF)]
    if args.all_cells:
        plan.append(("fp64_all_cells", pkg.FP64, pkg.IND_ALL))
    if args.form in ("mixed", "mixed_nn"):
        plan.append(("mixed", pkg.MIXED, pkg.IND_KXRCF))
    if args.form == "mixed_nn":
        plan.append(("mixed_nn", pkg.MIXED, pkg.IND_MLP))
    results = {}
    for name, prec, ind in plan:
        r = run(prec, ind, args.steps, args.warmup, args.profile_steps)
        results[name] = r
        print(f"[bench] {name}: {args.steps} steps at {N}^2: {r['ms']:.3f} ms/step, t={r['t']:.6f}",
              file=sys.stderr, flush=True)
        print(f"[bench] {name} kernel ms per step: " +
              ", ".join(f"{k} {v[0] / args.profile_steps:.2f}" for k, v in r["prof"].items() if v[1]),
              file=sys.stderr, flush=True)
    head = results[args.form]
    ms_per_step, value, prof, d = head["ms"], head["value"], head["prof"], head["diag"]
    errors = err_screen = None
    if args.form != "fp64" and args.error:
        errors = rel_l2_fields(head["state"][0], results["fp64"]["state"][0])
    if args.form == "mixed_nn" and args.error and "mixed" in results:
        # what the learned screen alone changes: same mixed arithmetic, KXRCF on every cell
        err_screen = rel_l2_fields(head["state"][0], results["mixed"]["state"][0])
    storage = "fp64" if args.form == "fp64" else "mixed"
    dom_k = max(prof, key=lambda k: prof[k][0])
    ms_tot, launches = prof[dom_k]
    avg_ms = ms_tot / max(launches, 1)
    # roofline numerator: the algorithmic bytes of one RK stage (the dominant
    # kernel runs once per stage) over its mean launch time
    bytes_launch = STAGE_ALG_BYTES[storage] * N * N
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    io_launch = io_bytes(storage)[dom_k] * N * N
    step_gbs = STEP_ALG_BYTES[storage] * N * N / (ms_per_step * 1e-3) / 1e9
    fp64_step_gbs = STEP_ALG_BYTES["fp64"] * N * N / (results["fp64"]["ms"] * 1e-3) / 1e9
    traffic = measured_traffic(dom_k, args.form, N)
    ceiling = triad_ceiling(pkg)
    cpu = None if args.no_cpu else cpu_baseline()
    err_run = error_run(pkg, args.error_n, args.error_t, dom, nn) if (args.error_n and args.form != "fp64") else None
    per_step = lambda r: {k: v[0] / args.profile_steps for k, v in r["prof"].items() if v[1]}
    out = {
        "metric": "P2 MHD DOF-updates/s @2048^2: mixed+NN vs plain FP64, same GPU, at stated L2 err",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f32+f64" + ("+bf16" if args.form == "mixed_nn" else ""),
        "data": "synthetic",
        "config": {"workload": f"perturbed Orszag-Tang {N}x{N}, P2 DG (48 DOF/cell), HLLD, TVB M=50 gated by "
                               f"KXRCF (Ck=1), positivity, SSP-RK3, CFL 0.2", "N": N, "form": args.form,
                   "indicator": "learned MLP, KXRCF in the band" if args.form == "mixed_nn" else "KXRCF",
                   "stepping": "HIP graph replays, one stream" if args.graph else
                               ("queued launches, one stream" if args.one_stream else "queued launches, two streams"),
                   "totaldiv": d["totaldiv"], "divB_face_jump": d["face_jump"], "pmin": d["pmin"],
                   "fp64_value": results["fp64"]["value"], "fp64_ms_per_step": results["fp64"]["ms"],
                   "speedup_vs_fp64": value / results["fp64"]["value"],
                   "rel_l2_vs_fp64": errors,
                   "rel_l2_vs_mixed_kxrcf": err_screen,
                   "fp64_all_cells_value": results.get("fp64_all_cells", {}).get("value"),
                   "fp64_all_cells_ms_per_step": results.get("fp64_all_cells", {}).get("ms"),
                   "mixed_kxrcf_value": results.get("mixed", {}).get("value"),
                   "indicator_stats": head.get("istats"),
                   "mlp_band": [nn[1], nn[2]] if nn else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": dom_k, "alg_bytes_per_launch": bytes_launch,
                     "alg_bytes_per_cell_stage": STAGE_ALG_BYTES[storage],
                     "traffic_over_alg": traffic / bytes_launch if traffic else None,
                     "kernel_io_bytes_per_launch": io_launch,
                     "kernel_avg_ms": avg_ms, "step_achieved": step_gbs,
                     "step_alg_bytes_per_cell": STEP_ALG_BYTES[storage],
                     "step_frac": step_gbs / HBM_PEAK_GBS,
                     "fp64_step_achieved": fp64_step_gbs, "fp64_step_frac": fp64_step_gbs / HBM_PEAK_GBS,
                     "ceiling": ceiling["gbs"], "ceiling_kernel": f"mhd2d_stream_triad (unroll "
                                                                  f"{ceiling['unroll']}, 6 GiB per launch)",
                     "frac_of_ceiling": achieved / ceiling["gbs"],
                     "step_frac_of_ceiling": step_gbs / ceiling["gbs"]},
        "error_run": err_run,
        "cpu_baseline": cpu,
        "kernels_ms_per_step": per_step(head),
        "kernels_ms_per_step_by_run": {k: per_step(r) for k, r in results.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
