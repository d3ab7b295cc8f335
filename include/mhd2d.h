/*
 * mhd2d.h -- C ABI of the MI355X (gfx950) implementation of one time step of
 * the 2D ideal-MHD P2 globally divergence-free DG solver of
 * GDF_Balsara_P2_0316 (L1ssandra/Freljord-CFD-arsenal).
 *
 * The reference exposes argument-less Fortran subroutines that share state
 * through COMMON blocks (GDF_Balsara_P2_0316/com.txt:19-54).  Here that state
 * lives in an opaque context on one GPU; each entry point below replaces one
 * of those subroutines (cited per function) and keeps its meaning, including
 * the order-dependent semantics of the reference's whole-array statements.
 * Host arrays cross the boundary in the reference's own column-major layout:
 *     uh(0:Nx+1, 0:Ny+1, 10, 8)   Bx(0:Nx, 1:Ny, 3)   By(1:Nx, 0:Ny, 3)
 * so a Fortran caller passes its COMMON arrays unchanged (iso_c_binding).
 *
 * Every function returns 0 on success, a negative value on a HIP or usage
 * error (message in mhd2d_last_error), and is stream-ordered on the
 * context's stream; functions that return scalars synchronise.  Those that
 * read the time loop's device scalars (mhd2d_step and mhd2d_step_graph with
 * t_out, mhd2d_get_time, mhd2d_diagnostics, mhd2d_calculate_dt) return the
 * context's numerical flags instead of 0 when any is set (a positive
 * MHD2D_FLAG_* mask, outputs filled as usual): the reference has no error
 * path, NaNs propagate and a collapsed dt ends its run.  The flags stay set
 * until mhd2d_set_time (and so mhd2d_rk3_prologue) clears them.
 * One host thread per context; contexts are independent.
 */
#ifndef MHD2D_H
#define MHD2D_H

#include <stdint.h>

/* numerical flags (positive returns, see above) */
#define MHD2D_FLAG_NONFINITE 1     /* a non-finite wave speed, dt or diagnostic */
#define MHD2D_FLAG_DT_SMALL 2      /* the CFL dt fell below 1e-7 (RK3.f90:69-71 ends the run) */
#define MHD2D_FLAG_NEG_PRESSURE 4  /* negative pressure at a quadrature point (diagnostics) */

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhd2d_ctx mhd2d_ctx;

/* Riemann solver (com.txt:1 flux_type) */
enum { MHD2D_LF = 1, MHD2D_HLL = 2, MHD2D_HLLC = 3, MHD2D_HLLD = 4 };
/* which cells the TVB limiter touches */
enum {
  MHD2D_IND_NONE = 0,      /* no limiting                                   */
  MHD2D_IND_ALL = 1,       /* every cell: TVB_Limiter_P2.f90 as shipped     */
  MHD2D_IND_KXRCF = 2,     /* cells flagged by KXRCF_Detector (Ck)          */
  MHD2D_IND_MLP = 3        /* learned indicator, KXRCF fallback in the band */
};
enum { MHD2D_FP64 = 0, MHD2D_MIXED = 1 };
/* initial conditions (init_data.f90 + initN.txt) */
enum {
  MHD2D_IC_VORTEX = 1,           /* init1.txt                                */
  MHD2D_IC_ORSZAG_TANG = 3,      /* init3.txt                                */
  MHD2D_IC_ROTOR = 4,            /* init4.txt + Rotor_functions.f90          */
  MHD2D_IC_BLAST = 5,            /* init5.txt + p0.f90 (shock = 1: P0 data)  */
  MHD2D_IC_CLOUD_SHOCK = 6,      /* init6.txt + Cloud_Shock_functions.f90    */
  MHD2D_IC_CLOUD_SHOCK_2 = 61,   /* init6b.txt + Cloud_Shock_functions2.f90  */
  MHD2D_IC_MAGNETIC_LOOP = 7,    /* init7.txt + Loop_functions.f90           */
  MHD2D_IC_BRIO_WU = 8,          /* init8.txt + BrioWu_functions.f90         */
  MHD2D_IC_BRIO_WU_Y = 81,       /* init8b.txt (shock = 1)                   */
  MHD2D_IC_RYU_JONES = 9,        /* init9.txt + RyuJones_functions.f90       */
  MHD2D_IC_RYU_JONES_Y = 91,     /* init9b.txt                               */
  MHD2D_IC_OT_PERTURBED = 103,   /* Orszag-Tang + seeded 16-mode velocity    */
  MHD2D_IC_RIEMANN4 = 104        /* four constant quadrants                  */
};
/* what initN.txt sets besides the fields (domain, boundaries, TVB M, tend,
   the P0 flag `shock` of init_data.f90:92) */
typedef struct mhd2d_ic_info {
  double xa, xb, ya, yb, tend, M;
  int32_t bcR, bcL, bcU, bcD;
  int32_t shock;
  int32_t pad;
} mhd2d_ic_info;
/* arrays readable with mhd2d_debug_get */
enum { MHD2D_DU = 0, MHD2D_DEZ1 = 1, MHD2D_DEZ2 = 2, MHD2D_EZVERTEX = 3, MHD2D_TROUBLE = 4 };

typedef struct mhd2d_config {
  int32_t Nx, Ny;                 /* com.txt:1 (each >= 2)                    */
  int32_t flux_type;              /* MHD2D_LF .. MHD2D_HLLD                   */
  int32_t bcR, bcL, bcU, bcD;     /* 1 periodic, 2 outflow (set_bc.f90)       */
  int32_t pp;                     /* pp_Limiter inside TVB (TVB_Limiter_P2.f90:123) */
  int32_t indicator;              /* MHD2D_IND_*                              */
  int32_t precision;              /* MHD2D_FP64 or MHD2D_MIXED                */
  int32_t reserved;
  double xa, xb, ya, yb;          /* domain (initN.txt)                       */
  double M;                       /* TVB constant                             */
  double CFL;                     /* calculate_dt.f90:8 (0.15 as shipped)     */
  double Ck;                      /* KXRCF threshold                          */
  double mlp_lo, mlp_hi;          /* MLP confidence band -> KXRCF fallback    */
} mhd2d_config;

/* context life cycle; allocates all device memory (no COMMON equivalent) */
int mhd2d_create(const mhd2d_config *cfg, mhd2d_ctx **out);
int mhd2d_destroy(mhd2d_ctx *ctx);
const char *mhd2d_last_error(const mhd2d_ctx *ctx);
int mhd2d_sync(mhd2d_ctx *ctx);

/* host <-> device state in the reference layout (com.txt:22,48) */
int mhd2d_upload(mhd2d_ctx *ctx, const double *uh, const double *Bx, const double *By);
int mhd2d_download(mhd2d_ctx *ctx, double *uh, double *Bx, double *By);

/* init_data.f90:21-121: L2 projection of an initial condition on the GPU.
   extra: IC parameters (perturbed OT: amplitude, seed; rotor: r0, r1, spin,
   p, Bx, rho_in, xc, yc; Riemann4: x0, y0, 4 x (rho,u,v,p,Bx,By)). */
int mhd2d_init_data(mhd2d_ctx *ctx, int32_t ic, const double *extra, int32_t nextra);
/* Binary checkpoint of the state the time loop carries (RK3.f90:58-174):
   uh, the face fields Bx/By and (t, count), so that a run restarts from
   it bit for bit (the reference writes only uh and has no restart).
   Little-endian file = 128-byte header, payload, CRC-32 trailer:
     header  "MHD2DCK1", u32 version (1; 2 from a mixed-form context),
             u32 header bytes 128, i32 Nx, Ny,
             polynomial degree 2, planes 52, flux_type, precision, bcR, bcL,
             bcU, bcD, then f64 xa, xb, ya, yb, t, dt, i64 count,
             u64 payload bytes, zero padding
     payload the 52 active (mode, variable) planes of uh(0:Nx+1, 0:Ny+1, :, :)
             in the order mode 1-6 of variables 1-8, then (7, Bx), (9, Bx),
             (8, By), (10, By), each plane i fastest; then Bx(0:Nx, 1:Ny, 3)
             and By(1:Nx, 0:Ny, 3) in Fortran order (all f64); version 2
             appends the low parts of the double-double mode-1 face values
             the mixed form's CT update carries: Bx(0:Nx, 1:Ny) then
             By(1:Nx, 0:Ny), Fortran order
     trailer u32 CRC-32 (IEEE 802.3, as zlib.crc32) of header + payload
   load checks magic, version, Nx, Ny, size and CRC before touching the
   context, then uploads the state and sets (t, count); a version-1 file
   gives a mixed-form context zero low parts, and a plain context ignores
   the low parts of a version-2 file. */
int mhd2d_checkpoint_save(mhd2d_ctx *ctx, const char *path);
int mhd2d_checkpoint_load(mhd2d_ctx *ctx, const char *path);
/* defaults of one initial condition (initN.txt); -1 for an unknown ic */
int mhd2d_ic_defaults(int32_t ic, mhd2d_ic_info *out);
/* calculate_Az.f90:1-26: mode 1 of Bx and By from the magnetic loop's
   vector potential (Loop_functions.f90:37-51), modes 2-3 untouched */
int mhd2d_calculate_az(mhd2d_ctx *ctx);
/* calculate_L2error.f90:1-37: L2 error of each conserved variable against
   the smooth vortex (init1.txt) translated by (t20, t20) */
int mhd2d_calculate_l2error(mhd2d_ctx *ctx, double t20, double L2[8]);

/* the reference's subroutines */
int mhd2d_set_bc(mhd2d_ctx *ctx);                        /* set_bc.f90:5-27          */
int mhd2d_lh(mhd2d_ctx *ctx);                            /* Lh.f90:1-305             */
int mhd2d_lhez(mhd2d_ctx *ctx);                          /* LhEz.f90:1-96 (after lh) */
int mhd2d_rk_stage(mhd2d_ctx *ctx, int32_t stage, double dt); /* RK3.f90:74-139      */
int mhd2d_tvb_limiter_p2(mhd2d_ctx *ctx);                /* TVB_Limiter_P2.f90:1-191 */
int mhd2d_pp_limiter(mhd2d_ctx *ctx);                    /* pp_Limiter.f90:1-83      */
int mhd2d_kxrcf_detector(mhd2d_ctx *ctx, int64_t *nflag);/* KXRCF_Detector.f90:1-108 */
int mhd2d_div_free_balsara(mhd2d_ctx *ctx);              /* div_free_Balsara.f90     */
int mhd2d_calculate_dt(mhd2d_ctx *ctx, double *dt);      /* calculate_dt.f90:5-27    */
int mhd2d_diagnostics(mhd2d_ctx *ctx, double *umax, double *totaldiv, double *pmin);
                                   /* calculate_umax/_totaldiv/_pmin.f90          */
/* largest |face polynomial - in-cell trace| of normal B over the Gauss points
   of every face, both interior neighbours (no counterpart in the reference:
   the face-jump half of the divergence report; div_free_Balsara.f90:43-58
   makes it zero up to rounding) */
int mhd2d_face_jump(mhd2d_ctx *ctx, double *jump);
/* mhd2d_diagnostics without a context and without allocating: no
   counterpart in the reference; for callers that cannot allocate (code under
   HIP graph capture, serving loops, harnesses that forbid hipMalloc).  uh is
   the caller's device copy of uh(0:Nx+1, 0:Ny+1, 10, 8) in the reference
   layout (column-major doubles), out the caller's device array of 3 doubles
   that receives umax, totaldiv and pmin (calculate_umax/_totaldiv/_pmin.f90);
   the domain gives hx, hy.  The kernels run on `stream` (a hipStream_t; NULL
   is the null stream) and the call does not synchronise: out is ready when
   the stream reaches that point.  The reductions' scratch (the per-block
   partial sums and the atomics) is a static device array of 512.5 KiB that
   each GPU the library runs on holds until the process ends, so Nx * Ny is
   limited to 4096 * 4096 cells (-3 beyond, with a message on stderr).  All
   calls on one GPU share that array: calls that overlap on one GPU (two
   streams, or two host threads) race.  Returns 0, -1 for a null pointer, a
   grid below 2 x 2 or an empty domain, -2 for a HIP error (the numerical
   flags of mhd2d_diagnostics are not reported). */
int mhd2d_diagnostics_noalloc(int32_t Nx, int32_t Ny, double xa, double xb, double ya, double yb, const double *uh,
                              double *out, void *stream);
/* calculate_Ez.f90:1-7: Ez[k] = u2[k]*B1[k] - u1[k]*B2[k] for k < n, on the
   host (the reference's pointwise helper; the vertex kernels evaluate the
   same function on the device).  Needs no context; -1 on a null pointer or
   n < 0. */
int mhd2d_calculate_ez(int64_t n, const double *u1, const double *u2, const double *B1, const double *B2,
                       double *Ez);
int mhd2d_rk3_prologue(mhd2d_ctx *ctx);                  /* RK3.f90:8-18             */
/* one pass of the RK3.f90:58-174 loop body (dt, clip to tend, 3 stages)
   with t and dt kept on the device; t_out (may be NULL) synchronises.  A
   pass that starts at t >= tend changes nothing (the loop test of
   RK3.f90:58), so passes may be queued without reading t back. */
int mhd2d_step(mhd2d_ctx *ctx, double tend, double *t_out);
/* nsteps passes of the same loop body, each replayed as one HIP graph
   launch (captured on first use, on one stream, once per phase of the
   six-step rotation of the context's buffers; kept until mhd2d_destroy);
   same results as nsteps calls of mhd2d_step, bit for bit (mhd2d_step runs
   the vertex EMFs, CT and detection on a second stream beside the rest).  Passes that start at t >= tend change
   nothing, so a caller advances to tend with chunks of passes and one read
   of t per chunk.  nsteps = 0 captures the graphs of all six phases ahead
   and runs nothing.  While mhd2d_profile_enable is on the passes run
   uncaptured. */
int mhd2d_step_graph(mhd2d_ctx *ctx, double tend, int64_t nsteps, double *t_out);
int mhd2d_get_time(mhd2d_ctx *ctx, double *t, double *dt, int64_t *count);
int mhd2d_set_time(mhd2d_ctx *ctx, double t, int64_t count);
/* copies of internal arrays in the reference layout (which = MHD2D_DU ..) */
int mhd2d_debug_get(mhd2d_ctx *ctx, int32_t which, double *host);

/* Live kernel timing (no counterpart in the reference; for benchmarks):
   HIP events bracket every launch on the context's stream while enabled.
   mhd2d_profile_get returns, per kernel class, the summed milliseconds and
   the launch count: 0 x faces, 1 y faces, 2 vertices, 3 cells, 4 CT,
   5 limiter, 6 set_bc, 7 dt, 8 KXRCF, 9 diagnostics. */
int mhd2d_profile_enable(mhd2d_ctx *ctx, int32_t on);
int mhd2d_profile_get(mhd2d_ctx *ctx, double *ms, int64_t *launches, int32_t nclass);
/* The measured HBM ceiling (no counterpart in the reference; for
   benchmarks): a STREAM triad a = b + 1.5 c over three arrays of n doubles
   (n even; the library allocates and frees them), `unroll` (1, 2, 4 or 8)
   16-byte elements per lane, timed with HIP events over `reps` launches
   after one warm-up.  gbs = 24 n bytes / mean launch time; ms_per_launch
   may be NULL.  Needs no context. */
int mhd2d_stream_triad(int64_t n, int32_t unroll, int32_t reps, double *gbs, double *ms_per_launch);

/* ---- learned troubled-cell indicator (no counterpart in the reference) ----
   A 60-64-64-1 MLP (GELU hidden layers, sigmoid output) over a cell's and its
   four neighbours' normalised modal coefficients gives p = P(the plain method
   limits the cell).  With indicator = MHD2D_IND_MLP the network decides:
   cells with p >= mlp_hi are limited (TVB) without KXRCF, the band of cells
   with p >= mlp_lo, dilated by one cell, goes to KXRCF_Detector (the
   fallback for low confidence), and the rest are not limited (mlp_hi >= 1
   leaves every decision in the band to KXRCF).  Inside mhd2d_step the
   network runs once per step, on the first stage's limiter input, and its
   marks serve all three stages.  Parameters are one FP32 vector of MHD2D_MLP_NPARAM
   values: W1[60][64], b1[64], W2[64][64], b2[64], w3[64], b3 (W rows =
   inputs).  Inference runs in BF16 on the matrix cores with FP32
   accumulation. */
#define MHD2D_MLP_NPARAM 8129
#define MHD2D_MLP_NIN 60
int mhd2d_mlp_set_weights(mhd2d_ctx *ctx, const float *params, int32_t n);
/* indicator probability of every interior cell of the current state;
   host_p is (Nx, Ny) column-major (may be NULL) */
int mhd2d_mlp_probability(mhd2d_ctx *ctx, float *host_p);
/* indicator statistics on the current state, read as a limiter input (its
   halos as they are: after mhd2d_rk_stage 1 that is the stage-1 limiter's
   input), out[8] = { cells, label
   (KXRCF flags AND TVB changes), KXRCF flags, decisions (candidates KXRCF
   flags, or the network decides: TVB applied), candidates (cells handed to
   KXRCF or decided), missed (label and
   not a candidate), over (candidate and not label), missed_hydro (missed
   cells whose TVB change includes a non-magnetic variable: changes to Bx, By
   alone are overwritten by div_free_Balsara) }.  In the band the screen can
   only skip TVB applications; only the cells it decides (p >= mlp_hi) may
   be limited where the plain method would not limit them. */
int mhd2d_indicator_stats(mhd2d_ctx *ctx, int64_t *out);

/* FP32 training of the indicator on the GPU.  A trainer owns a device
   dataset of (60 features, label) rows, collected from solver states with
   the plain method's own decision as the label, and the training buffers. */
typedef struct mhd2d_trainer mhd2d_trainer;
int mhd2d_trainer_create(int64_t capacity, int32_t max_batch, mhd2d_trainer **out);
int mhd2d_trainer_destroy(mhd2d_trainer *tr);
const char *mhd2d_trainer_last_error(const mhd2d_trainer *tr);
/* append the labelled cells of ctx's current state (every positive, each
   negative with probability neg_keep) */
int mhd2d_trainer_collect(mhd2d_trainer *tr, mhd2d_ctx *ctx, double neg_keep, uint64_t seed, int64_t *n_total,
                          int64_t *n_pos);
int mhd2d_trainer_size(mhd2d_trainer *tr, int64_t *n, int64_t *npos);
int mhd2d_trainer_clear(mhd2d_trainer *tr);
int mhd2d_trainer_get_data(mhd2d_trainer *tr, int64_t offset, int64_t count, float *X, float *Y, int64_t *idx);
int mhd2d_trainer_set_data(mhd2d_trainer *tr, const float *X, const float *Y, int64_t n);
/* gradient and loss of params over rows [offset, offset + count) (loss =
   mean over rows of w * binary cross-entropy, w = pos_weight for label 1) */
int mhd2d_trainer_grad(mhd2d_trainer *tr, const float *params, int64_t offset, int32_t count, float pos_weight,
                       float *grad, float *loss);
/* Adam (beta 0.9, 0.999, eps 1e-8) for `steps` batches of `batch` random
   rows; params in/out; losses[steps] (may be NULL) */
int mhd2d_trainer_fit(mhd2d_trainer *tr, float *params, int32_t steps, int32_t batch, float lr, float pos_weight,
                      uint64_t seed, float *losses);

#ifdef __cplusplus
}
#endif
#endif
