// mhd2d_internal.hpp -- declarations shared by the translation units of
// libmhd2d.so: device scalars, the context, launch helpers and the launch
// functions each kernel file provides.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/mhd2d.h"
#include "mhd2d_physics.hpp"
#include "mhd2d_tables.hpp"
#include "mhd2d_tvb.hpp"

namespace mhd2d_impl {
using namespace mhd;

constexpr int BS = 256;

// device scalars of the time loop (COMMON /time/ and the diagnostics)
struct DevScal {
  double t, dt, tend, CFL;
  double alphax, alphay, umax, pmin, totaldiv;
  unsigned long long ax_bits, ay_bits, umax_bits;
  long long pmin_key;
  long long count;
  long long nflag;
  unsigned int nlist;   // length of the troubled-cell list of the current limiter pass
  unsigned int nband;   // length of the indicator's candidate list (cells handed to KXRCF)
  unsigned int npp;     // length of the list of cells the positivity limiter must look at (fused stages)
  int active;
  int health;           // MHD2D_FLAG_* bits, sticky until mhd2d_set_time
  long long epoch;      // passes that advanced, since the context was created (never reset): the
                        // fused stages' limiter lists mark cells with 4 epoch + stage (lflag)
};

__device__ __forceinline__ long long dkey(double x) {   // order-preserving key of a double
  long long b = __double_as_longlong(x);
  return b >= 0 ? b : b ^ 0x7fffffffffffffffLL;
}
__device__ __forceinline__ double dunkey(long long k) {
  return __longlong_as_double(k >= 0 ? k : k ^ 0x7fffffffffffffffLL);
}

// a step that starts at t >= tend changes nothing (the reference's loop
// test): inside mhd2d_step every kernel reads the step's active flag, set by
// k_dt_final, through g.gate and returns at once when it is 0
__device__ __forceinline__ bool step_off(const Geo& g) { return g.gate && *g.gate == 0; }
__device__ __forceinline__ bool step_off(const int* gate) { return gate && *gate == 0; }

// blocks of consecutive ids go to the same XCD (round-robin dispatch over 8)
__device__ __forceinline__ long long xcd_block() {
  const unsigned b = blockIdx.x, n = gridDim.x;
  const unsigned full = n & ~7u;
  if (b >= full) return b;
  return (long long)(b & 7u) * (full >> 3) + (b >> 3);
}

// select one of five compile-time node values by a lane-varying index
__device__ __forceinline__ double sel5(int q, double a0, double a1, double a2, double a3, double a4) {
  double r = a0;
  r = q == 1 ? a1 : r;
  r = q == 2 ? a2 : r;
  r = q == 3 ? a3 : r;
  r = q == 4 ? a4 : r;
  return r;
}

// trace of all 8 variables of one cell on one side at node q, in type T
template <int SIDE, typename T, class SV>
__device__ __forceinline__ void trace_at(const SV& U, long long c, int q, T out[8]) {
  constexpr double L0 = lam(0), L1 = lam(1), L2 = lam(2), L3 = lam(3), L4 = lam(4);
  const double s = sel5(q, L0, L1, L2, L3, L4);
  const double p2 = sel5(q, P2(L0), P2(L1), P2(L2), P2(L3), P2(L4));
  const double p3 = sel5(q, P3(L0), P3(L1), P3(L2), P3(L3), P3(L4));
  const double t23 = sel5(q, T23(L0), T23(L1), T23(L2), T23(L3), T23(L4));
#pragma unroll
  for (int n = 0; n < 8; n++) {
    T acc = T(0);
#pragma unroll
    for (int b = 0; b < 10; b++) {
      const int p = PL(b, n);
      if (p >= 0) acc += T(U(p, c)) * T(phiTr(SIDE, b, s, p2, p3, t23));
    }
    out[n] = acc;
  }
}

// ---------------------------------------------------------------------------
// initial conditions (init_data.f90:21-121; init1/3/4.txt; Rotor_functions.f90)
struct ICPar {
  int ic;
  int flags;            // IC_FULL_ENERGY | IC_SHOCK
  double par[32];
  double kx[16], ky[16], phu[16], phv[16];
};

// kinetic and magnetic energy include w and Bz (init5-9b); P0 data (shock = 1)
constexpr int IC_FULL_ENERGY = 1, IC_SHOCK = 2;

// the indicator MLP 60-64-64-1: parameter vector layout (FP32)
constexpr int MLP_NIN = 60;
constexpr int MLP_OFF_W1 = 0;                        // [60][64], row = input
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + 60 * 64;     // [64]
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + 64;          // [64][64]
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + 64 * 64;     // [64]
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + 64;          // [64]
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + 64;          // [1]
constexpr int MLP_NPARAM = MLP_OFF_B3 + 1;           // 8129

// mixed form: face moments M_k = sum_q w_q F(s_q) L_k(s_q) of the fluxes,
// L = {1, s, P2(s)}, plane FM_N * n + k for variable n; plane FM_P3 holds
// the P3(s) moment of the one variable that needs it (By on x faces, Bx on
// y faces); EMF planes 0 and 1 hold its L_0 and L_1 moments
constexpr int FM_N = 3, FM_P3 = 24;

// face buffers: planes 0..2 Bx modes on x faces, 3..5 By modes on y faces;
// the mixed form adds the low parts of the double-double mode-1 values
// (plane FACE_LO_X: Bx on x faces, FACE_LO_Y: By on y faces) that its CT
// update carries, so that the discrete divergence of the faces does not
// accumulate rounding from step to step
constexpr int FACE_PL = 6, FACE_PL_MIX = 8, FACE_LO_X = 6, FACE_LO_Y = 7;

// kernel classes for mhd2d_profile_get
enum { KC_XFACE = 0, KC_YFACE, KC_VERTEX, KC_CELL, KC_CT, KC_LIMIT, KC_SETBC, KC_DT, KC_KXRCF, KC_DIAG, KC_N };

}  // namespace mhd2d_impl

// ===========================================================================
// host side

namespace mhd2d_impl {
// one captured step: the buffer roles it was captured for and the ones it
// leaves (the roles of S, U0, UI, UII, T, F, F0, FI, FII)
struct StepGraph {
  double* before[9];
  double* after[9];
  hipGraphExec_t exec;
};
}  // namespace mhd2d_impl

struct mhd2d_ctx {
  mhd2d_config cfg;
  mhd::Geo g;
  hipStream_t stream;
  hipStream_t stream2;   // inside mhd2d_step: the vertex EMFs and CT beside the face and cell kernels
  hipEvent_t ev_fork, ev_faces, ev_join, ev_cell, ev_det;
  int device;
  // state buffers (52 planes each, or the mixed-form storage of
  // mhd2d_tables.hpp); roles rotate inside mhd2d_step
  double *S, *U0, *UI, *UII, *T;
  double* du;
  double* stage;       // mixed form only: 52 FP64 planes for upload/download (allocated on first use)
  // face buffers (6 planes each)
  double *F, *F0, *FI, *FII, *Ft, *dEz;
  // face fluxes and EMFs
  double *FX, *FY, *EZX, *EZY, *EZV;
  int* trouble;        // detector decision per interior cell (1 = TVB applied)
  int* tmask;          // TVB change bits per interior cell (0 = unchanged)
  int* tlist;          // interior-cell indices the detector flagged (length sc->nlist)
  int* blist;          // interior-cell indices the indicator hands to KXRCF (sc->nband)
  int* screen;         // indicator: 1 where p >= mlp_lo (before the one-cell dilation)
  int* pplist;         // fused stages: interior cells the cell kernel's epilogue leaves to positivity (sc->npp)
  int* lflag;          // fused stages: 4 epoch + stage where a list kernel limited the cell in place
  double* trec;        // modified edge values of changed cells: 32 planes over the interior cells
  void* mlpw;          // indicator MLP weights in MFMA operand order (mhd2d_mlp.hip)
  float* prob;         // indicator MLP probability per interior cell
  mhd2d_impl::DevScal* sc;
  double* partial;
  mhd2d_impl::ICPar* icp;
  long long nblk_cells;
  bool lh_done;
  bool capturing;      // inside capture_step: the stage stays on one stream
  // live kernel timing (mhd2d_profile_*): event pairs on the context stream
  bool prof;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_ev;
  std::vector<hipEvent_t> ev_pool;
  std::vector<mhd2d_impl::StepGraph> graphs;   // mhd2d_step_graph: one per buffer-role phase
  char err[512];
};

namespace mhd2d_impl {

inline int face_planes(const mhd2d_ctx* ctx) { return ctx->cfg.precision == MHD2D_MIXED ? FACE_PL_MIX : FACE_PL; }

inline int fail(mhd2d_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof ctx->err, fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HIPCHK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, -2, "%s failed: %s", #call, hipGetErrorString(e_));             \
  } while (0)
#define LAUNCHCHK(ctx)                                                                 \
  do {                                                                                 \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess) return fail(ctx, -3, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

inline unsigned nblk(long long n) { return (unsigned)((n + BS - 1) / BS); }


inline hipEvent_t pool_event(mhd2d_ctx* ctx) {
  if (!ctx->ev_pool.empty()) {
    hipEvent_t e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// brackets one launch with events when profiling is on
struct Timed {
  mhd2d_ctx* ctx;
  int kc;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(mhd2d_ctx* c, int k) : ctx(c), kc(k) {
    if (ctx->prof) {
      a = pool_event(ctx);
      b = pool_event(ctx);
      (void)hipEventRecord(a, ctx->stream);
    }
  }
  ~Timed() {
    if (ctx->prof && a && b) {
      (void)hipEventRecord(b, ctx->stream);
      ctx->prof_ev.push_back({kc, {a, b}});
    }
  }
};

// calls f(std::true_type) for the mixed-form state storage, f(std::false_type) otherwise
template <typename F>
int with_mix(const mhd2d_ctx* ctx, F&& f) {
  return ctx->cfg.precision == MHD2D_MIXED ? f(std::true_type{}) : f(std::false_type{});
}

// launch functions (stream-ordered on ctx->stream; 0 or a negative status)
// calls f(integral_constant<int, flux_type>)
template <typename F>
int dispatch_ft(int ft, F&& f) {
  switch (ft) {
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    default: return f(std::integral_constant<int, 4>{});
  }
}

int launch_fluxes(mhd2d_ctx* ctx, const double* U, bool with_vertex, const DevScal* sc);     // mhd2d_faces.hip
int launch_xface_mixed(mhd2d_ctx* ctx, const double* U, int ft);                            // mhd2d_faces_mix.hip
int launch_yface_mixed(mhd2d_ctx* ctx, const double* U, int ft);                            // mhd2d_faces_mix.hip
int launch_vertex(mhd2d_ctx* ctx, const double* U, const DevScal* sc);                      // mhd2d_faces.hip
// lout != nullptr: the epilogue also writes the stage value limited without
// TVB (positivity when pp, divergence-free reconstruction from faces Fn)
int launch_cell(mhd2d_ctx* ctx, int mode, const double* S, const double* A, const double* U0, double* out,
                const DevScal* sc, double dt, double* lout = nullptr, const double* Fn = nullptr,
                bool pp = false);                                                             // mhd2d_cell.hip
int launch_ct(mhd2d_ctx* ctx, int mode, const double* F, const double* A, const double* F0, double* out,
              const DevScal* sc, double dt);                                                  // mhd2d_cell.hip
int launch_axpy(mhd2d_ctx* ctx, double* out, const double* X, double a, const double* Y, double b,
                const double* Z, double dt, long long n);                                     // mhd2d_cell.hip
int launch_axpy_state(mhd2d_ctx* ctx, double* out, const double* X, double a, const double* Y, double b,
                      const double* Z, double dt);                                            // mhd2d_cell.hip
int launch_kxrcf(mhd2d_ctx* ctx, const double* U, DevScal* sc);                               // mhd2d_limit.hip
int launch_tvb_list(mhd2d_ctx* ctx, const double* in, const double* inB = nullptr);                                       // mhd2d_limit.hip
int launch_detect_tvb(mhd2d_ctx* ctx, const double* in, DevScal* sc = nullptr, const double* inB = nullptr);              // mhd2d_limit.hip
int launch_tvb_all(mhd2d_ctx* ctx, const double* in, int* kx);                               // mhd2d_limit.hip
int launch_mlp_screen(mhd2d_ctx* ctx, const double* in, const double* inB = nullptr);   // indicator: prob[], screen, candidate list   // mhd2d_mlp.hip
int launch_candidates_tvb(mhd2d_ctx* ctx, const double* in, DevScal* sc, const double* inB = nullptr);  // KXRCF + TVB on the candidates  // mhd2d_mlp.hip
// stage: 1..3 inside mhd2d_step (the learned indicator screens once, at
// stage 1, and reuses the screen for stages 2 and 3); 0 for a standalone call
int launch_limit(mhd2d_ctx* ctx, const double* in, double* out, const double* F, bool tvb, bool pp, bool div,
                 DevScal* sc, int stage = 0);                                                                // mhd2d_limit.hip
// after a limiter pass without TVB into out (the fused cell epilogue, or
// k_limit<false, ..>): detection and TVB on the unlimited stage value `in`
// (KXRCF-gated and learned indicators; inB: a second buffer holding its Bx,
// By planes), then the cells TVB changed limited again from `in` into out
// (lean: in place, out holding the stage value's other planes and `in` its
// Bx, By planes, the fused stages of both forms).
int launch_detection(mhd2d_ctx* ctx, const double* in, DevScal* sc, int stage, const double* inB = nullptr);                    // mhd2d_limit.hip
int launch_limit_relist(mhd2d_ctx* ctx, double* in, double* out, const double* F, bool pp, int stage,
                        bool lean = false);  // mhd2d_limit.hip
// the cells the fused cell kernel's epilogue left to the positivity limiter
// (the list of its bound test), limited from `in` into out
int launch_pp_list(mhd2d_ctx* ctx, double* q, double* out, const double* F, int stage);                 // mhd2d_limit.hip
// div_free_Balsara into the B planes of out from faces F (halo B zeroed);
// the other planes are left as they are
int launch_divfree(mhd2d_ctx* ctx, double* out, const double* F);                                  // mhd2d_limit.hip

}  // namespace mhd2d_impl
