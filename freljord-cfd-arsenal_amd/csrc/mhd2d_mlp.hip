// mhd2d_mlp.hip -- the learned troubled-cell indicator: a 60-64-64-1 MLP
// (GELU, sigmoid output) over a cell's and its four neighbours' normalised
// modal coefficients.
//
//   screen     k_mlp_screen: BF16 MFMA (v_mfma_f32_16x16x16_bf16) per wave of
//              64 cells, FP32 accumulation, p = P(the plain method limits the
//              cell).  Cells with p >= mlp_lo, dilated by one cell, are the
//              candidates; the rest are confidently smooth and skip the
//              classical detector.  Inside mhd2d_step the screen runs once per
//              step (on the first stage's limiter input) and serves all three
//              stages: a feature moves by at most a fifth of a cell per step at
//              the CFL numbers used, and the dilation covers it.
//   detector   KXRCF_Detector on the candidates only, at every stage; its
//              flags feed the same sparse TVB and apply passes as the plain
//              method, so on candidate cells the decision is the plain one.
//   training   FP32 kernels: dense forward, output + weighted cross-entropy,
//              backward through GELU, weight gradients, Adam.
//   labels     the plain method's own decision on the same state: KXRCF
//              flags the cell AND the TVB limiter changes it.
//
// Features of cell c (60 = 5 stencil cells x 12): stencil k = C, E, W, N, S.
// With s_r, s_e the largest |mean| of density and total energy over the
// stencil and s_m the largest |mean| of rho u, rho v:
//   [0]      k = C: log10(M hx hy / s_r) / 4 (the TVB threshold on the local
//            scale); k != C: (rho_k - rho_C) / s_r
//   [1..4]   density modes 2, 3, 4, 6 (x, y, P2(x), P2(y)) / s_r
//   [5]      k = C: log10(sqrt(hx hy)) / 2 (the resolution, which sets the
//            KXRCF threshold); k != C: (E_k - E_C) / s_e
//   [6..9]   energy modes 2, 3, 4, 6 / s_e
//   [10, 11] mean rho u / s_m, mean rho v / s_m (the inflow directions)
#include "mhd2d_detect.hpp"
#include "mhd2d_internal.hpp"

static_assert(mhd2d_impl::MLP_NPARAM == MHD2D_MLP_NPARAM && mhd2d_impl::MLP_NIN == MHD2D_MLP_NIN,
              "indicator MLP layout");

namespace mhd2d_impl {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int XP = 68;          // LDS row pitch (bf16) of a 64-wide activation row: conflict-free A reads

template <class SV>
__device__ __forceinline__ void mlp_features(const SV& U, const Geo& g, int i, int j,
                                             double Mh2, float f[MLP_NIN]) {
  const long long cs[5] = {cix(g, i, j), cix(g, i + 1, j), cix(g, i - 1, j), cix(g, i, j + 1), cix(g, i, j - 1)};
  constexpr int dm[4] = {1, 2, 3, 5};
  double r[5], e[5], mu[5], mv[5];
  double sr = 0.0, se = 0.0, sm = 0.0;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    r[k] = U(PL(0, 0), cs[k]);
    e[k] = U(PL(0, 4), cs[k]);
    mu[k] = U(PL(0, 1), cs[k]);
    mv[k] = U(PL(0, 2), cs[k]);
    sr = fmax(sr, fabs(r[k]));
    se = fmax(se, fabs(e[k]));
    sm = fmax(sm, fmax(fabs(mu[k]), fabs(mv[k])));
  }
  const double ir = 1.0 / (sr + 1e-12), ie = 1.0 / (se + 1e-12), im = 1.0 / (sm + 1e-12);
#pragma unroll
  for (int k = 0; k < 5; k++) {
    float* o = f + 12 * k;
    o[0] = k == 0 ? (float)(log10(Mh2 * ir) * 0.25) : (float)((r[k] - r[0]) * ir);
    o[5] = k == 0 ? (float)(log10(sqrt(g.hx * g.hy)) * 0.5) : (float)((e[k] - e[0]) * ie);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      o[1 + q] = (float)(U(PL(dm[q], 0), cs[k]) * ir);
      o[6 + q] = (float)(U(PL(dm[q], 4), cs[k]) * ie);
    }
    o[10] = (float)(mu[k] * im);
    o[11] = (float)(mv[k] * im);
  }
}

// append the band lanes' cell ids to blist (one atomic per wave)
__device__ __forceinline__ void band_append(bool band, int t, int* __restrict__ blist, DevScal* lst) {
  const unsigned long long b = __ballot(band);
  if (b) {
    const int lane = __lane_id(), leader = __ffsll((long long)b) - 1;
    unsigned int base = 0u;
    if (lane == leader) base = atomicAdd(&lst->nband, (unsigned int)__popcll(b));
    base = (unsigned int)__shfl((int)base, leader);
    if (band) blist[base + (unsigned int)__popcll(b & ((1ull << lane) - 1ull))] = t;
  }
}

// the same features in FP32 for inference (l10Mh2 = log10(M hx hy),
// l10h = log10(sqrt(hx hy)) / 2 precomputed)
template <class SV>
__device__ __forceinline__ void mlp_features_f32(const SV& U, const Geo& g, int i, int j,
                                                 float l10Mh2, float l10h, float f[MLP_NIN]) {
  const long long cs[5] = {cix(g, i, j), cix(g, i + 1, j), cix(g, i - 1, j), cix(g, i, j + 1), cix(g, i, j - 1)};
  constexpr int dm[4] = {1, 2, 3, 5};
  float r[5], e[5], mu[5], mv[5];
  float sr = 0.f, se = 0.f, sm = 0.f;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    r[k] = (float)U(PL(0, 0), cs[k]);
    e[k] = (float)U(PL(0, 4), cs[k]);
    mu[k] = (float)U(PL(0, 1), cs[k]);
    mv[k] = (float)U(PL(0, 2), cs[k]);
    sr = fmaxf(sr, fabsf(r[k]));
    se = fmaxf(se, fabsf(e[k]));
    sm = fmaxf(sm, fmaxf(fabsf(mu[k]), fabsf(mv[k])));
  }
  const float ir = __builtin_amdgcn_rcpf(sr + 1e-12f), ie = __builtin_amdgcn_rcpf(se + 1e-12f);
  const float im = __builtin_amdgcn_rcpf(sm + 1e-12f);
#pragma unroll
  for (int k = 0; k < 5; k++) {
    float* o = f + 12 * k;
    o[0] = k == 0 ? 0.25f * (l10Mh2 - __log10f(sr + 1e-12f)) : (r[k] - r[0]) * ir;
    o[5] = k == 0 ? l10h : (e[k] - e[0]) * ie;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      o[1 + q] = (float)U(PL(dm[q], 0), cs[k]) * ir;
      o[6 + q] = (float)U(PL(dm[q], 4), cs[k]) * ie;
    }
    o[10] = mu[k] * im;
    o[11] = mv[k] * im;
  }
}

__device__ __forceinline__ unsigned short bf16_bits(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
__device__ __forceinline__ unsigned int bf16x2_bits(float a, float b) {
  typedef __bf16 b2 __attribute__((ext_vector_type(2)));
  const b2 v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(unsigned int, v);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
// inference GELU: the tanh form x / (1 + exp(-2 sqrt(2/pi) (x + 0.044715 x^3)));
// it differs from the erf form by < 5e-4, below the BF16 rounding of the
// activations it feeds
__device__ __forceinline__ float gelu_fast(float x) {
  const float u = x * (1.5957691216057308f + 0.0713548162726f * x * x);
  return x * __builtin_amdgcn_rcpf(1.0f + __expf(-u));
}
__device__ __forceinline__ float gelu_d(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__device__ __forceinline__ unsigned short f2bf(float x) {    // round to nearest even
  unsigned int u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// Inference weights (prepared by the host from the FP32 parameters): per
// layer 16 B-operand tiles (kk, nt) of 64 lanes x 4 bf16, lane l holding
// W[kk*16 + 4*(l/16) + j][nt*16 + l%16], j = 0..3; then b1, b2, w3, b3.
struct MlpDev {
  s4 w1[16][64];
  s4 w2[16][64];
  float b1[64], b2[64], w3[64], b3;
};

// the screen's LDS tiles are per wave (sX[wv], sZ[wv]): a wave orders its
// own LDS writes before its reads, without waiting for the block's other waves
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool MIX>
__global__ __launch_bounds__(256) void k_mlp_screen(SRdB<MIX> U, int* __restrict__ screen,
                                                    float* __restrict__ prob, const MlpDev* __restrict__ W,
                                                    Geo g, float l10Mh2, float l10h, float lo, float hi) {
  if (step_off(g)) return;
  __shared__ __attribute__((aligned(16))) unsigned short sX[4][64 * XP];
  __shared__ float sZ[4][64];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, grp = l >> 4, col = l & 15;
  const long long ncell = (long long)g.Nx * g.Ny;
  const long long t = xcd_block() * 256 + threadIdx.x;
  const bool valid = t < ncell;
  const int i = valid ? 1 + (int)(t % g.Nx) : 1, j = valid ? 1 + (int)(t / g.Nx) : 1;
  unsigned short* X = sX[wv];
  {
    float f[MLP_NIN];
    if (valid) mlp_features_f32(U, g, i, j, l10Mh2, l10h, f);
    unsigned int* X32 = reinterpret_cast<unsigned int*>(X + l * XP);
#pragma unroll
    for (int k = 0; k < MLP_NIN / 2; k++) X32[k] = valid ? bf16x2_bits(f[2 * k], f[2 * k + 1]) : 0u;
#pragma unroll
    for (int k = MLP_NIN / 2; k < 32; k++) X32[k] = 0u;
  }
  wave_lds_sync();
  // layer 1: [64 cells x 64] x [64 x 64]
  f4 acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
    s4 b[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) b[nt] = W->w1[kk * 4 + nt][l];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const s4 a = *reinterpret_cast<const s4*>(&X[(mt * 16 + col) * XP + kk * 16 + 4 * grp]);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[nt], acc[mt][nt], 0, 0, 0);
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int q = 0; q < 4; q++)
        X[(mt * 16 + 4 * grp + q) * XP + nt * 16 + col] = bf16_bits(gelu_fast(acc[mt][nt][q] + W->b1[nt * 16 + col]));
  wave_lds_sync();
  // layer 2
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
    s4 b[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) b[nt] = W->w2[kk * 4 + nt][l];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const s4 a = *reinterpret_cast<const s4*>(&X[(mt * 16 + col) * XP + kk * 16 + 4 * grp]);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[nt], acc[mt][nt], 0, 0, 0);
    }
  }
  // output layer: z(row) = b3 + sum_n w3[n] gelu(h2(row, n) + b2[n]); the
  // 16 lanes of a group hold the 16 columns of a tile row
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      float s = 0.f;
#pragma unroll
      for (int nt = 0; nt < 4; nt++) s += W->w3[nt * 16 + col] * gelu_fast(acc[mt][nt][q] + W->b2[nt * 16 + col]);
      s += __shfl_xor(s, 1);
      s += __shfl_xor(s, 2);
      s += __shfl_xor(s, 4);
      s += __shfl_xor(s, 8);
      if (col == 0) sZ[wv][mt * 16 + 4 * grp + q] = s + W->b3;
    }
  wave_lds_sync();
  if (!valid) return;
  const float p = __builtin_amdgcn_rcpf(1.0f + __expf(-sZ[wv][l]));
  prob[t] = p;
  // 2: the network decides the cell is troubled (TVB without KXRCF);
  // 1: the confidence band, decided by KXRCF; 0: not troubled
  screen[t] = p >= hi ? 2 : p >= lo ? 1 : 0;
}

// the screen mark (0, 1 or 2) of interior cell (a, b) (0-based); outside
// the domain the periodic image, or 0 for outflow boundaries
__device__ __forceinline__ int screen_at(const int* __restrict__ screen, const Geo& g, int a, int b) {
  if (a < 0) { if (g.bcL != 1) return 0; a += g.Nx; }
  if (a >= g.Nx) { if (g.bcR != 1) return 0; a -= g.Nx; }
  if (b < 0) { if (g.bcD != 1) return 0; b += g.Ny; }
  if (b >= g.Ny) { if (g.bcU != 1) return 0; b -= g.Ny; }
  return screen[(long long)b * g.Nx + a];
}

__device__ __forceinline__ bool candidate(const int* __restrict__ screen, const Geo& g, int i, int j) {
  return screen_at(screen, g, i, j) | screen_at(screen, g, i + 1, j) | screen_at(screen, g, i - 1, j) |
         screen_at(screen, g, i, j + 1) | screen_at(screen, g, i, j - 1);
}

// candidates = screen dilated by one cell; every cell's TVB marks are
// cleared for the step
__global__ __launch_bounds__(BS) void k_dilate(const int* __restrict__ screen, int* __restrict__ trouble,
                                               int* __restrict__ tmask, int* __restrict__ blist, Geo g,
                                               DevScal* lst) {
  if (step_off(g)) return;
  const long long t = xcd_block() * BS + threadIdx.x;
  if (t >= (long long)g.Nx * g.Ny) return;
  const bool cand = candidate(screen, g, (int)(t % g.Nx), (int)(t / g.Nx));
  trouble[t] = 0;
  tmask[t] = 0;
  band_append(cand, (int)t, blist, lst);
}

__global__ void k_reset_nlist(DevScal* lst, const int* gate) {
  if (step_off(gate)) return;
  if (threadIdx.x == 0) lst->nlist = 0u;
}

// KXRCF_Detector on the candidates; flagged ones join the troubled-cell
// list, and so do the cells the network itself decides (screen mark 2)
// without the KXRCF evaluation
template <bool MIX>
__global__ __launch_bounds__(BS) void k_kxrcf_cands(SRdB<MIX> U, int* __restrict__ trouble,
                                                    int* __restrict__ tmask, int* __restrict__ list,
                                                    const int* __restrict__ blist, const int* __restrict__ screen,
                                                    Geo g, double Ck, double hk, DevScal* sc, DevScal* lst) {
  if (step_off(g)) return;
  const unsigned int nb = lst->nband;
  for (unsigned int k0 = blockIdx.x * BS; k0 < nb; k0 += gridDim.x * BS) {
    const unsigned int k = k0 + threadIdx.x;
    bool flag = false;
    int t = 0;
    if (k < nb) {
      t = blist[k];
      const int i = 1 + t % g.Nx, j = 1 + t / g.Nx;
      flag = screen[t] == 2 ? true : kxrcf_cell(U, g, i, j, Ck, hk);
      trouble[t] = flag ? 1 : 0;
      tmask[t] = 0;
    }
    if (sc) count_flags(flag, sc);
    list_append(flag, t, list, lst);
  }
}

__global__ void k_reset_nband(DevScal* lst, const int* gate) {
  if (step_off(gate)) return;
  if (threadIdx.x == 0) lst->nband = 0u;
}

// ---------------------------------------------------------------------------
// training data: features and label of every interior cell whose label is 1
// or whose hash falls below neg_keep; label = KXRCF flag AND TVB change
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

template <bool MIX>
__global__ __launch_bounds__(BS) void k_collect(SRd<MIX> U, const int* __restrict__ trouble,
                                                const int* __restrict__ tmask, Geo g, double Mh2, double neg_keep,
                                                unsigned long long seed, float* __restrict__ X,
                                                float* __restrict__ Y, long long* __restrict__ idx,
                                                unsigned long long* __restrict__ count,
                                                unsigned long long* __restrict__ npos, long long cap) {
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  if (t >= (long long)g.Nx * g.Ny) return;
  const bool label = trouble[t] != 0 && tmask[t] != 0;
  const double u = (double)(mix64(seed ^ (0x9e3779b97f4a7c15ull * (unsigned long long)(t + 1))) >> 11) * 0x1.0p-53;
  if (!label && !(u < neg_keep)) return;
  const int i = 1 + (int)(t % g.Nx), j = 1 + (int)(t / g.Nx);
  float f[MLP_NIN];
  mlp_features(U, g, i, j, Mh2, f);
  bool finite = true;
#pragma unroll
  for (int q = 0; q < MLP_NIN; q++) finite = finite && isfinite(f[q]);
  if (!finite) return;           // states that lost positivity carry no label
  const unsigned long long k = atomicAdd(count, 1ull);
  if ((long long)k >= cap) return;
  if (label) atomicAdd(npos, 1ull);
#pragma unroll
  for (int q = 0; q < MLP_NIN; q++) X[k * MLP_NIN + q] = f[q];
  Y[k] = label ? 1.f : 0.f;
  idx[k] = t;
}

// ---------------------------------------------------------------------------
// FP32 training kernels over a batch of B samples (rows gathered by rows[])
// Z = In W + b, A = gelu(Z); In rows are gathered through rows[] when given
__global__ __launch_bounds__(BS) void k_dense_fwd(const float* __restrict__ In, const long long* __restrict__ rows,
                                                  int K, const float* __restrict__ Wt, const float* __restrict__ b,
                                                  float* __restrict__ Z, float* __restrict__ A, int B) {
  const long long e = (long long)blockIdx.x * BS + threadIdx.x;
  if (e >= (long long)B * 64) return;
  const int r = (int)(e / 64), n = (int)(e % 64);
  const float* x = In + (rows ? rows[r] : (long long)r) * K;
  float z = b[n];
  for (int k = 0; k < K; k++) z += x[k] * Wt[k * 64 + n];
  Z[e] = z;
  A[e] = gelu(z);
}

// z3 = A2 w3 + b3, p = sigmoid(z3), per-sample weighted cross-entropy
// loss_b = w_b (-(y log p + (1 - y) log(1 - p))) / B and dz3 = w_b (p - y) / B
__global__ __launch_bounds__(BS) void k_out_loss(const float* __restrict__ A2, const float* __restrict__ w3,
                                                 const float* __restrict__ b3, const float* __restrict__ Y,
                                                 const long long* __restrict__ rows, float pos_w, int B,
                                                 float* __restrict__ dz3, float* __restrict__ loss) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= B) return;
  float z = b3[0];
  for (int n = 0; n < 64; n++) z += A2[(long long)r * 64 + n] * w3[n];
  const float y = Y[rows ? rows[r] : r];
  const float w = y > 0.5f ? pos_w : 1.0f;
  const float p = 1.0f / (1.0f + expf(-z));
  // log(1 + e^-|z|) form for both terms
  const float lse = log1pf(expf(-fabsf(z)));
  const float lp = -(fmaxf(-z, 0.f) + lse), l1p = -(fmaxf(z, 0.f) + lse);   // log p, log(1 - p)
  loss[r] = -w * (y * lp + (1.f - y) * l1p) / (float)B;
  dz3[r] = w * (p - y) / (float)B;
}

// dZ = (dNext Wn^T) .* gelu'(Z); dNext is [B x Nn], Wn is [64 x Nn]
__global__ __launch_bounds__(BS) void k_dense_bwd(const float* __restrict__ dNext, int Nn, const float* __restrict__ Wn,
                                                  const float* __restrict__ Z, float* __restrict__ dZ, int B) {
  const long long e = (long long)blockIdx.x * BS + threadIdx.x;
  if (e >= (long long)B * 64) return;
  const int r = (int)(e / 64), k = (int)(e % 64);
  float s = 0.f;
  for (int n = 0; n < Nn; n++) s += dNext[(long long)r * Nn + n] * Wn[k * Nn + n];
  dZ[e] = s * gelu_d(Z[e]);
}

// dW[k][n] = sum_r In[r][k] dZ[r][n] (k < K), db[n] = sum_r dZ[r][n], split
// over the batch: block (x, c) takes WG_ROWS rows (chunk c) for BS entries
// of the (K + 1) x N table (the bias row is k = K) and writes its partial
// sums; k_wgrad_sum adds the chunks in a fixed order.  A batch of 8192 rows
// then runs as 64 x 17 blocks instead of 17 blocks that each walk all rows.
constexpr int WG_ROWS = 128;
__global__ __launch_bounds__(BS) void k_wgrad_part(const float* __restrict__ In, const long long* __restrict__ rows,
                                                   int K, const float* __restrict__ dZ, int N,
                                                   float* __restrict__ part, int B) {
  const int ne = (K + 1) * N;
  const int e = blockIdx.x * BS + threadIdx.x;
  if (e >= ne) return;
  const int k = e / N, n = e % N;
  const int r0 = blockIdx.y * WG_ROWS, r1 = min(B, r0 + WG_ROWS);
  float s = 0.f;
  if (k < K) {
    for (int r = r0; r < r1; r++) s += In[(rows ? rows[r] : (long long)r) * K + k] * dZ[(long long)r * N + n];
  } else {
    for (int r = r0; r < r1; r++) s += dZ[(long long)r * N + n];
  }
  part[(long long)blockIdx.y * ne + e] = s;
}
__global__ __launch_bounds__(BS) void k_wgrad_sum(const float* __restrict__ part, int nchunk, int K, int N,
                                                  float* __restrict__ dW, float* __restrict__ db) {
  const int ne = (K + 1) * N;
  const int e = blockIdx.x * BS + threadIdx.x;
  if (e >= ne) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; c++) s += part[(long long)c * ne + e];
  if (e < K * N) dW[e] = s;
  else db[e - K * N] = s;
}

// fixed-order sum of a vector (one block)
__global__ void k_sum(const float* __restrict__ v, int n, float* __restrict__ out) {
  __shared__ float s[BS];
  float a = 0.f;
  for (int k = threadIdx.x; k < n; k += BS) a += v[k];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = BS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

__global__ __launch_bounds__(BS) void k_adam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                                             float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                                             float c1, float c2) {
  const int e = blockIdx.x * BS + threadIdx.x;
  if (e >= n) return;
  const float gg = gr[e];
  const float mm = b1 * m[e] + (1.f - b1) * gg;
  const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// random rows of an epoch (with replacement)
__global__ __launch_bounds__(BS) void k_rows(long long* __restrict__ rows, int B, long long N, unsigned long long seed) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= B) return;
  rows[r] = (long long)(mix64(seed + 0x9e3779b97f4a7c15ull * (unsigned long long)(r + 1)) % (unsigned long long)N);
}

// indicator statistics over the interior cells (see mhd2d_indicator_stats);
// the decision on a candidate is the network's where it marks the cell 2,
// KXRCF's elsewhere
__global__ __launch_bounds__(BS) void k_istats(const int* __restrict__ screen, const int* __restrict__ kx,
                                               const int* __restrict__ tm, Geo g,
                                               unsigned long long* __restrict__ out) {
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  if (t >= (long long)g.Nx * g.Ny) return;
  const bool cand = candidate(screen, g, (int)(t % g.Nx), (int)(t / g.Nx));
  const bool k = kx[t] != 0, c = tm[t] != 0, lab = k && c;
  const bool d = cand && (k || screen[t] == 2);
  const bool hydro = (tm[t] & ~((1 << 5) | (1 << 6))) != 0;   // change bits of rho .. E, Bz
  const bool v[8] = {true, lab, k, d, cand, lab && !cand, cand && !lab, lab && !cand && hydro};
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const unsigned long long b = __ballot(v[q]);
    if (__lane_id() == __ffsll((long long)__ballot(1)) - 1 && b) atomicAdd(&out[q], (unsigned long long)__popcll(b));
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side

static inline unsigned short host_f2bf(float x) {
  unsigned int u;
  memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

int mlp_upload(mhd2d_ctx* ctx, const float* params) {
  MlpDev h;
  memset(&h, 0, sizeof h);
  const float* W1 = params + MLP_OFF_W1;
  const float* W2 = params + MLP_OFF_W2;
  for (int kk = 0; kk < 4; kk++)
    for (int nt = 0; nt < 4; nt++)
      for (int l = 0; l < 64; l++)
        for (int q = 0; q < 4; q++) {
          const int k = kk * 16 + 4 * (l / 16) + q, n = nt * 16 + (l % 16);
          h.w1[kk * 4 + nt][l][q] = (short)(k < MLP_NIN ? host_f2bf(W1[k * 64 + n]) : 0);
          h.w2[kk * 4 + nt][l][q] = (short)host_f2bf(W2[k * 64 + n]);
        }
  memcpy(h.b1, params + MLP_OFF_B1, 64 * sizeof(float));
  memcpy(h.b2, params + MLP_OFF_B2, 64 * sizeof(float));
  memcpy(h.w3, params + MLP_OFF_W3, 64 * sizeof(float));
  h.b3 = params[MLP_OFF_B3];
  if (!ctx->mlpw) HIPCHK(ctx, hipMalloc(&ctx->mlpw, sizeof(MlpDev)));
  HIPCHK(ctx, hipMemcpyAsync(ctx->mlpw, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

static inline unsigned nb_blocks(long long n) { return (unsigned)((n + BS - 1) / BS); }

static int launch_screen_only(mhd2d_ctx* ctx, const double* in, const double* inB = nullptr) {
  const Geo& g = ctx->g;
  if (!ctx->mlpw) return fail(ctx, -6, "indicator MLP: no weights (mhd2d_mlp_set_weights)");
  const long long ncell = (long long)g.Nx * g.Ny;
  return with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_mlp_screen<MIX>), dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, ctx->stream,
                       SRdB<MIX>{in, g.plane, inB}, ctx->screen, ctx->prob, (const MlpDev*)ctx->mlpw, g,
                       (float)log10(ctx->cfg.M * g.hx * g.hy), (float)(0.5 * log10(sqrt(g.hx * g.hy))),
                       (float)ctx->cfg.mlp_lo, (float)ctx->cfg.mlp_hi);
    LAUNCHCHK(ctx);
    return 0;
  });
}

int launch_mlp_screen(mhd2d_ctx* ctx, const double* in, const double* inB) {
  const Geo& g = ctx->g;
  int rc = launch_screen_only(ctx, in, inB);
  if (rc) return rc;
  hipLaunchKernelGGL(k_reset_nband, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, ctx->g.gate);
  hipLaunchKernelGGL(k_dilate, dim3(nb_blocks((long long)g.Nx * g.Ny)), dim3(BS), 0, ctx->stream,
                     (const int*)ctx->screen, ctx->trouble, ctx->tmask, ctx->blist, g, ctx->sc);
  LAUNCHCHK(ctx);
  return 0;
}

int launch_candidates_tvb(mhd2d_ctx* ctx, const double* in, DevScal* sc, const double* inB) {
  const Geo& g = ctx->g;
  const long long ncell = (long long)g.Nx * g.Ny;
  const unsigned nb = std::min<unsigned>(nb_blocks(ncell), 1024u);
  hipLaunchKernelGGL(k_reset_nlist, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, ctx->g.gate);
  int rc = with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_kxrcf_cands<MIX>), dim3(nb), dim3(BS), 0, ctx->stream, SRdB<MIX>{in, g.plane, inB}, ctx->trouble,
                       ctx->tmask, ctx->tlist, (const int*)ctx->blist, (const int*)ctx->screen, g, ctx->cfg.Ck,
                       kx_hk(g), sc, ctx->sc);
    LAUNCHCHK(ctx);
    return 0;
  });
  if (rc) return rc;
  return launch_tvb_list(ctx, in, inB);
}

}  // namespace mhd2d_impl

// ---------------------------------------------------------------------------
// trainer: a device dataset of (features, label) rows and the FP32 training
// loop (no counterpart in the reference)
struct mhd2d_trainer {
  int device;
  hipStream_t stream;
  long long cap, n, npos;
  int bmax;
  float *X, *Y;
  long long* idx;
  unsigned long long* cnt;     // [0] rows, [1] positives (device)
  long long* rows;
  float *Z1, *A1, *Z2, *A2, *dz3, *loss, *dZ2, *dZ1, *lsum;
  float *P, *G, *Mo, *Vo;
  float* wpart;                // weight-gradient partial sums, ceil(bmax / WG_ROWS) x 65 x 64
  char err[512];
};

namespace {
using namespace mhd2d_impl;

int tfail(mhd2d_trainer* tr, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(tr->err, sizeof tr->err, fmt, ap);
  va_end(ap);
  return code;
}
#define TCHK(tr, call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return tfail(tr, -2, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

inline unsigned nb_of(long long n) { return (unsigned)((n + BS - 1) / BS); }

void wgrad(mhd2d_trainer* tr, const float* In, const long long* rows, int K, const float* dZ, int N, float* dW,
           float* db, int B) {
  const int nchunk = (B + WG_ROWS - 1) / WG_ROWS;
  hipLaunchKernelGGL(k_wgrad_part, dim3(nb_of((long long)(K + 1) * N), nchunk), dim3(BS), 0, tr->stream, In, rows, K,
                     dZ, N, tr->wpart, B);
  hipLaunchKernelGGL(k_wgrad_sum, dim3(nb_of((long long)(K + 1) * N)), dim3(BS), 0, tr->stream,
                     (const float*)tr->wpart, nchunk, K, N, dW, db);
}

// forward + backward of one batch (rows of X given by rows[], or the B rows
// from X itself when rows == nullptr); gradients into tr->G, loss into lsum
int batch_grad(mhd2d_trainer* tr, const float* X, const float* Y, const long long* rows, int B, float pos_w) {
  hipStream_t st = tr->stream;
  float* P = tr->P;
  float* G = tr->G;
  hipLaunchKernelGGL(k_dense_fwd, dim3(nb_of((long long)B * 64)), dim3(BS), 0, st, X, rows, MLP_NIN,
                     P + MLP_OFF_W1, P + MLP_OFF_B1, tr->Z1, tr->A1, B);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nb_of((long long)B * 64)), dim3(BS), 0, st, (const float*)tr->A1,
                     (const long long*)nullptr, 64, P + MLP_OFF_W2, P + MLP_OFF_B2, tr->Z2, tr->A2, B);
  hipLaunchKernelGGL(k_out_loss, dim3(nb_of(B)), dim3(BS), 0, st, (const float*)tr->A2, P + MLP_OFF_W3,
                     P + MLP_OFF_B3, Y, rows, pos_w, B, tr->dz3, tr->loss);
  wgrad(tr, tr->A2, nullptr, 64, tr->dz3, 1, G + MLP_OFF_W3, G + MLP_OFF_B3, B);
  hipLaunchKernelGGL(k_dense_bwd, dim3(nb_of((long long)B * 64)), dim3(BS), 0, st, (const float*)tr->dz3, 1,
                     P + MLP_OFF_W3, (const float*)tr->Z2, tr->dZ2, B);
  wgrad(tr, tr->A1, nullptr, 64, tr->dZ2, 64, G + MLP_OFF_W2, G + MLP_OFF_B2, B);
  hipLaunchKernelGGL(k_dense_bwd, dim3(nb_of((long long)B * 64)), dim3(BS), 0, st, (const float*)tr->dZ2, 64,
                     P + MLP_OFF_W2, (const float*)tr->Z1, tr->dZ1, B);
  wgrad(tr, X, rows, MLP_NIN, tr->dZ1, 64, G + MLP_OFF_W1, G + MLP_OFF_B1, B);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(BS), 0, st, (const float*)tr->loss, B, tr->lsum);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tfail(tr, -3, "kernel launch: %s", hipGetErrorString(e));
  return 0;
}

}  // namespace

extern "C" {

int mhd2d_mlp_set_weights(mhd2d_ctx* ctx, const float* params, int32_t n) {
  if (!ctx || !params) return -1;
  if (n != MLP_NPARAM) return fail(ctx, -1, "mlp_set_weights: expected %d parameters, got %d", MLP_NPARAM, n);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return mlp_upload(ctx, params);
}

int mhd2d_mlp_probability(mhd2d_ctx* ctx, float* host_p) {
  if (!ctx) return -1;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc = launch_screen_only(ctx, ctx->S);
  if (rc) return rc;
  const long long ncell = (long long)ctx->g.Nx * ctx->g.Ny;
  if (host_p) HIPCHK(ctx, hipMemcpyAsync(host_p, ctx->prob, ncell * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int mhd2d_indicator_stats(mhd2d_ctx* ctx, int64_t* out) {
  if (!ctx || !out) return -1;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const long long ncell = (long long)ctx->g.Nx * ctx->g.Ny;
  int* kx = nullptr;
  unsigned long long* cnt = nullptr;
  HIPCHK(ctx, hipMalloc(&kx, ncell * sizeof(int)));
  HIPCHK(ctx, hipMalloc(&cnt, 8 * sizeof(unsigned long long)));
  int rc = launch_screen_only(ctx, ctx->S);
  if (!rc) rc = launch_tvb_all(ctx, ctx->S, kx);   // KXRCF flags into kx, TVB change bits of every cell into tmask
  if (!rc) {
    (void)hipMemsetAsync(cnt, 0, 8 * sizeof(unsigned long long), ctx->stream);
    hipLaunchKernelGGL(k_istats, dim3(nb_of(ncell)), dim3(BS), 0, ctx->stream, (const int*)ctx->screen,
                       (const int*)kx, (const int*)ctx->tmask, ctx->g, cnt);
    unsigned long long h[8];
    (void)hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = fail(ctx, -2, "indicator_stats: %s", hipGetErrorString(e));
    for (int q = 0; q < 8; q++) out[q] = (int64_t)h[q];
  }
  (void)hipFree(kx);
  (void)hipFree(cnt);
  return rc;
}

const char* mhd2d_trainer_last_error(const mhd2d_trainer* tr) { return tr ? tr->err : "null trainer"; }

int mhd2d_trainer_destroy(mhd2d_trainer* tr) {
  if (!tr) return 0;
  (void)hipSetDevice(tr->device);
  if (tr->stream) (void)hipStreamSynchronize(tr->stream);
  void* ps[] = {tr->X, tr->Y, tr->idx, tr->cnt, tr->rows, tr->Z1, tr->A1, tr->Z2, tr->A2, tr->dz3, tr->loss,
                tr->dZ2, tr->dZ1, tr->lsum, tr->P, tr->G, tr->Mo, tr->Vo, tr->wpart};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (tr->stream) (void)hipStreamDestroy(tr->stream);
  free(tr);
  return 0;
}

int mhd2d_trainer_create(int64_t capacity, int32_t max_batch, mhd2d_trainer** out) {
  if (!out || capacity <= 0 || max_batch <= 0) return -1;
  *out = nullptr;
  mhd2d_trainer* tr = (mhd2d_trainer*)calloc(1, sizeof(mhd2d_trainer));
  if (!tr) return -1;
  tr->cap = capacity;
  tr->bmax = max_batch;
  if (hipGetDevice(&tr->device) != hipSuccess) { free(tr); return -2; }
  const size_t B = (size_t)max_batch;
  struct { void** p; size_t n; } al[] = {
      {(void**)&tr->X, (size_t)capacity * MLP_NIN * sizeof(float)}, {(void**)&tr->Y, (size_t)capacity * sizeof(float)},
      {(void**)&tr->idx, (size_t)capacity * sizeof(long long)}, {(void**)&tr->cnt, 2 * sizeof(unsigned long long)},
      {(void**)&tr->rows, B * sizeof(long long)}, {(void**)&tr->Z1, B * 64 * sizeof(float)},
      {(void**)&tr->A1, B * 64 * sizeof(float)}, {(void**)&tr->Z2, B * 64 * sizeof(float)},
      {(void**)&tr->A2, B * 64 * sizeof(float)}, {(void**)&tr->dz3, B * sizeof(float)},
      {(void**)&tr->loss, B * sizeof(float)}, {(void**)&tr->dZ2, B * 64 * sizeof(float)},
      {(void**)&tr->dZ1, B * 64 * sizeof(float)}, {(void**)&tr->lsum, sizeof(float)},
      {(void**)&tr->P, MLP_NPARAM * sizeof(float)}, {(void**)&tr->G, MLP_NPARAM * sizeof(float)},
      {(void**)&tr->Mo, MLP_NPARAM * sizeof(float)}, {(void**)&tr->Vo, MLP_NPARAM * sizeof(float)},
      {(void**)&tr->wpart, (B + WG_ROWS - 1) / WG_ROWS * 65 * 64 * sizeof(float)},
  };
  for (auto& a : al) {
    if (hipMalloc(a.p, a.n) != hipSuccess) {
      mhd2d_trainer_destroy(tr);
      return -2;
    }
  }
  if (hipStreamCreateWithFlags(&tr->stream, hipStreamNonBlocking) != hipSuccess) {
    mhd2d_trainer_destroy(tr);
    return -2;
  }
  (void)hipMemsetAsync(tr->cnt, 0, 2 * sizeof(unsigned long long), tr->stream);
  if (hipStreamSynchronize(tr->stream) != hipSuccess) {
    mhd2d_trainer_destroy(tr);
    return -2;
  }
  *out = tr;
  return 0;
}

int mhd2d_trainer_collect(mhd2d_trainer* tr, mhd2d_ctx* ctx, double neg_keep, uint64_t seed, int64_t* n_total,
                          int64_t* n_pos) {
  if (!tr || !ctx) return -1;
  TCHK(tr, hipSetDevice(tr->device));
  const Geo& g = ctx->g;
  // the plain method's decision on the current state: KXRCF flags + TVB change bits
  int rc = launch_detect_tvb(ctx, ctx->S);
  if (rc) return tfail(tr, rc, "collect: %s", ctx->err);
  rc = with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_collect<MIX>), dim3(nb_of((long long)g.Nx * g.Ny)), dim3(BS), 0, ctx->stream,
                       SRd<MIX>{ctx->S, g.plane}, (const int*)ctx->trouble, (const int*)ctx->tmask, g,
                       ctx->cfg.M * g.hx * g.hy, neg_keep, (unsigned long long)seed, tr->X, tr->Y, tr->idx, tr->cnt,
                       tr->cnt + 1, tr->cap);
    return 0;
  });
  unsigned long long h[2];
  TCHK(tr, hipMemcpyAsync(h, tr->cnt, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  TCHK(tr, hipStreamSynchronize(ctx->stream));
  tr->n = std::min<long long>((long long)h[0], tr->cap);
  tr->npos = (long long)h[1];
  if (n_total) *n_total = tr->n;
  if (n_pos) *n_pos = tr->npos;
  return (long long)h[0] > tr->cap ? tfail(tr, 1, "collect: capacity %lld reached", tr->cap) : 0;
}

int mhd2d_trainer_size(mhd2d_trainer* tr, int64_t* n, int64_t* npos) {
  if (!tr) return -1;
  if (n) *n = tr->n;
  if (npos) *npos = tr->npos;
  return 0;
}

int mhd2d_trainer_clear(mhd2d_trainer* tr) {
  if (!tr) return -1;
  TCHK(tr, hipSetDevice(tr->device));
  TCHK(tr, hipMemsetAsync(tr->cnt, 0, 2 * sizeof(unsigned long long), tr->stream));
  TCHK(tr, hipStreamSynchronize(tr->stream));
  tr->n = tr->npos = 0;
  return 0;
}

int mhd2d_trainer_get_data(mhd2d_trainer* tr, int64_t offset, int64_t count, float* X, float* Y, int64_t* idx) {
  if (!tr || offset < 0 || count < 0 || offset + count > tr->n) return -1;
  TCHK(tr, hipSetDevice(tr->device));
  if (X) TCHK(tr, hipMemcpy(X, tr->X + offset * MLP_NIN, count * MLP_NIN * sizeof(float), hipMemcpyDeviceToHost));
  if (Y) TCHK(tr, hipMemcpy(Y, tr->Y + offset, count * sizeof(float), hipMemcpyDeviceToHost));
  if (idx) TCHK(tr, hipMemcpy(idx, tr->idx + offset, count * sizeof(long long), hipMemcpyDeviceToHost));
  return 0;
}

int mhd2d_trainer_set_data(mhd2d_trainer* tr, const float* X, const float* Y, int64_t n) {
  if (!tr || !X || !Y || n < 0 || n > tr->cap) return -1;
  TCHK(tr, hipSetDevice(tr->device));
  TCHK(tr, hipMemcpy(tr->X, X, n * MLP_NIN * sizeof(float), hipMemcpyHostToDevice));
  TCHK(tr, hipMemcpy(tr->Y, Y, n * sizeof(float), hipMemcpyHostToDevice));
  tr->n = n;
  long long np = 0;
  for (int64_t k = 0; k < n; k++) np += Y[k] > 0.5f;
  tr->npos = np;
  unsigned long long h[2] = {(unsigned long long)n, (unsigned long long)np};
  TCHK(tr, hipMemcpy(tr->cnt, h, sizeof h, hipMemcpyHostToDevice));
  return 0;
}

int mhd2d_trainer_grad(mhd2d_trainer* tr, const float* params, int64_t offset, int32_t count, float pos_weight,
                       float* grad, float* loss) {
  if (!tr || !params || count <= 0 || count > tr->bmax || offset < 0 || offset + count > tr->n) return -1;
  TCHK(tr, hipSetDevice(tr->device));
  TCHK(tr, hipMemcpyAsync(tr->P, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, tr->stream));
  int rc = batch_grad(tr, tr->X + offset * MLP_NIN, tr->Y + offset, nullptr, count, pos_weight);
  if (rc) return rc;
  if (grad) TCHK(tr, hipMemcpyAsync(grad, tr->G, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, tr->stream));
  if (loss) TCHK(tr, hipMemcpyAsync(loss, tr->lsum, sizeof(float), hipMemcpyDeviceToHost, tr->stream));
  TCHK(tr, hipStreamSynchronize(tr->stream));
  return 0;
}

int mhd2d_trainer_fit(mhd2d_trainer* tr, float* params, int32_t steps, int32_t batch, float lr, float pos_weight,
                      uint64_t seed, float* losses) {
  if (!tr || !params || steps <= 0 || batch <= 0 || batch > tr->bmax || tr->n <= 0) return -1;
  TCHK(tr, hipSetDevice(tr->device));
  hipStream_t st = tr->stream;
  TCHK(tr, hipMemcpyAsync(tr->P, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, st));
  TCHK(tr, hipMemsetAsync(tr->Mo, 0, MLP_NPARAM * sizeof(float), st));
  TCHK(tr, hipMemsetAsync(tr->Vo, 0, MLP_NPARAM * sizeof(float), st));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  std::vector<float> lh(losses ? steps : 0);
  float c1 = 1.f, c2 = 1.f;
  for (int s = 0; s < steps; s++) {
    hipLaunchKernelGGL(k_rows, dim3(nb_of(batch)), dim3(BS), 0, st, tr->rows, batch, (long long)tr->n,
                       (unsigned long long)(seed * 1000003ull + (unsigned long long)s));
    int rc = batch_grad(tr, tr->X, tr->Y, tr->rows, batch, pos_weight);
    if (rc) return rc;
    c1 *= b1;
    c2 *= b2;
    hipLaunchKernelGGL(k_adam, dim3(nb_of(MLP_NPARAM)), dim3(BS), 0, st, tr->P, (const float*)tr->G, tr->Mo, tr->Vo,
                       MLP_NPARAM, lr, b1, b2, eps, 1.f - c1, 1.f - c2);
    if (losses) TCHK(tr, hipMemcpyAsync(&lh[s], tr->lsum, sizeof(float), hipMemcpyDeviceToHost, st));
  }
  TCHK(tr, hipMemcpyAsync(params, tr->P, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, st));
  TCHK(tr, hipStreamSynchronize(st));
  if (losses) memcpy(losses, lh.data(), steps * sizeof(float));
  return 0;
}

}  // extern "C"
