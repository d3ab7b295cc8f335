// mhd2d_limit.hip -- the limiter pass (TVB_Limiter_P2 with the optional
// positivity limiter and the Balsara divergence-free reconstruction fused)
// and the KXRCF troubled-cell detector.
#include "mhd2d_internal.hpp"

namespace mhd2d_impl {
namespace {

// ---------------------------------------------------------------------------
// positivity limiter of one cell (pp_Limiter.f90:10-81) on modes m[]
__device__ __forceinline__ double glleval(const double* m, int n, int i1, int j1, int fam) {
  const double x = fam == 0 ? lamL(i1) : lam(i1);
  const double y = fam == 0 ? lam(j1) : lamL(j1);
  double a = 0.0;
#pragma unroll
  for (int b = 0; b < 10; b++) {
    const int p = PL(b, n);
    if (p >= 0) a += m[p] * phi(b, x, y);
  }
  return a;
}

__device__ __forceinline__ void pp_cell(double m[NPL]) {
  const double eps = 1e-13;
  // density: eta = the LAST |(rbar - eps)/(rbar - rq)| below 1.  For
  // doubles, RN(|a/b|) < 1 exactly when |a| < |b|, so only that point's
  // quotient is formed.
  double eta = 1.0;
  const double rbar = m[PL(0, 0)];
  const double na = rbar - eps;
  double nb = 0.0;
  bool any = false;
#pragma unroll
  for (int i1 = 0; i1 < 5; i1++)
#pragma unroll
    for (int j1 = 0; j1 < 5; j1++)
#pragma unroll
      for (int f = 0; f < 2; f++) {
        const double b = rbar - glleval(m, 0, i1, j1, f);
        if (fabs(na) < fabs(b)) { nb = b; any = true; }
      }
  if (any) eta = fabs(na / nb);
  if (eta < 1.0) {
#pragma unroll
    for (int b = 1; b < 6; b++) m[PL(b, 0)] = 0.5 * eta * m[PL(b, 0)];
  }
  // pressure.  The reference keeps eta1 from the last negative-pressure
  // point and takes eta = min(eta, eta1) at every point, i.e. eta is the
  // minimum of eta1 over the negative-pressure points.  With the nodes the
  // reference holds (-1, 0, 0, 0, 1) only 30 of its 50 points are distinct;
  // their pressures are evaluated unrolled and the bisection
  // (pp_Limiter.f90:101-125, at most 47 halvings) runs only where p < 0.
  double ubar[8];
#pragma unroll
  for (int n = 0; n < 8; n++) ubar[n] = m[PL(0, n)];
  unsigned neg = 0u;   // bit k: distinct point k has negative pressure
#pragma unroll
  for (int k = 0; k < 30; k++) {
    // k < 15: family 1 (x = -1, 0, 1; y = Gauss); k >= 15: family 2
    const int f = k / 15, a = (k % 15) / 5, b = k % 5;
    const double x = f == 0 ? (a == 0 ? -1.0 : a == 1 ? 0.0 : 1.0) : lam(b);
    const double y = f == 0 ? lam(b) : (a == 0 ? -1.0 : a == 1 ? 0.0 : 1.0);
    double uq[8];
#pragma unroll
    for (int n = 0; n < 8; n++) {
      double acc = 0.0;
#pragma unroll
      for (int bb = 0; bb < 10; bb++) {
        const int p = PL(bb, n);
        if (p >= 0) acc += m[p] * phi(bb, x, y);
      }
      uq[n] = acc;
    }
    if (pressure<double>(uq) < 0.0) neg |= 1u << k;
  }
  eta = 1.0;
  while (neg) {
    const int k = __builtin_ctz(neg);
    neg &= neg - 1u;
    const int f = k / 15, a = (k % 15) / 5, b = k % 5;
    const double g = a == 0 ? -1.0 : a == 1 ? 0.0 : 1.0;
    const double lb = b == 0 ? lam(0) : b == 1 ? lam(1) : b == 2 ? lam(2) : b == 3 ? lam(3) : lam(4);
    const double x = f == 0 ? g : lb, y = f == 0 ? lb : g;
    double uq[8];
    {
      MHD_NOCONTRACT
      const double pv[10] = {1.0, x, y, x * x - 1.0 / 3.0, x * y, y * y - 1.0 / 3.0, x * x * x - 3.0 * x / 5.0,
                             (x * x - 1.0 / 3.0) * (y), (x) * (y * y - 1.0 / 3.0), y * y * y - 3.0 * y / 5.0};
#pragma unroll
      for (int n = 0; n < 8; n++) {
        double acc = 0.0;
#pragma unroll
        for (int bb = 0; bb < 10; bb++) {
          const int p = PL(bb, n);
          if (p >= 0) acc = fma(m[p], pv[bb], acc);
        }
        uq[n] = acc;
      }
    }
    double ta = 0.0, tb = 1.0, ut[8];
    while (tb - ta > 1e-14) {
      const double tq = 0.5 * (ta + tb);
#pragma unroll
      for (int n = 0; n < 8; n++) ut[n] = tq * uq[n] + (1 - tq) * ubar[n];
      if (pressure<double>(ut) < 1e-12) tb = tq;
      else ta = tq;
    }
    const double tq = ta;
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const double sq = tq * uq[n] + (1 - tq) * ubar[n];
      s1 += (sq - ubar[n]) * (sq - ubar[n]);
      s2 += (uq[n] - ubar[n]) * (uq[n] - ubar[n]);
    }
    const double eta1 = sqrt(s1) / sqrt(s2);
    if (eta1 < eta) eta = eta1;
  }
  if (eta < 1.0) eta = 0.5 * eta;
#pragma unroll
  for (int n = 0; n < 8; n++)
#pragma unroll
    for (int b = 1; b < 10; b++) {
      const int p = PL(b, n);
      if (p >= 0) m[p] = eta * m[p];
    }
}

// div_free_Balsara.f90:15-78: in-cell B from the four faces
__device__ __forceinline__ void divfree_cell(double m[NPL], const double* __restrict__ F, const Geo& g, int i, int j) {
  const long long pl = g.plane;
  const long long cR = cix(g, i, j), cL = cix(g, i - 1, j), cD = cix(g, i, j - 1);
  const double a0R = F[0 * pl + cR], a1R = F[1 * pl + cR], a2R = F[2 * pl + cR];
  const double a0L = F[0 * pl + cL], a1L = F[1 * pl + cL], a2L = F[2 * pl + cL];
  const double b0U = F[3 * pl + cR], b1U = F[4 * pl + cR], b2U = F[5 * pl + cR];
  const double b0D = F[3 * pl + cD], b1D = F[4 * pl + cD], b2D = F[5 * pl + cD];
  const double rxy = g.hx / g.hy, ryx = g.hy / g.hx;
  m[PL(0, 5)] = 0.5 * (a0R + a0L) + (1.0 / 6.0) * rxy * (b1U - b1D);
  m[PL(1, 5)] = 0.5 * (a0R - a0L) + (1.0 / 15.0) * rxy * (b2U - b2D);
  m[PL(2, 5)] = 0.5 * (a1R + a1L);
  m[PL(3, 5)] = -(1.0 / 4.0) * rxy * (b1U - b1D);
  m[PL(4, 5)] = 0.5 * (a1R - a1L);
  m[PL(5, 5)] = 0.5 * (a2R + a2L);
  m[PL(6, 5)] = -(1.0 / 6.0) * rxy * (b2U - b2D);
  m[PL(8, 5)] = 0.5 * (a2R - a2L);
  m[PL(0, 6)] = 0.5 * (b0U + b0D) + (1.0 / 6.0) * ryx * (a1R - a1L);
  m[PL(2, 6)] = 0.5 * (b0U - b0D) + (1.0 / 15.0) * ryx * (a2R - a2L);
  m[PL(1, 6)] = 0.5 * (b1U + b1D);
  m[PL(5, 6)] = -(1.0 / 4.0) * ryx * (a1R - a1L);
  m[PL(4, 6)] = 0.5 * (b1U - b1D);
  m[PL(3, 6)] = 0.5 * (b2U + b2D);
  m[PL(9, 6)] = -(1.0 / 6.0) * ryx * (a2R - a2L);
  m[PL(7, 6)] = 0.5 * (b2U - b2D);
}

__device__ __forceinline__ bool is_bplane(int p) { return (p >= 30 && p < 42) || p >= 48; }

// limiter pass (TVB_Limiter_P2.f90:16-121 + pp + div_free_Balsara) over
// the whole array: interior cells are limited, halo cells copied (or, with
// DIV, their B planes zeroed as div_free_Balsara.f90:13 does).
// flags: indicator == 2/3 reads trouble[] (1 = limit).
template <bool APPLY, bool PP, bool DIV>
__global__ __launch_bounds__(BS) void k_limit(const double* __restrict__ in, double* __restrict__ out,
                                              const double* __restrict__ F, const int* __restrict__ tmask,
                                              const double* __restrict__ trec, Geo g) {
  const long long t = xcd_block() * BS + threadIdx.x;
  const int W = g.Nx + 2;
  if (t >= (long long)W * (g.Ny + 2)) return;
  const int i = (int)(t % W), j = (int)(t / W);
  const long long c = cix(g, i, j), pl = g.plane;
  const bool interior = i >= 1 && i <= g.Nx && j >= 1 && j <= g.Ny;
  if (!interior) {
#pragma unroll 4
    for (int p = 0; p < NPL; p++) out[p * pl + c] = (DIV && is_bplane(p)) ? 0.0 : in[p * pl + c];
    return;
  }
  double m[NPL];
#pragma unroll
  for (int p = 0; p < NPL; p++) m[p] = in[p * pl + c];
  if (APPLY) {
    // TVB_Limiter_P2.f90:106-114 for the variables k_tvb marked changed
    const long long ncell = (long long)g.Nx * g.Ny, tc = (long long)(j - 1) * g.Nx + (i - 1);
    const int change = tmask[tc];
    if (change) {
#pragma unroll
      for (int n = 0; n < 8; n++) {
        if (!(change & (1 << n))) continue;
        const double Rm = trec[(0 + n) * ncell + tc], Lm = trec[(8 + n) * ncell + tc];
        const double Um = trec[(16 + n) * ncell + tc], Dm = trec[(24 + n) * ncell + tc];
#pragma unroll
        for (int b = 6; b < 10; b++) {
          const int p = PL(b, n);
          if (p >= 0) m[p] = 0.0;
        }
        m[PL(4, n)] = 0.0;
        m[PL(1, n)] = 0.5 * (Rm + Lm);
        m[PL(2, n)] = 0.5 * (Um + Dm);
        m[PL(3, n)] = (3.0 / 4.0) * (Rm - Lm);
        m[PL(5, n)] = (3.0 / 4.0) * (Um - Dm);
      }
    }
  }
  if (PP) pp_cell(m);
  if (DIV) divfree_cell(m, F, g, i, j);
#pragma unroll
  for (int p = 0; p < NPL; p++) out[p * pl + c] = m[p];
}


// KXRCF_Detector.f90 (uG per cell, flags reset each call, Ck given).
// The inflow-face integral sum_q w_q (u_self(q) - u_nb(q)) of a variable is
// linear in the modes: sum_b (U_self[b] W(side, b) - U_nb[b] W(opp, b)) with
// W(side, b) = sum_q w_q phi_b(side, s_q), the same Gauss sums the reference
// evaluates per point (KXRCF_Detector.f90:62-88), folded at compile time.
// Sums that vanish exactly (odd or orthogonal tangential factors) are
// taken as 0 rather than as their ~1e-17 rounding residue.
__host__ __device__ constexpr double WTR(int side, int b) {
  double s = 0.0;
  for (int q = 0; q < 5; q++) s += wgt(q) * phiTr(side, b, lam(q), P2(lam(q)), P3(lam(q)), T23(lam(q)));
  return (s < 1e-13 && s > -1e-13) ? 0.0 : s;
}

template <int SIDE>
__device__ __forceinline__ double face_jump_int(const double* __restrict__ U, long long pl, long long cself,
                                                long long cnb, int n) {
  constexpr int OPP = SIDE == 0 ? 1 : SIDE == 1 ? 0 : SIDE == 2 ? 3 : 2;
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int bb = 0; bb < 10; bb++) {
    const int p = PL(bb, n);
    if (p < 0) continue;
    if (WTR(SIDE, bb) != 0.0) a += U[p * pl + cself] * WTR(SIDE, bb);
    if (WTR(OPP, bb) != 0.0) b += U[p * pl + cnb] * WTR(OPP, bb);
  }
  return a - b;
}

// KXRCF_Detector.f90:55-105 for cell (i, j); hk = (hx hy)^((k+1)/4), k = 2
__device__ __forceinline__ bool kxrcf_cell(const double* __restrict__ U, const Geo& g, int i, int j, double Ck,
                                           double hk) {
  const long long pl = g.plane, c = cix(g, i, j);
  const long long cE = cix(g, i + 1, j), cW = cix(g, i - 1, j), cN = cix(g, i, j + 1), cS = cix(g, i, j - 1);
  const bool inE = U[PL(0, 1) * pl + cE] < 0.0, inW = U[PL(0, 1) * pl + cW] > 0.0;
  const bool inN = U[PL(0, 2) * pl + cN] < 0.0, inS = U[PL(0, 2) * pl + cS] > 0.0;
  double Lij = 1e-12;
  if (inE) Lij += 5 * g.hy;
  if (inW) Lij += 5 * g.hy;
  if (inN) Lij += 5 * g.hx;
  if (inS) Lij += 5 * g.hx;
  const double den0 = hk * Lij;
  bool flag = false;
#pragma unroll
  for (int n = 0; n < 8; n++) {
    double I = 0.0;
    if (inE) I += g.hy1 * face_jump_int<0>(U, pl, c, cE, n);
    if (inW) I += g.hy1 * face_jump_int<1>(U, pl, c, cW, n);
    if (inN) I += g.hx1 * face_jump_int<2>(U, pl, c, cN, n);
    if (inS) I += g.hx1 * face_jump_int<3>(U, pl, c, cS, n);
    // max over the 25 Gauss points of |u_n| + 1e-12, row by row in eta:
    // u(x, y_j) = c0 + c1 x + c2 P2(x) + c3 P3(x)
    const double m0 = U[PL(0, n) * pl + c], m1 = U[PL(1, n) * pl + c], m2 = U[PL(2, n) * pl + c];
    const double m3 = U[PL(3, n) * pl + c], m4 = U[PL(4, n) * pl + c], m5 = U[PL(5, n) * pl + c];
    const double e6 = n == 5 ? U[PL(6, 5) * pl + c] : n == 6 ? U[PL(7, 6) * pl + c] : 0.0;
    const double e8 = n == 5 ? U[PL(8, 5) * pl + c] : n == 6 ? U[PL(9, 6) * pl + c] : 0.0;
    double uinf = 0.0;
#pragma unroll
    for (int j1 = 0; j1 < 5; j1++) {
      const double y = lam(j1);
      double c0 = m0 + m2 * y + m5 * P2(y), c1 = m1 + m4 * y, c2 = m3, c3 = 0.0;
      if (n == 5) { c1 += e8 * P2(y); c3 = e6; }           // Bx: phi9 = x P2(y), phi7 = P3(x)
      if (n == 6) { c2 += e6 * y; c0 += e8 * P3(y); }       // By: phi8 = P2(x) y, phi10 = P3(y)
#pragma unroll
      for (int i1 = 0; i1 < 5; i1++) {
        const double x = lam(i1);
        const double a = c0 + c1 * x + c2 * P2(x) + c3 * P3(x);
        uinf = fmax(uinf, fabs(a) + 1e-12);
      }
    }
    if (fabs(I) / (den0 * uinf) >= Ck) flag = true;
  }
  return flag;
}

// one count per wave of the flagged lanes (nflag of the diagnostics)
__device__ __forceinline__ void count_flags(bool flag, DevScal* sc) {
  const unsigned long long b = __ballot(flag);
  if (b && __lane_id() == __ffsll((long long)__ballot(1)) - 1)
    atomicAdd((unsigned long long*)&sc->nflag, (unsigned long long)__popcll(b));
}

__global__ __launch_bounds__(BS) void k_kxrcf(const double* __restrict__ U, int* __restrict__ trouble, Geo g,
                                              double Ck, double hk, DevScal* sc) {
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  if (t >= (long long)g.Nx * g.Ny) return;
  const int i = 1 + (int)(t % g.Nx), j = 1 + (int)(t / g.Nx);
  const bool flag = kxrcf_cell(U, g, i, j, Ck, hk);
  trouble[t] = flag ? 1 : 0;
  if (sc) count_flags(flag, sc);
}

__global__ void k_list_reset(DevScal* lst) {
  if (threadIdx.x == 0) lst->nlist = 0u;
}

// KXRCF over the interior cells for the limiter pass: trouble[] and the
// compacted list of flagged cells (one atomic per wave), tmask[] cleared
__global__ __launch_bounds__(BS) void k_detect(const double* __restrict__ U, int* __restrict__ trouble,
                                               int* __restrict__ tmask, int* __restrict__ list, Geo g, double Ck,
                                               double hk, DevScal* sc, DevScal* lst) {
  const long long t = xcd_block() * BS + threadIdx.x;
  if (t >= (long long)g.Nx * g.Ny) return;
  const int i = 1 + (int)(t % g.Nx), j = 1 + (int)(t / g.Nx);
  const bool flag = kxrcf_cell(U, g, i, j, Ck, hk);
  trouble[t] = flag ? 1 : 0;
  tmask[t] = 0;
  if (sc) count_flags(flag, sc);
  const unsigned long long b = __ballot(flag);
  if (b) {
    const int lane = __lane_id(), leader = __ffsll((long long)b) - 1;
    unsigned int base = 0u;
    if (lane == leader) base = atomicAdd(&lst->nlist, (unsigned int)__popcll(b));
    base = (unsigned int)__shfl((int)base, leader);
    if (flag) list[base + (unsigned int)__popcll(b & ((1ull << lane) - 1ull))] = (int)t;
  }
}

// TVB decision and edge values of one cell (TVB_Limiter_P2.f90:24-104):
// returns the change bits per variable; Rm/Lm (x) and Um/Dm (y) are the
// modified edge deviations DeltaUR1mod .. DeltaUD1mod.
__device__ __forceinline__ int tvb_cell(const double* __restrict__ in, const Geo& g, int i, int j, double M,
                                        double Rm[8], double Lm[8], double Um[8], double Dm[8]) {
  const long long pl = g.plane, c = cix(g, i, j);
  int change = 0;
  double U[8];
#pragma unroll
  for (int n = 0; n < 8; n++) U[n] = in[PL(0, n) * pl + c];
  {
    double dR[8], dL[8], r1[8], l1[8];
    const long long cE = cix(g, i + 1, j), cW = cix(g, i - 1, j);
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const double m1 = in[PL(1, n) * pl + c], m3 = in[PL(3, n) * pl + c];
      dR[n] = in[PL(0, n) * pl + cE] - U[n];
      dL[n] = U[n] - in[PL(0, n) * pl + cW];
      r1[n] = m1 + (2.0 / 3.0) * m3;
      l1[n] = m1 - (2.0 / 3.0) * m3;
    }
    tvb_dir<1>(U, dR, dL, r1, l1, M * (g.hx * g.hx), Rm, Lm);
#pragma unroll
    for (int n = 0; n < 8; n++)
      if (fabs(Rm[n] - r1[n]) > 1e-6 || fabs(Lm[n] - l1[n]) > 1e-6) change |= 1 << n;
  }
  {
    double dR[8], dL[8], r1[8], l1[8];
    const long long cN = cix(g, i, j + 1), cS = cix(g, i, j - 1);
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const double m2 = in[PL(2, n) * pl + c], m5 = in[PL(5, n) * pl + c];
      dR[n] = in[PL(0, n) * pl + cN] - U[n];
      dL[n] = U[n] - in[PL(0, n) * pl + cS];
      r1[n] = m2 + (2.0 / 3.0) * m5;
      l1[n] = m2 - (2.0 / 3.0) * m5;
    }
    tvb_dir<2>(U, dR, dL, r1, l1, M * (g.hy * g.hy), Um, Dm);
#pragma unroll
    for (int n = 0; n < 8; n++)
      if (fabs(Um[n] - r1[n]) > 1e-6 || fabs(Dm[n] - l1[n]) > 1e-6) change |= 1 << n;
  }
  return change;
}

// TVB decision over the interior cells that the detector marked: IND 1
// (every cell, TVB_Limiter_P2.f90 as shipped) marks all of them; otherwise
// trouble[] holds the detector's decision (KXRCF_Detector, the reference's
// commented gating at TVB_Limiter_P2.f90:12,22).  tmask[] gets the change
// bits and trec[] the edge values for k_limit<true, ..> to apply.  Kept
// apart from the detector and from k_limit so that the eigen-decompositions'
// register demand does not lower the occupancy of the passes every cell
// takes.
template <int IND>
__global__ __launch_bounds__(BS) void k_tvb(const double* __restrict__ in, int* __restrict__ trouble,
                                            int* __restrict__ tmask, double* __restrict__ trec, Geo g, double M,
                                            double Ck, double hk, DevScal* sc) {
  const long long ncell = (long long)g.Nx * g.Ny;
  const long long t = xcd_block() * BS + threadIdx.x;
  if (t >= ncell) return;
  const int i = 1 + (int)(t % g.Nx), j = 1 + (int)(t / g.Nx);
  const bool flag = IND == MHD2D_IND_ALL ? true : trouble[t] != 0;
  if (IND == MHD2D_IND_ALL) trouble[t] = 1;
  int change = 0;
  if (flag) {
    double Rm[8], Lm[8], Um[8], Dm[8];
    change = tvb_cell(in, g, i, j, M, Rm, Lm, Um, Dm);
    if (change) {
#pragma unroll
      for (int n = 0; n < 8; n++) {
        trec[(0 + n) * ncell + t] = Rm[n];
        trec[(8 + n) * ncell + t] = Lm[n];
        trec[(16 + n) * ncell + t] = Um[n];
        trec[(24 + n) * ncell + t] = Dm[n];
      }
    }
  }
  tmask[t] = change;
}

// TVB on the listed (flagged) cells only; a fixed grid strides the list
__global__ __launch_bounds__(BS) void k_tvb_list(const double* __restrict__ in, const int* __restrict__ list,
                                                 int* __restrict__ tmask, double* __restrict__ trec, Geo g,
                                                 double M, const DevScal* __restrict__ lst) {
  const long long ncell = (long long)g.Nx * g.Ny;
  const unsigned int n = lst->nlist;
  for (unsigned int k = blockIdx.x * BS + threadIdx.x; k < n; k += gridDim.x * BS) {
    const long long t = list[k];
    const int i = 1 + (int)(t % g.Nx), j = 1 + (int)(t / g.Nx);
    double Rm[8], Lm[8], Um[8], Dm[8];
    const int change = tvb_cell(in, g, i, j, M, Rm, Lm, Um, Dm);
    if (change) {
#pragma unroll
      for (int q = 0; q < 8; q++) {
        trec[(0 + q) * ncell + t] = Rm[q];
        trec[(8 + q) * ncell + t] = Lm[q];
        trec[(16 + q) * ncell + t] = Um[q];
        trec[(24 + q) * ncell + t] = Dm[q];
      }
    }
    tmask[t] = change;
  }
}

}  // namespace

// (hx hy)^(0.25 (k + 1)) with k = 2 (KXRCF_Detector.f90:96)
static double kx_hk(const Geo& g) { return pow(g.hx * g.hy, 0.25 * (2 + 1)); }

int launch_kxrcf(mhd2d_ctx* ctx, const double* U, DevScal* sc) {
  const Geo& g = ctx->g;
  Timed tm(ctx, KC_KXRCF);
  hipLaunchKernelGGL(k_kxrcf, dim3(nblk((long long)g.Nx * g.Ny)), dim3(BS), 0, ctx->stream, U, ctx->trouble, g,
                     ctx->cfg.Ck, kx_hk(g), sc);
  LAUNCHCHK(ctx);
  return 0;
}

// limiter pass in -> out: detector + TVB decisions (k_tvb), then one pass
// over every cell that applies them, the positivity limiter and the
// divergence-free reconstruction (k_limit)
int launch_limit(mhd2d_ctx* ctx, const double* in, double* out, const double* F, bool tvb, bool pp, bool div,
                 DevScal* sc) {
  const Geo& g = ctx->g;
  const long long n = (long long)(g.Nx + 2) * (g.Ny + 2);
  const int ind = ctx->cfg.indicator;
  if (tvb && ind == MHD2D_IND_NONE) tvb = false;
  Timed tm(ctx, KC_LIMIT);
  if (tvb) {
    const dim3 gt(nblk((long long)g.Nx * g.Ny)), bt(BS);
    if (ind == MHD2D_IND_ALL) {
      hipLaunchKernelGGL(k_tvb<MHD2D_IND_ALL>, gt, bt, 0, ctx->stream, in, ctx->trouble, ctx->tmask, ctx->trec, g,
                         ctx->cfg.M, ctx->cfg.Ck, kx_hk(g), sc);
    } else {
      // detector with a compacted list, then TVB over the list only
      hipLaunchKernelGGL(k_list_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->sc);
      hipLaunchKernelGGL(k_detect, gt, bt, 0, ctx->stream, in, ctx->trouble, ctx->tmask, ctx->tlist, g, ctx->cfg.Ck,
                         kx_hk(g), sc, ctx->sc);
      const unsigned nb = std::min<unsigned>(nblk((long long)g.Nx * g.Ny), 1024u);
      hipLaunchKernelGGL(k_tvb_list, dim3(nb), bt, 0, ctx->stream, in, (const int*)ctx->tlist, ctx->tmask,
                         ctx->trec, g, ctx->cfg.M, (const DevScal*)ctx->sc);
    }
    LAUNCHCHK(ctx);
  }
  const dim3 gr(nblk(n)), bl(BS);
#define LIM(a, b, c) \
  hipLaunchKernelGGL((k_limit<a, b, c>), gr, bl, 0, ctx->stream, in, out, F, ctx->tmask, ctx->trec, g)
  if (tvb && pp && div) LIM(true, true, true);
  else if (tvb && pp) LIM(true, true, false);
  else if (tvb && div) LIM(true, false, true);
  else if (tvb) LIM(true, false, false);
  else if (pp && div) LIM(false, true, true);
  else if (pp) LIM(false, true, false);
  else if (div) LIM(false, false, true);
  else LIM(false, false, false);
#undef LIM
  LAUNCHCHK(ctx);
  return 0;
}

}  // namespace mhd2d_impl
