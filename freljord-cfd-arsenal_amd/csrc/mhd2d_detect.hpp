// mhd2d_detect.hpp -- troubled-cell detection shared by the limiter and the
// learned indicator: KXRCF_Detector per cell, flag counting and the
// compacted list of flagged cells.
#pragma once
#include "mhd2d_internal.hpp"

namespace mhd2d_impl {

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

template <int SIDE, class SV>
__device__ __forceinline__ double face_jump_int(const SV& U, long long cself, long long cnb, int n) {
  constexpr int OPP = SIDE == 0 ? 1 : SIDE == 1 ? 0 : SIDE == 2 ? 3 : 2;
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int bb = 0; bb < 10; bb++) {
    const int p = PL(bb, n);
    if (p < 0) continue;
    if (WTR(SIDE, bb) != 0.0) a += U(p, cself) * WTR(SIDE, bb);
    if (WTR(OPP, bb) != 0.0) b += U(p, cnb) * WTR(OPP, bb);
  }
  return a - b;
}

// KXRCF_Detector.f90:55-105 for cell (i, j); hk = (hx hy)^((k+1)/4), k = 2
template <class SV>
__device__ __forceinline__ bool kxrcf_cell(const SV& U, const Geo& g, int i, int j, double Ck, double hk) {
  const long long c = cix(g, i, j);
  const long long cE = cix(g, i + 1, j), cW = cix(g, i - 1, j), cN = cix(g, i, j + 1), cS = cix(g, i, j - 1);
  const bool inE = U(PL(0, 1), cE) < 0.0, inW = U(PL(0, 1), cW) > 0.0;
  const bool inN = U(PL(0, 2), cN) < 0.0, inS = U(PL(0, 2), cS) > 0.0;
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
    if (inE) I += g.hy1 * face_jump_int<0>(U, c, cE, n);
    if (inW) I += g.hy1 * face_jump_int<1>(U, c, cW, n);
    if (inN) I += g.hx1 * face_jump_int<2>(U, c, cN, n);
    if (inS) I += g.hx1 * face_jump_int<3>(U, c, cS, n);
    // max over the 25 Gauss points of |u_n| + 1e-12, row by row in eta:
    // u(x, y_j) = c0 + c1 x + c2 P2(x) + c3 P3(x)
    const double m0 = U(PL(0, n), c), m1 = U(PL(1, n), c), m2 = U(PL(2, n), c);
    const double m3 = U(PL(3, n), c), m4 = U(PL(4, n), c), m5 = U(PL(5, n), c);
    const double e6 = n == 5 ? U(PL(6, 5), c) : n == 6 ? U(PL(7, 6), c) : 0.0;
    const double e8 = n == 5 ? U(PL(8, 5), c) : n == 6 ? U(PL(9, 6), c) : 0.0;
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

// append the flagged lanes' cell ids to list (one atomic per wave)
__device__ __forceinline__ void list_append(bool flag, int t, int* __restrict__ list, DevScal* lst) {
  const unsigned long long b = __ballot(flag);
  if (b) {
    const int lane = __lane_id(), leader = __ffsll((long long)b) - 1;
    unsigned int base = 0u;
    if (lane == leader) base = atomicAdd(&lst->nlist, (unsigned int)__popcll(b));
    base = (unsigned int)__shfl((int)base, leader);
    if (flag) list[base + (unsigned int)__popcll(b & ((1ull << lane) - 1ull))] = t;
  }
}

// (hx hy)^(0.25 (k + 1)) with k = 2 (KXRCF_Detector.f90:96)
inline double kx_hk(const Geo& g) { return pow(g.hx * g.hy, 0.25 * (2 + 1)); }

}  // namespace mhd2d_impl
