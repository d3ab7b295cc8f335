// mhd2d_limit_cell.hpp -- the per-cell parts of the limiter pass that both
// the limiter kernels and the cell kernel's fused epilogue apply: the
// positivity limiter and the Balsara divergence-free reconstruction.
#pragma once
#include "mhd2d_internal.hpp"

namespace mhd2d_impl {

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

// Bound that settles a cell without its point evaluations.  On [-1,1]^2 the
// basis functions (1, x, y, x^2-1/3, xy, y^2-1/3, x^3-3x/5, (x^2-1/3)y,
// x(y^2-1/3), y^3-3y/5) are at most (1, 1, 1, 2/3, 1, 2/3, 2/5, 2/3, 2/3, 2/5)
// in magnitude, so at every point a variable lies within d_n = sum_b |m_b|
// max|phi_b| of its average, and the point sums round to within
// 1e-12 (|ubar_n| + d_n) of the exact value.  When the density cannot move
// by |rbar - eps| and the pressure's lower bound over those boxes is positive
// (by more than 1e-10 of its terms, far above their rounding), no point of
// either pass below fires, eta stays 1 and the final scaling by 1 leaves the
// modes as they are: returning early gives the same bits.
__host__ __device__ constexpr double pp_pmax(int b) {
  return b == 0 || b == 1 || b == 2 || b == 4 ? 1.0 : b == 6 || b == 9 ? 0.4 : 2.0 / 3.0;
}
// the test from the sums a[n] = sum_{b >= 1} |m_b| max|phi_b| (summed over b
// in increasing order) and the averages avg[n]
__device__ __forceinline__ bool pp_settled_sums(const double a[8], const double avg[8], double eps) {
  double d[8];
#pragma unroll
  for (int n = 0; n < 8; n++) d[n] = a[n] * (1.0 + 1e-12) + 1e-12 * (fabs(avg[n]) + a[n]);
  const double rbar = avg[0];
  const double rlo = rbar - d[0];
  if (!(d[0] < fabs(rbar - eps)) || !(rlo > 0.0)) return false;
  double kin = 0.0, mag = 0.0;
#pragma unroll
  for (int n = 1; n < 4; n++) {
    const double v = fabs(avg[n]) + d[n];
    kin += v * v;
  }
#pragma unroll
  for (int n = 5; n < 8; n++) {
    const double v = fabs(avg[n]) + d[n];
    mag += v * v;
  }
  const double k = 0.5 * kin / rlo, b2 = 0.5 * mag;
  const double plo = (avg[4] - d[4]) - k - b2;
  return plo > 1e-10 * (fabs(avg[4]) + d[4] + k + b2);
}
__device__ __forceinline__ bool pp_settled(const double m[NPL], double eps) {
  double a[8], avg[8];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    double s = 0.0;
#pragma unroll
    for (int b = 1; b < 10; b++) {
      const int p = PL(b, n);
      if (p >= 0) s += fabs(m[p]) * pp_pmax(b);
    }
    a[n] = s;
    avg[n] = m[PL(0, n)];
  }
  return pp_settled_sums(a, avg, eps);
}

// returns whether the limiter changed the cell (a scaling by eta = 1 leaves
// every value's bits as they were)
__device__ __forceinline__ bool pp_cell(double m[NPL]) {
  const double eps = 1e-13;
  if (pp_settled(m, eps)) return false;
  bool changed = false;
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
    changed = true;
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
  if (eta < 1.0) {
    eta = 0.5 * eta;
    changed = true;
  }
#pragma unroll
  for (int n = 0; n < 8; n++)
#pragma unroll
    for (int b = 1; b < 10; b++) {
      const int p = PL(b, n);
      if (p >= 0) m[p] = eta * m[p];
    }
  return changed;
}

// div_free_Balsara.f90:15-78: in-cell B from the four faces; set(p, v)
// receives plane p of the reconstruction
// div_free_Balsara.f90:15-78 for one Bx or By plane p of interior cell (i,
// j), from its four faces (Bx modes on the x faces i-1/2, i+1/2, By modes on
// the y faces j-1/2, j+1/2).  Contraction is off, so every kernel that
// reconstructs (the limiter passes, the fused cell epilogue, k_divfree)
// gets the same bits, each operation rounded as the reference rounds it.
// p is a constant wherever this is inlined, so the choice below folds away.
__device__ __forceinline__ double divfree_plane(int p, const double* __restrict__ F, const Geo& g, int i, int j) {
  MHD_NOCONTRACT
  const long long pl = g.plane;
  const long long cR = cix(g, i, j), cL = cix(g, i - 1, j), cD = cix(g, i, j - 1);
  auto a = [&](int k, long long c) { return F[k * pl + c]; };         // Bx mode k + 1 on an x face
  auto b = [&](int k, long long c) { return F[(3 + k) * pl + c]; };   // By mode k + 1 on a y face
  const double rxy = g.hx / g.hy, ryx = g.hy / g.hx;
  if (p == PL(0, 5)) return 0.5 * (a(0, cR) + a(0, cL)) + (1.0 / 6.0) * rxy * (b(1, cR) - b(1, cD));
  if (p == PL(1, 5)) return 0.5 * (a(0, cR) - a(0, cL)) + (1.0 / 15.0) * rxy * (b(2, cR) - b(2, cD));
  if (p == PL(2, 5)) return 0.5 * (a(1, cR) + a(1, cL));
  if (p == PL(3, 5)) return -(1.0 / 4.0) * rxy * (b(1, cR) - b(1, cD));
  if (p == PL(4, 5)) return 0.5 * (a(1, cR) - a(1, cL));
  if (p == PL(5, 5)) return 0.5 * (a(2, cR) + a(2, cL));
  if (p == PL(6, 5)) return -(1.0 / 6.0) * rxy * (b(2, cR) - b(2, cD));
  if (p == PL(8, 5)) return 0.5 * (a(2, cR) - a(2, cL));
  if (p == PL(0, 6)) return 0.5 * (b(0, cR) + b(0, cD)) + (1.0 / 6.0) * ryx * (a(1, cR) - a(1, cL));
  if (p == PL(2, 6)) return 0.5 * (b(0, cR) - b(0, cD)) + (1.0 / 15.0) * ryx * (a(2, cR) - a(2, cL));
  if (p == PL(1, 6)) return 0.5 * (b(1, cR) + b(1, cD));
  if (p == PL(5, 6)) return -(1.0 / 4.0) * ryx * (a(1, cR) - a(1, cL));
  if (p == PL(4, 6)) return 0.5 * (b(1, cR) - b(1, cD));
  if (p == PL(3, 6)) return 0.5 * (b(2, cR) + b(2, cD));
  if (p == PL(9, 6)) return -(1.0 / 6.0) * ryx * (a(2, cR) - a(2, cL));
  return 0.5 * (b(2, cR) - b(2, cD));                                   // PL(7, 6)
}
template <class Set>
__device__ __forceinline__ void divfree_apply(Set&& set, const double* __restrict__ F, const Geo& g, int i, int j) {
  constexpr int planes[16] = {PL(0, 5), PL(1, 5), PL(2, 5), PL(3, 5), PL(4, 5), PL(5, 5), PL(6, 5), PL(8, 5),
                              PL(0, 6), PL(2, 6), PL(1, 6), PL(5, 6), PL(4, 6), PL(3, 6), PL(9, 6), PL(7, 6)};
#pragma unroll
  for (int k = 0; k < 16; k++) set(planes[k], divfree_plane(planes[k], F, g, i, j));
}
__device__ __forceinline__ void divfree_cell(double m[NPL], const double* __restrict__ F, const Geo& g, int i, int j) {
  divfree_apply([&](int p, double v) { m[p] = v; }, F, g, i, j);
}

// append the flagged lanes' interior-cell ids to the positivity list (one
// atomic per wave)
__device__ __forceinline__ void pp_append(bool flag, int t, int* __restrict__ list, DevScal* lst) {
  const unsigned long long b = __ballot(flag);
  if (b) {
    const int lane = __lane_id(), leader = __ffsll((long long)b) - 1;
    unsigned int base = 0u;
    if (lane == leader) base = atomicAdd(&lst->npp, (unsigned int)__popcll(b));
    base = (unsigned int)__shfl((int)base, leader);
    if (flag) list[base + (unsigned int)__popcll(b & ((1ull << lane) - 1ull))] = t;
  }
}

}  // namespace mhd2d_impl
