// mhd2d_tables.hpp -- quadrature and basis tables of get_GLP.f90:38-56 and
// get_basis.f90:5-189 as constexpr functions, plus the device-state layout.
//
// All table entries are evaluated exactly as the reference evaluates them
// (same formulas, round-to-nearest per operation, no contraction), so that
// fully unrolled kernels fold them to the same doubles the reference holds.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mhd {

#define MHD_NOCONTRACT _Pragma("clang fp contract(off)")

// 5-point Gauss-Legendre nodes / weights and Gauss-Lobatto nodes
__host__ __device__ constexpr double lam(int q) {
  return q == 0 ? -0.9061798459386639927976269
       : q == 1 ? -0.5384693101056830910363144
       : q == 2 ? 0.0
       : q == 3 ? 0.5384693101056830910363144
                : 0.9061798459386639927976269;
}
__host__ __device__ constexpr double wgt(int q) {
  return q == 0 ? 0.2369268850561890875142640
       : q == 1 ? 0.4786286704993664680412915
       : q == 2 ? 0.5688888888888888888888889
       : q == 3 ? 0.4786286704993664680412915
                : 0.2369268850561890875142640;
}
// The reference's Gauss-Lobatto nodes as it actually holds them: lambdaL is
// undeclared in com.txt, hence an implicit INTEGER array, and get_GLP.f90:52-56
// stores -1, trunc(-0.6547) = 0, 0, trunc(0.6547) = 0, 1.  Only the
// positivity limiter (phiGLL) reads them.
__host__ __device__ constexpr double lamL(int q) {
  return q == 0 ? -1.0 : q == 4 ? 1.0 : 0.0;
}

// Legendre-product basis phi_{b+1}(x, y), b = 0..9 (get_basis.f90:49-159)
__host__ __device__ constexpr double phi(int b, double x, double y) {
  MHD_NOCONTRACT
  return b == 0 ? 1.0
       : b == 1 ? x
       : b == 2 ? y
       : b == 3 ? x * x - 1.0 / 3.0
       : b == 4 ? x * y
       : b == 5 ? y * y - 1.0 / 3.0
       : b == 6 ? x * x * x - 3.0 * x / 5.0
       : b == 7 ? (x * x - 1.0 / 3.0) * (y)
       : b == 8 ? (x) * (y * y - 1.0 / 3.0)
                : y * y * y - 3.0 * y / 5.0;
}
// hx1 * d/dx and hy1 * d/dy of phi (phixG, phiyG without the 1/hx1, 1/hy1)
__host__ __device__ constexpr double dphix(int b, double x, double y) {
  MHD_NOCONTRACT
  return b == 1 ? 1.0
       : b == 3 ? 2.0 * x
       : b == 4 ? y
       : b == 6 ? 3 * (x * x) - 3.0 / 5.0
       : b == 7 ? 2.0 * x * y
       : b == 8 ? y * y - 1.0 / 3.0
                : 0.0;
}
__host__ __device__ constexpr double dphiy(int b, double x, double y) {
  MHD_NOCONTRACT
  return b == 2 ? 1.0
       : b == 4 ? x
       : b == 5 ? 2.0 * y
       : b == 7 ? x * x - 1.0 / 3.0
       : b == 8 ? 2.0 * x * y
       : b == 9 ? 3 * (y * y) - 3.0 / 5.0
                : 0.0;
}
// mm(d) = mean of phi^2
__host__ __device__ constexpr double mmass(int b) {
  return b == 0 ? 1.0 : b == 1 ? 1.0 / 3.0 : b == 2 ? 1.0 / 3.0 : b == 3 ? 4.0 / 45.0 : b == 4 ? 1.0 / 9.0
       : b == 5 ? 4.0 / 45.0 : b == 6 ? 4.0 / 175.0 : b == 7 ? 4.0 / 135.0 : b == 8 ? 4.0 / 135.0 : 4.0 / 175.0;
}
// face-trace tables phiGR/phiGL/phiGU/phiGD (get_basis.f90), as the
// reference writes them (2/3 and 2/5 are literals, not 1 - 1/3 etc.)
__host__ __device__ constexpr double P2(double x) { MHD_NOCONTRACT return x * x - 1.0 / 3.0; }
__host__ __device__ constexpr double P3(double x) { MHD_NOCONTRACT return x * x * x - 3.0 * x / 5.0; }
__host__ __device__ constexpr double T23(double x) { MHD_NOCONTRACT return (2.0 / 3.0) * (x); }
// side: 0 = R (xi=+1), 1 = L (xi=-1), 2 = U (eta=+1), 3 = D (eta=-1); s = node value
__host__ __device__ constexpr double phiTr(int side, int b, double s, double p2, double p3, double t23) {
  return side == 0 ? (b == 0 ? 1.0 : b == 1 ? 1.0 : b == 2 ? s : b == 3 ? 2.0 / 3.0 : b == 4 ? s : b == 5 ? p2
                      : b == 6 ? 2.0 / 5.0 : b == 7 ? t23 : b == 8 ? p2 : p3)
       : side == 1 ? (b == 0 ? 1.0 : b == 1 ? -1.0 : b == 2 ? s : b == 3 ? 2.0 / 3.0 : b == 4 ? -s : b == 5 ? p2
                      : b == 6 ? -2.0 / 5.0 : b == 7 ? t23 : b == 8 ? -p2 : p3)
       : side == 2 ? (b == 0 ? 1.0 : b == 1 ? s : b == 2 ? 1.0 : b == 3 ? p2 : b == 4 ? s : b == 5 ? 2.0 / 3.0
                      : b == 6 ? p3 : b == 7 ? p2 : b == 8 ? t23 : 2.0 / 5.0)
                   : (b == 0 ? 1.0 : b == 1 ? s : b == 2 ? -1.0 : b == 3 ? p2 : b == 4 ? -s : b == 5 ? 2.0 / 3.0
                      : b == 6 ? p3 : b == 7 ? -p2 : b == 8 ? t23 : -2.0 / 5.0);
}
// phiTr(side, b, s) written in the basis {1, s, P2(s), P3(s)}: coefficient k
__host__ __device__ constexpr double trk(int side, int b, int k) {
  const double t23 = 2.0 / 3.0, t25 = 2.0 / 5.0;
  const double sg = (side == 0 || side == 2) ? 1.0 : -1.0;   // the face's normal coordinate, +-1
  if (side < 2) {   // x faces: normal xi = sg, tangential s = eta
    switch (b) {
      case 0: return k == 0 ? 1.0 : 0.0;
      case 1: return k == 0 ? sg : 0.0;
      case 2: return k == 1 ? 1.0 : 0.0;
      case 3: return k == 0 ? t23 : 0.0;
      case 4: return k == 1 ? sg : 0.0;
      case 5: return k == 2 ? 1.0 : 0.0;
      case 6: return k == 0 ? sg * t25 : 0.0;
      case 7: return k == 1 ? t23 : 0.0;
      case 8: return k == 2 ? sg : 0.0;
      default: return k == 3 ? 1.0 : 0.0;
    }
  }
  switch (b) {      // y faces: normal eta = sg, tangential s = xi
    case 0: return k == 0 ? 1.0 : 0.0;
    case 1: return k == 1 ? 1.0 : 0.0;
    case 2: return k == 0 ? sg : 0.0;
    case 3: return k == 2 ? 1.0 : 0.0;
    case 4: return k == 1 ? sg : 0.0;
    case 5: return k == 0 ? t23 : 0.0;
    case 6: return k == 3 ? 1.0 : 0.0;
    case 7: return k == 2 ? sg : 0.0;
    case 8: return k == 1 ? t23 : 0.0;
    default: return k == 0 ? sg * t25 : 0.0;
  }
}

// vertex values phiRU/LU/RD/LD (get_basis.f90:5-47); v: 0 RU, 1 LU, 2 RD, 3 LD
__host__ __device__ constexpr double phiV(int v, int b) {
  const double t23 = 2.0 / 3.0, t25 = 2.0 / 5.0;
  const double sx = (v == 0 || v == 2) ? 1.0 : -1.0;   // xi = +-1
  const double sy = (v == 0 || v == 1) ? 1.0 : -1.0;   // eta = +-1
  return b == 0 ? 1.0 : b == 1 ? sx : b == 2 ? sy : b == 3 ? t23 : b == 4 ? sx * sy : b == 5 ? t23
       : b == 6 ? sx * t25 : b == 7 ? sy * t23 : b == 8 ? sx * t23 : sy * t25;
}

// ----------------------------------------------------------------------
// device layout: 52 active (mode, variable) planes, each an (Nx+2) x (Ny+2)
// array with halos, row pitch P (multiple of 16 doubles), interior i = 1
// starting 128-byte aligned.
constexpr int NPL = 52;
constexpr int OFF = 15;

// plane of (mode b = 0..9, variable n = 0..7); -1 if the slot is always zero
// (active mask of Lh.f90:59)
__host__ __device__ constexpr int PL(int b, int n) {
  return b < 6 ? n * 6 + b
       : (n == 5 && b == 6) ? 48
       : (n == 5 && b == 8) ? 49
       : (n == 6 && b == 7) ? 50
       : (n == 6 && b == 9) ? 51
                            : -1;
}

// Storage of a state buffer.  Plain form: the 52 planes in FP64.  Mixed
// form: the 22 planes of the cell averages and of both magnetic-field
// components in FP64 (slots 0..21), then the 30 higher-mode planes of the
// other six variables in FP32 (296 instead of 416 bytes per cell).  The
// cell averages and every B plane stay FP64 because the averages carry the
// conservation and the B planes the discrete divergence.
constexpr int NPL_D = 22, NPL_F = 30;
__host__ __device__ constexpr bool pl_f64(int p) { return (p < 48 && p % 6 == 0) || (p >= 30 && p < 42) || p >= 48; }
__host__ __device__ constexpr int slot_d(int p) {
  return p < 30 ? p / 6 : p < 42 ? 5 + (p - 30) : p == 42 ? 17 : 18 + (p - 48);
}
__host__ __device__ constexpr int slot_f(int p) { return p < 30 ? (p / 6) * 5 + (p % 6 - 1) : 25 + (p - 43); }
// the planes of Bx and By (all their modes)
__host__ __device__ constexpr bool is_bplane(int p) { return (p >= 30 && p < 42) || p >= 48; }

struct Geo {
  int Nx, Ny, P;
  int bcR, bcL, bcU, bcD;
  int pad;
  long long plane;               // doubles per plane
  double hx, hy, hx1, hy1;
  const int* gate;               // inside mhd2d_step: the step's active flag (nullptr elsewhere)
};
__host__ __device__ __forceinline__ long long cix(const Geo& g, int i, int j) {
  return (long long)j * g.P + OFF + i;
}

// error-free transformations and double-double arithmetic (Dekker; Knuth's
// two-sum); the FMA gives the exact product error.  Contraction stays off:
// a product fused into a later sum would break the error terms.
struct dd {
  double hi, lo;
};
__host__ __device__ __forceinline__ dd two_sum(double a, double b) {
  MHD_NOCONTRACT
  const double s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ __forceinline__ dd fast_two_sum(double a, double b) {   // |a| >= |b|
  MHD_NOCONTRACT
  const double s = a + b;
  return {s, b - (s - a)};
}
__host__ __device__ __forceinline__ dd dd_add(dd a, dd b) {
  MHD_NOCONTRACT
  const dd s = two_sum(a.hi, b.hi);
  return fast_two_sum(s.hi, s.lo + (a.lo + b.lo));
}
__host__ __device__ __forceinline__ dd dd_scale(double c, dd a) {   // c * a
  MHD_NOCONTRACT
  const double p = c * a.hi;
  const double e = fma(c, a.hi, -p);
  return fast_two_sum(p, fma(c, a.lo, e));
}

// the one k with trk(side, b, k) != 0
__host__ __device__ constexpr int trace_k(int side, int b) {
  return trk(side, b, 0) != 0.0 ? 0 : trk(side, b, 1) != 0.0 ? 1 : trk(side, b, 2) != 0.0 ? 2 : 3;
}

// view of a state buffer (MIX: the mixed-form storage above); p is the plane
// of PL(), c the cell index of cix()
template <bool MIX>
struct SRd {
  const double* b;
  long long pl;
  __device__ __forceinline__ double operator()(int p, long long c) const {
    if (!MIX) return b[(long long)p * pl + c];
    if (pl_f64(p)) return b[(long long)slot_d(p) * pl + c];
    return (double)reinterpret_cast<const float*>(b + NPL_D * pl)[(long long)slot_f(p) * pl + c];
  }
};
// the same over two buffers of that layout: the Bx, By planes from bq (when
// set), the others from b.  The fused stages' unlimited stage value is read
// so: its planes other than Bx, By share the limited value's storage.
template <bool MIX>
struct SRdB {
  const double* b;
  long long pl;
  const double* bq;
  __device__ __forceinline__ double operator()(int p, long long c) const {
    return SRd<MIX>{(bq && is_bplane(p)) ? bq : b, pl}(p, c);
  }
};
template <bool MIX>
struct SWr {
  double* b;
  long long pl;
  __device__ __forceinline__ double operator()(int p, long long c) const { return SRd<MIX>{b, pl}(p, c); }
  __device__ __forceinline__ void set(int p, long long c, double v) const {
    if (!MIX) b[(long long)p * pl + c] = v;
    else if (pl_f64(p)) b[(long long)slot_d(p) * pl + c] = v;
    else reinterpret_cast<float*>(b + NPL_D * pl)[(long long)slot_f(p) * pl + c] = (float)v;
  }
  __device__ __forceinline__ operator SRd<MIX>() const { return SRd<MIX>{b, pl}; }
};
// f(integral_constant<int, I>) for I = B..E-1: a loop whose index is a
// compile-time constant in every iteration (so the storage-map branches of
// SRd/SWr fold away)
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}
// bytes of one state buffer
__host__ __device__ constexpr long long state_bytes(bool mix, long long plane) {
  return mix ? (NPL_D * 8LL + NPL_F * 4LL) * plane : 52LL * 8 * plane;
}

}  // namespace mhd
