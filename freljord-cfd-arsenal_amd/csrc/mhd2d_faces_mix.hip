// mhd2d_faces_mix.hip -- the mixed form's FP32 face kernels (traces, physical
// fluxes, Riemann solvers and face EMFs; Lh.f90:72-264, LhEz.f90:6-28).
// This file is compiled with reciprocal-based FP32 division (x / y as
// x * rcp(y)); it holds no FP64 arithmetic of the plain form.
#include "mhd2d_internal.hpp"

namespace mhd2d_impl {
namespace {

// ---------------------------------------------------------------------------
// Face kernels: one lane per face; the lane runs the five Gauss points in
// turn.  Mixed form: what the cell and CT kernels need of a face is the
// Gauss sums of its fluxes against the face Legendre polynomials,
// M_k = sum_q w_q F(s_q) L_k(s_q), L = {1, s, P2(s), P3(s)} (every trace of a
// mode of G/get_basis.f90 is one L_k times a constant), so the face kernel
// stores those moments (layout in mhd2d_internal.hpp: FM_*) instead of the
// five point values: 25 flux and 2 EMF values per face instead of 40 and 5.
// Each side's trace along the face is collapsed once into a
// cubic in the tangential coordinate s,
//   u(s) = c0 + c1 s + c2 P2(s) + c3 P3(s)   (per variable),
// from the cell's modes and the reference's trace tables (phiGR/GL/GU/GD,
// get_basis.f90), so the 52 modes of the two cells are read once per face.
// The plain form keeps the reference's per-point sums (k_xface_pt /
// k_yface_pt below).
template <int SIDE, typename T, class SV>
__device__ __forceinline__ void trace_coef(const SV& U, long long c, T cf[8][4]) {
#pragma unroll
  for (int n = 0; n < 8; n++) {
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
#pragma unroll
    for (int b = 0; b < 10; b++) {
      const int p = PL(b, n);
      if (p < 0) continue;
      const T m = T(U(p, c));
      if (trk(SIDE, b, 0) != 0.0) a0 += m * T(trk(SIDE, b, 0));
      if (trk(SIDE, b, 1) != 0.0) a1 += m * T(trk(SIDE, b, 1));
      if (trk(SIDE, b, 2) != 0.0) a2 += m * T(trk(SIDE, b, 2));
      if (trk(SIDE, b, 3) != 0.0) a3 += m * T(trk(SIDE, b, 3));
    }
    cf[n][0] = a0; cf[n][1] = a1; cf[n][2] = a2; cf[n][3] = a3;
  }
}

template <typename T>
__device__ __forceinline__ void trace_eval(const T cf[8][4], T s, T p2, T p3, T u[8]) {
#pragma unroll
  for (int n = 0; n < 8; n++) u[n] = cf[n][0] + cf[n][1] * s + cf[n][2] * p2 + cf[n][3] * p3;
}

// x faces: Lh.f90:72-136 (traces, fluxes), :206-234 (Riemann),
// LhEz.f90:6-16 (face EMF).  Face i (0..Nx) of row j (1..Ny) lies between
// cells i and i+1.
template <int FT, typename T>
__device__ __forceinline__ void xface_at(const SRd<true>& U, T* __restrict__ FX, T* __restrict__ EZ,
                                         const Geo& g, int i, int j) {
  const long long o = cix(g, i, j);
  T cL[8][4], cR[8][4];
  trace_coef<0, T>(U, o, cL);                 // right trace of cell i
  trace_coef<1, T>(U, cix(g, i + 1, j), cR);  // left trace of cell i+1
  T m[8][3] = {}, m3 = T(0), e0 = T(0), e1 = T(0);
#pragma unroll 1
  for (int q = 0; q < 5; q++) {
    constexpr double L0 = lam(0), L1 = lam(1), L2 = lam(2), L3 = lam(3), L4 = lam(4);
    const T sq = T(sel5(q, L0, L1, L2, L3, L4));
    const T p2 = T(sel5(q, P2(L0), P2(L1), P2(L2), P2(L3), P2(L4)));
    const T p3 = T(sel5(q, P3(L0), P3(L1), P3(L2), P3(L3), P3(L4)));
    T UL[8], UR[8], FL[8], FR[8], F[8];
    trace_eval<T>(cL, sq, p2, p3, UL);
    trace_eval<T>(cR, sq, p2, p3, UR);
    lh_flux<T, 1>(UL, FL);
    lh_flux<T, 1>(UR, FR);
    T a1, b1, a2, b2;
    eigenvalueMm<T, 1>(UL, a1, b1);
    eigenvalueMm<T, 1>(UR, a2, b2);
    const T SR = tmax(a1, a2), SL = tmin(b1, b2);
    riemann<T, 1, FT>(SR, SL, UL, UR, FL, FR, F);
    const T S = tmax(fabs(UL[1] / UL[0]), fabs(UR[1] / UR[0]));
    const T ez = T(-0.5) * (FL[6] + FR[6] - S * (UR[6] - UL[6]));
    const T w = T(sel5(q, wgt(0), wgt(1), wgt(2), wgt(3), wgt(4)));
    const T ws = w * sq, wp2 = w * p2;
#pragma unroll
    for (int n = 0; n < 8; n++) {
      m[n][0] += w * F[n];
      m[n][1] += ws * F[n];
      m[n][2] += wp2 * F[n];
    }
    m3 += (w * p3) * F[6];           // By on x faces: mode 10 = P3(eta)
    e0 += w * ez;
    e1 += ws * ez;
  }
#pragma unroll
  for (int n = 0; n < 8; n++)
#pragma unroll
    for (int k = 0; k < 3; k++) FX[(FM_N * n + k) * g.plane + o] = m[n][k];
  FX[FM_P3 * g.plane + o] = m3;
  EZ[o] = e0;
  EZ[g.plane + o] = e1;
}
template <int FT, typename T>
__global__ __launch_bounds__(BS) void k_xface(SRd<true> U, T* __restrict__ FX,
                                              T* __restrict__ EZ, Geo g, const DevScal* __restrict__ sc) {
  if (step_off(g)) return;
  const long long f = xcd_block() * BS + threadIdx.x;
  const long long nf = (long long)(g.Nx + 1) * g.Ny;
  if (f >= nf) return;
  const int i = (int)(f % (g.Nx + 1)), j = 1 + (int)(f / (g.Nx + 1));
  xface_at<FT, T>(U, FX, EZ, g, i, j);
}

// y faces: Lh.f90:139-203, :237-264; LhEz.f90:18-28.  Face j (0..Ny) of
// column i (1..Nx) lies between cells j and j+1.
template <int FT, typename T>
__device__ __forceinline__ void yface_at(const SRd<true>& U, T* __restrict__ FY, T* __restrict__ EZ,
                                         const Geo& g, int i, int j) {
  const long long o = cix(g, i, j);
  T cD[8][4], cU[8][4];
  trace_coef<2, T>(U, o, cD);                 // upper trace of cell j (reference UU)
  trace_coef<3, T>(U, cix(g, i, j + 1), cU);  // lower trace of cell j+1 (reference UD)
  T m[8][3] = {}, m3 = T(0), e0 = T(0), e1 = T(0);
#pragma unroll 1
  for (int q = 0; q < 5; q++) {
    constexpr double L0 = lam(0), L1 = lam(1), L2 = lam(2), L3 = lam(3), L4 = lam(4);
    const T sq = T(sel5(q, L0, L1, L2, L3, L4));
    const T p2 = T(sel5(q, P2(L0), P2(L1), P2(L2), P2(L3), P2(L4)));
    const T p3 = T(sel5(q, P3(L0), P3(L1), P3(L2), P3(L3), P3(L4)));
    T UD[8], UU[8], FD[8], FU[8], F[8];
    trace_eval<T>(cD, sq, p2, p3, UD);
    trace_eval<T>(cU, sq, p2, p3, UU);
    lh_flux<T, 2, false>(UD, FD);
    lh_flux<T, 2, true>(UU, FU);
    T a1, b1, a2, b2;
    eigenvalueMm<T, 2>(UD, a1, b1);
    eigenvalueMm<T, 2>(UU, a2, b2);
    const T SR = tmax(a1, a2), SL = tmin(b1, b2);
    riemann<T, 2, FT>(SR, SL, UD, UU, FD, FU, F);
    const T S = tmax(fabs(UD[2] / UD[0]), fabs(UU[2] / UU[0]));
    const T ez = T(0.5) * (FD[5] + FU[5] - S * (UU[5] - UD[5]));
    const T w = T(sel5(q, wgt(0), wgt(1), wgt(2), wgt(3), wgt(4)));
    const T ws = w * sq, wp2 = w * p2;
#pragma unroll
    for (int n = 0; n < 8; n++) {
      m[n][0] += w * F[n];
      m[n][1] += ws * F[n];
      m[n][2] += wp2 * F[n];
    }
    m3 += (w * p3) * F[5];           // Bx on y faces: mode 7 = P3(xi)
    e0 += w * ez;
    e1 += ws * ez;
  }
#pragma unroll
  for (int n = 0; n < 8; n++)
#pragma unroll
    for (int k = 0; k < 3; k++) FY[(FM_N * n + k) * g.plane + o] = m[n][k];
  FY[FM_P3 * g.plane + o] = m3;
  EZ[o] = e0;
  EZ[g.plane + o] = e1;
}
template <int FT, typename T>
__global__ __launch_bounds__(BS) void k_yface(SRd<true> U, T* __restrict__ FY,
                                              T* __restrict__ EZ, Geo g, const DevScal* __restrict__ sc) {
  if (step_off(g)) return;
  const long long f = xcd_block() * BS + threadIdx.x;
  const long long nf = (long long)g.Nx * (g.Ny + 1);
  if (f >= nf) return;
  const int i = 1 + (int)(f % g.Nx), j = (int)(f / g.Nx);
  yface_at<FT, T>(U, FY, EZ, g, i, j);
}

}  // namespace

int launch_xface_mixed(mhd2d_ctx* ctx, const double* U, int ft) {
  const Geo& g = ctx->g;
  const long long nxf = (long long)(g.Nx + 1) * g.Ny;
  return dispatch_ft(ft, [&](auto FTc) -> int {
    constexpr int FT = decltype(FTc)::value;
    hipLaunchKernelGGL((k_xface<FT, float>), dim3(nblk(nxf)), dim3(BS), 0, ctx->stream, SRd<true>{U, g.plane},
                       (float*)ctx->FX, (float*)ctx->EZX, g, nullptr);
    return 0;
  });
}

int launch_yface_mixed(mhd2d_ctx* ctx, const double* U, int ft) {
  const Geo& g = ctx->g;
  const long long nyf = (long long)g.Nx * (g.Ny + 1);
  return dispatch_ft(ft, [&](auto FTc) -> int {
    constexpr int FT = decltype(FTc)::value;
    hipLaunchKernelGGL((k_yface<FT, float>), dim3(nblk(nyf)), dim3(BS), 0, ctx->stream, SRd<true>{U, g.plane},
                       (float*)ctx->FY, (float*)ctx->EZY, g, nullptr);
    return 0;
  });
}

}  // namespace mhd2d_impl
