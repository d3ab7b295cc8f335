// mhd2d_faces.hip -- face and vertex kernels of one RK stage: traces at the
// face Gauss points, physical fluxes, the 1D Riemann solvers, face EMFs
// (Lh.f90:72-264, LhEz.f90:6-28) and the corner EMF (LhEz.f90:33-69).
#include "mhd2d_internal.hpp"

namespace mhd2d_impl {
namespace {

// ---------------------------------------------------------------------------
// Face kernels: one lane per face; the lane runs the five Gauss points in
// turn and writes each point's flux as one coalesced row.
// Mixed form: each side's trace along the face is collapsed once into a
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
#pragma unroll
    for (int n = 0; n < 8; n++) FX[(q * 8 + n) * g.plane + o] = F[n];
    const T S = tmax(fabs(UL[1] / UL[0]), fabs(UR[1] / UR[0]));
    EZ[q * g.plane + o] = T(-0.5) * (FL[6] + FR[6] - S * (UR[6] - UL[6]));
  }
}
template <int FT, typename T>
__global__ __launch_bounds__(BS) void k_xface(SRd<true> U, T* __restrict__ FX,
                                              T* __restrict__ EZ, Geo g, const DevScal* __restrict__ sc) {
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
#pragma unroll
    for (int n = 0; n < 8; n++) FY[(q * 8 + n) * g.plane + o] = F[n];
    const T S = tmax(fabs(UD[2] / UD[0]), fabs(UU[2] / UU[0]));
    EZ[q * g.plane + o] = T(0.5) * (FD[5] + FU[5] - S * (UU[5] - UD[5]));
  }
}
template <int FT, typename T>
__global__ __launch_bounds__(BS) void k_yface(SRd<true> U, T* __restrict__ FY,
                                              T* __restrict__ EZ, Geo g, const DevScal* __restrict__ sc) {
  const long long f = xcd_block() * BS + threadIdx.x;
  const long long nf = (long long)g.Nx * (g.Ny + 1);
  if (f >= nf) return;
  const int i = 1 + (int)(f % g.Nx), j = (int)(f / g.Nx);
  yface_at<FT, T>(U, FY, EZ, g, i, j);
}

// Plain form: one lane per (face, Gauss point); the trace at the point is
// summed over the modes in the reference's order (Lh.f90:81-84).  The
// collapsed cubic of the mixed form rounds differently, and the HLLD
// branches of magnetically dominated states amplify that past 1e-12.
// x faces:
template <int FT, typename T>
__global__ __launch_bounds__(BS) void k_xface_pt(SRd<false> U, T* __restrict__ FX,
                                              T* __restrict__ EZ, Geo g, const DevScal* __restrict__ sc) {
  const long long t = xcd_block() * BS + threadIdx.x;
  const long long nf = (long long)(g.Nx + 1) * g.Ny;
  if (t >= nf * 5) return;
  const int q = (int)(t % 5);
  const long long f = t / 5;
  const int i = (int)(f % (g.Nx + 1)), j = 1 + (int)(f / (g.Nx + 1));
  T UL[8], UR[8], FL[8], FR[8], F[8];
  trace_at<0, T>(U, cix(g, i, j), q, UL);       // right trace of cell i
  trace_at<1, T>(U, cix(g, i + 1, j), q, UR);   // left trace of cell i+1
  lh_flux<T, 1>(UL, FL);
  lh_flux<T, 1>(UR, FR);
  T a1, b1, a2, b2;
  eigenvalueMm<T, 1>(UL, a1, b1);
  eigenvalueMm<T, 1>(UR, a2, b2);
  const T SR = tmax(a1, a2), SL = tmin(b1, b2);
  riemann<T, 1, FT>(SR, SL, UL, UR, FL, FR, F);
  const long long o = cix(g, i, j);
#pragma unroll
  for (int n = 0; n < 8; n++) FX[(q * 8 + n) * g.plane + o] = F[n];
  const T S = tmax(fabs(UL[1] / UL[0]), fabs(UR[1] / UR[0]));
  EZ[q * g.plane + o] = T(-0.5) * (FL[6] + FR[6] - S * (UR[6] - UL[6]));
}

// y faces, plain form, one lane per (face, Gauss point)
template <int FT, typename T>
__global__ __launch_bounds__(BS) void k_yface_pt(SRd<false> U, T* __restrict__ FY,
                                              T* __restrict__ EZ, Geo g, const DevScal* __restrict__ sc) {
  const long long t = xcd_block() * BS + threadIdx.x;
  const long long nf = (long long)g.Nx * (g.Ny + 1);
  if (t >= nf * 5) return;
  const int q = (int)(t % 5);
  const long long f = t / 5;
  const int i = 1 + (int)(f % g.Nx), j = (int)(f / g.Nx);
  T UD[8], UU[8], FD[8], FU[8], F[8];
  trace_at<2, T>(U, cix(g, i, j), q, UD);       // upper trace of cell j (reference UU)
  trace_at<3, T>(U, cix(g, i, j + 1), q, UU);   // lower trace of cell j+1 (reference UD)
  lh_flux<T, 2, false>(UD, FD);
  lh_flux<T, 2, true>(UU, FU);
  T a1, b1, a2, b2;
  eigenvalueMm<T, 2>(UD, a1, b1);
  eigenvalueMm<T, 2>(UU, a2, b2);
  const T SR = tmax(a1, a2), SL = tmin(b1, b2);
  riemann<T, 2, FT>(SR, SL, UD, UU, FD, FU, F);
  const long long o = cix(g, i, j);
#pragma unroll
  for (int n = 0; n < 8; n++) FY[(q * 8 + n) * g.plane + o] = F[n];
  const T S = tmax(fabs(UD[2] / UD[0]), fabs(UU[2] / UU[0]));
  EZ[q * g.plane + o] = T(0.5) * (FD[5] + FU[5] - S * (UU[5] - UD[5]));
}

// vertex (corner) EMF: LhEz.f90:33-69.  Vertex (i, j) is the upper-right
// corner of cell (i, j), i = 0..Nx, j = 0..Ny.
template <int V, class SV>
__device__ __forceinline__ void vertex_state(const SV& U, long long c, double out[8]) {
#pragma unroll
  for (int n = 0; n < 8; n++) {
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 10; b++) {
      const int p = PL(b, n);
      if (p >= 0) acc += U(p, c) * phiV(V, b);
    }
    out[n] = acc;
  }
}

template <int FT, bool MIX>
__device__ __forceinline__ void vertex_at(const SRd<MIX>& U, double* __restrict__ EZV, const Geo& g, int i, int j) {
  double URU[8], ULU[8], URD[8], ULD[8];
  vertex_state<3>(U, cix(g, i + 1, j + 1), URU);  // LD corner of the upper-right cell
  vertex_state<2>(U, cix(g, i, j + 1), ULU);      // RD corner of the upper-left cell
  vertex_state<1>(U, cix(g, i + 1, j), URD);      // LU corner of the lower-right cell
  vertex_state<0>(U, cix(g, i, j), ULD);          // RU corner of the lower-left cell
  EZV[cix(g, i, j)] = vertex_emf<double, FT>(URU, ULU, URD, ULD);
}
template <int FT, bool MIX>
__global__ __launch_bounds__(BS) void k_vertex(SRd<MIX> U, double* __restrict__ EZV, Geo g,
                                               const DevScal* __restrict__ sc) {
  const long long t = xcd_block() * BS + threadIdx.x;
  const long long nv = (long long)(g.Nx + 1) * (g.Ny + 1);
  if (t >= nv) return;
  const int i = (int)(t % (g.Nx + 1)), j = (int)(t / (g.Nx + 1));
  vertex_at<FT>(U, EZV, g, i, j);
}


}  // namespace

template <typename F>
int dispatch_ft(int ft, F&& f) {
  switch (ft) {
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    default: return f(std::integral_constant<int, 4>{});
  }
}

// face fluxes + face EMFs (+ vertex EMFs when with_vertex); the face
// quadrature runs in double (plain form) or float (mixed form)
int launch_fluxes(mhd2d_ctx* ctx, const double* U, bool with_vertex, const DevScal* sc) {
  const Geo& g = ctx->g;
  const bool mixed = ctx->cfg.precision == MHD2D_MIXED;
  return dispatch_ft(ctx->cfg.flux_type, [&](auto FTc) -> int {
    constexpr int FT = decltype(FTc)::value;
    const long long nxf = (long long)(g.Nx + 1) * g.Ny, nyf = (long long)g.Nx * (g.Ny + 1);
    {
      Timed tm(ctx, KC_XFACE);
      if (mixed)
        hipLaunchKernelGGL((k_xface<FT, float>), dim3(nblk(nxf)), dim3(BS), 0, ctx->stream, SRd<true>{U, g.plane},
                           (float*)ctx->FX,
                           (float*)ctx->EZX, g, sc);
      else
        hipLaunchKernelGGL((k_xface_pt<FT, double>), dim3(nblk(5 * nxf)), dim3(BS), 0, ctx->stream,
                           SRd<false>{U, g.plane}, ctx->FX,
                           ctx->EZX, g, sc);
    }
    {
      Timed tm(ctx, KC_YFACE);
      if (mixed)
        hipLaunchKernelGGL((k_yface<FT, float>), dim3(nblk(nyf)), dim3(BS), 0, ctx->stream, SRd<true>{U, g.plane},
                           (float*)ctx->FY,
                           (float*)ctx->EZY, g, sc);
      else
        hipLaunchKernelGGL((k_yface_pt<FT, double>), dim3(nblk(5 * nyf)), dim3(BS), 0, ctx->stream,
                           SRd<false>{U, g.plane}, ctx->FY,
                           ctx->EZY, g, sc);
    }
    if (with_vertex) {
      Timed tm(ctx, KC_VERTEX);
      const long long nv = (long long)(g.Nx + 1) * (g.Ny + 1);
      if (mixed)
        hipLaunchKernelGGL((k_vertex<FT, true>), dim3(nblk(nv)), dim3(BS), 0, ctx->stream, SRd<true>{U, g.plane},
                           ctx->EZV, g, sc);
      else
        hipLaunchKernelGGL((k_vertex<FT, false>), dim3(nblk(nv)), dim3(BS), 0, ctx->stream, SRd<false>{U, g.plane},
                           ctx->EZV, g, sc);
    }
    LAUNCHCHK(ctx);
    return 0;
  });
}

// vertex EMFs alone (LhEz after Lh in the per-subroutine API)
int launch_vertex(mhd2d_ctx* ctx, const double* U, const DevScal* sc) {
  const Geo& g = ctx->g;
  return dispatch_ft(ctx->cfg.flux_type, [&](auto FTc) -> int {
    constexpr int FT = decltype(FTc)::value;
    const long long nv = (long long)(g.Nx + 1) * (g.Ny + 1);
    Timed tm(ctx, KC_VERTEX);
    if (ctx->cfg.precision == MHD2D_MIXED)
      hipLaunchKernelGGL((k_vertex<FT, true>), dim3(nblk(nv)), dim3(BS), 0, ctx->stream, SRd<true>{U, g.plane},
                         ctx->EZV, g, sc);
    else
      hipLaunchKernelGGL((k_vertex<FT, false>), dim3(nblk(nv)), dim3(BS), 0, ctx->stream, SRd<false>{U, g.plane},
                         ctx->EZV, g, sc);
    LAUNCHCHK(ctx);
    return 0;
  });
}

}  // namespace mhd2d_impl
