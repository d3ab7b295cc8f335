// mhd2d_faces.hip -- face and vertex kernels of one RK stage: traces at the
// face Gauss points, physical fluxes, the 1D Riemann solvers, face EMFs
// (Lh.f90:72-264, LhEz.f90:6-28) and the corner EMF (LhEz.f90:33-69).
// The mixed form's FP32 face kernels live in mhd2d_faces_mix.hip.
#include "mhd2d_internal.hpp"

namespace mhd2d_impl {
namespace {

// Plain form: one lane per (face, Gauss point); the trace at the point is
// summed over the modes in the reference's order (Lh.f90:81-84).  The
// collapsed cubic of the mixed form rounds differently, and the HLLD
// branches of magnetically dominated states amplify that past 1e-12.
// x faces:
template <int FT, typename T>
__global__ __launch_bounds__(BS) void k_xface_pt(SRd<false> U, T* __restrict__ FX,
                                              T* __restrict__ EZ, Geo g, const DevScal* __restrict__ sc) {
  if (step_off(g)) return;
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
  if (step_off(g)) return;
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
  if (step_off(g)) return;
  const long long t = xcd_block() * BS + threadIdx.x;
  const long long nv = (long long)(g.Nx + 1) * (g.Ny + 1);
  if (t >= nv) return;
  const int i = (int)(t % (g.Nx + 1)), j = (int)(t / (g.Nx + 1));
  vertex_at<FT>(U, EZV, g, i, j);
}


}  // namespace


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
        launch_xface_mixed(ctx, U, FT);
      else
        hipLaunchKernelGGL((k_xface_pt<FT, double>), dim3(nblk(5 * nxf)), dim3(BS), 0, ctx->stream,
                           SRd<false>{U, g.plane}, ctx->FX,
                           ctx->EZX, g, sc);
    }
    {
      Timed tm(ctx, KC_YFACE);
      if (mixed)
        launch_yface_mixed(ctx, U, FT);
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
