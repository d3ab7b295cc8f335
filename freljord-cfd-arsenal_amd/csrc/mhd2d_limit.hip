// mhd2d_limit.hip -- the limiter pass (TVB_Limiter_P2 with the optional
// positivity limiter and the Balsara divergence-free reconstruction fused)
// and the KXRCF troubled-cell detector.
#include "mhd2d_internal.hpp"
#include "mhd2d_detect.hpp"
#include "mhd2d_limit_cell.hpp"

namespace mhd2d_impl {
namespace {



// limiter of one cell (TVB_Limiter_P2.f90:16-121 + pp + div_free_Balsara):
// interior cells are limited, halo cells copied (or, with DIV, their B
// planes zeroed as div_free_Balsara.f90:13 does)
template <bool APPLY, bool PP, bool DIV, bool MIX, class IV>
__device__ __forceinline__ void limit_cell(const IV& in, const SWr<MIX>& out, const double* __restrict__ F,
                                           const int* __restrict__ tmask, const double* __restrict__ trec,
                                           const Geo& g, int i, int j) {
  const long long c = cix(g, i, j);
  const bool interior = i >= 1 && i <= g.Nx && j >= 1 && j <= g.Ny;
  if (!interior) {
#pragma unroll 4
    for (int p = 0; p < NPL; p++) out.set(p, c, (DIV && is_bplane(p)) ? 0.0 : in(p, c));
    return;
  }
  double m[NPL];
#pragma unroll
  for (int p = 0; p < NPL; p++) m[p] = in(p, c);
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
  for (int p = 0; p < NPL; p++) out.set(p, c, m[p]);
}

// limiter pass over the whole array
template <bool APPLY, bool PP, bool DIV, bool MIX>
__global__ __launch_bounds__(BS) void k_limit(SRd<MIX> in, SWr<MIX> out,
                                              const double* __restrict__ F, const int* __restrict__ tmask,
                                              const double* __restrict__ trec, Geo g) {
  if (step_off(g)) return;
  const long long t = xcd_block() * BS + threadIdx.x;
  const int W = g.Nx + 2;
  if (t >= (long long)W * (g.Ny + 2)) return;
  limit_cell<APPLY, PP, DIV, MIX>(in, out, F, tmask, trec, g, (int)(t % W), (int)(t / W));
}

// The fused stages store the stage value's planes other than Bx, By once,
// in the limited value (out); its companion (q) holds the Bx, By planes.
// Before a list kernel limits interior cell tc in place, the cell's other
// unlimited planes are copied to the companion and the cell is marked
// with this stage's tag 4 epoch + stage (the next stage's cell kernel then
// reads its RK input there); the companion then holds the whole unlimited
// value of the cell.
template <bool MIX>
__device__ __forceinline__ void lean_take(const SWr<MIX>& q, const SWr<MIX>& out, int* __restrict__ lflag, int tag,
                                          const Geo& g, int tc) {
  if (lflag[tc] == tag) return;   // a list kernel of this stage took it already
  const long long c = cix(g, 1 + tc % g.Nx, 1 + tc / g.Nx);
#pragma unroll
  for (int p = 0; p < NPL; p++)
    if (!is_bplane(p)) q.set(p, c, out(p, c));
  lflag[tc] = tag;
}

// the cells of the flagged list whose TVB changed something: the cell
// kernel's fused epilogue limited every cell without TVB (pp and the
// divergence-free reconstruction), so these are limited again from the
// unlimited stage value, with TVB (a fixed grid strides the list).  LEAN:
// the fused stages' shared planes (in is the companion, lean_take)
template <bool PP, bool MIX, bool LEAN>
__global__ __launch_bounds__(BS) void k_limit_list(SWr<MIX> in, SWr<MIX> out, const double* __restrict__ F,
                                                   const int* __restrict__ list, const int* __restrict__ tmask,
                                                   const double* __restrict__ trec, Geo g,
                                                   const DevScal* __restrict__ lst, int* __restrict__ lflag,
                                                   int stage) {
  if (step_off(g)) return;
  const unsigned int n = lst->nlist;
  const int tag = (int)(4 * lst->epoch + stage);
  for (unsigned int k = blockIdx.x * BS + threadIdx.x; k < n; k += gridDim.x * BS) {
    const int tc = list[k];
    if (!tmask[tc]) continue;
    if (LEAN) lean_take(in, out, lflag, tag, g, tc);
    limit_cell<true, PP, true, MIX>(SRd<MIX>(in), out, F, tmask, trec, g, 1 + tc % g.Nx, 1 + tc / g.Nx);
  }
}

// the cells the fused cell kernel's epilogue left to the positivity limiter
// (mixed form: those its bound cannot settle; plain form: those positivity
// changes): the full limiter pass without TVB (positivity, then the
// divergence-free reconstruction) from the unlimited stage value (the
// companion q, after lean_take), in place; the cells TVB changes are limited
// again afterwards (k_limit_list)
template <bool MIX>
__global__ __launch_bounds__(BS) void k_pp_list(SWr<MIX> q, SWr<MIX> out, const double* __restrict__ F,
                                                const int* __restrict__ list, Geo g,
                                                const DevScal* __restrict__ lst, int* __restrict__ lflag,
                                                int stage) {
  if (step_off(g)) return;
  const unsigned int n = lst->npp;
  const int tag = (int)(4 * lst->epoch + stage);
  for (unsigned int k = blockIdx.x * BS + threadIdx.x; k < n; k += gridDim.x * BS) {
    const int tc = list[k];
    lean_take(q, out, lflag, tag, g, tc);
    limit_cell<false, true, true, MIX>(SRd<MIX>(q), out, F, nullptr, nullptr, g, 1 + tc % g.Nx, 1 + tc / g.Nx);
  }
}

// div_free_Balsara.f90:13-79 over the whole array, B planes only: the
// in-cell B of every interior cell from its four new faces, the halo B
// zeroed.  The limited B planes depend on the faces alone (positivity and
// TVB changes of Bx, By are overwritten here), so in the mixed form this pass
// runs after the CT update, apart from the cell kernel, whose lean epilogue
// writes the other planes of the limited value
template <bool MIX>
__global__ __launch_bounds__(BS) void k_divfree(SWr<MIX> out, const double* __restrict__ F, Geo g) {
  if (step_off(g)) return;
  const long long t = xcd_block() * BS + threadIdx.x;
  const int W = g.Nx + 2;
  if (t >= (long long)W * (g.Ny + 2)) return;
  const int i = (int)(t % W), j = (int)(t / W);
  const long long c = cix(g, i, j);
  if (i >= 1 && i <= g.Nx && j >= 1 && j <= g.Ny) {
    divfree_apply([&](int p, double v) { out.set(p, c, v); }, F, g, i, j);
  } else {
#pragma unroll
    for (int p = 0; p < NPL; p++)
      if (is_bplane(p)) out.set(p, c, 0.0);
  }
}

template <bool MIX>
__global__ __launch_bounds__(BS) void k_kxrcf(SRd<MIX> U, int* __restrict__ trouble, Geo g,
                                              double Ck, double hk, DevScal* sc) {
  if (step_off(g)) return;
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  if (t >= (long long)g.Nx * g.Ny) return;
  const int i = 1 + (int)(t % g.Nx), j = 1 + (int)(t / g.Nx);
  const bool flag = kxrcf_cell(U, g, i, j, Ck, hk);
  trouble[t] = flag ? 1 : 0;
  if (sc) count_flags(flag, sc);
}

__global__ void k_list_reset(DevScal* lst, const int* gate) {
  if (step_off(gate)) return;
  if (threadIdx.x == 0) lst->nlist = 0u;
}

// KXRCF over the interior cells for the limiter pass: trouble[] and the
// compacted list of flagged cells (one atomic per wave), tmask[] cleared
template <bool MIX>
__global__ __launch_bounds__(BS) void k_detect(SRdB<MIX> U, int* __restrict__ trouble,
                                               int* __restrict__ tmask, int* __restrict__ list, Geo g, double Ck,
                                               double hk, DevScal* sc, DevScal* lst) {
  if (step_off(g)) return;
  const long long t = xcd_block() * BS + threadIdx.x;
  if (t >= (long long)g.Nx * g.Ny) return;
  const int i = 1 + (int)(t % g.Nx), j = 1 + (int)(t / g.Nx);
  const bool flag = kxrcf_cell(U, g, i, j, Ck, hk);
  trouble[t] = flag ? 1 : 0;
  tmask[t] = 0;
  if (sc) count_flags(flag, sc);
  list_append(flag, (int)t, list, lst);
}

// TVB decision and edge values of one cell (TVB_Limiter_P2.f90:24-104):
// returns the change bits per variable; Rm/Lm (x) and Um/Dm (y) are the
// modified edge deviations DeltaUR1mod .. DeltaUD1mod.
template <class SV>
__device__ __forceinline__ int tvb_cell(const SV& in, const Geo& g, int i, int j, double M,
                                        double Rm[8], double Lm[8], double Um[8], double Dm[8]) {
  const long long c = cix(g, i, j);
  int change = 0;
  double U[8];
#pragma unroll
  for (int n = 0; n < 8; n++) U[n] = in(PL(0, n), c);
  {
    double dR[8], dL[8], r1[8], l1[8];
    const long long cE = cix(g, i + 1, j), cW = cix(g, i - 1, j);
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const double m1 = in(PL(1, n), c), m3 = in(PL(3, n), c);
      dR[n] = in(PL(0, n), cE) - U[n];
      dL[n] = U[n] - in(PL(0, n), cW);
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
      const double m2 = in(PL(2, n), c), m5 = in(PL(5, n), c);
      dR[n] = in(PL(0, n), cN) - U[n];
      dL[n] = U[n] - in(PL(0, n), cS);
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
template <int IND, bool MIX>
__global__ __launch_bounds__(BS) void k_tvb(SRd<MIX> in, int* __restrict__ trouble,
                                            int* __restrict__ tmask, double* __restrict__ trec, Geo g, double M,
                                            double Ck, double hk, DevScal* sc) {
  if (step_off(g)) return;
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
template <bool MIX>
__global__ __launch_bounds__(BS) void k_tvb_list(SRdB<MIX> in, const int* __restrict__ list,
                                                 int* __restrict__ tmask, double* __restrict__ trec, Geo g,
                                                 double M, const DevScal* __restrict__ lst) {
  if (step_off(g)) return;
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

int launch_kxrcf(mhd2d_ctx* ctx, const double* U, DevScal* sc) {
  const Geo& g = ctx->g;
  Timed tm(ctx, KC_KXRCF);
  return with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_kxrcf<MIX>), dim3(nblk((long long)g.Nx * g.Ny)), dim3(BS), 0, ctx->stream,
                       SRd<MIX>{U, g.plane}, ctx->trouble, g, ctx->cfg.Ck, kx_hk(g), sc);
    LAUNCHCHK(ctx);
    return 0;
  });
}

// TVB over the compacted list of flagged cells (a fixed grid strides it)
int launch_tvb_list(mhd2d_ctx* ctx, const double* in, const double* inB) {
  const Geo& g = ctx->g;
  const unsigned nb = std::min<unsigned>(nblk((long long)g.Nx * g.Ny), 1024u);
  return with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_tvb_list<MIX>), dim3(nb), dim3(BS), 0, ctx->stream, SRdB<MIX>{in, g.plane, inB},
                       (const int*)ctx->tlist, ctx->tmask, ctx->trec, g, ctx->cfg.M, (const DevScal*)ctx->sc);
    LAUNCHCHK(ctx);
    return 0;
  });
}

// KXRCF with a compacted list, then TVB over the list: trouble[] = KXRCF
// flags, tmask[] = TVB change bits of the flagged cells (0 elsewhere)
int launch_detect_tvb(mhd2d_ctx* ctx, const double* in, DevScal* sc, const double* inB) {
  const Geo& g = ctx->g;
  hipLaunchKernelGGL(k_list_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, ctx->g.gate);
  int rc = with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_detect<MIX>), dim3(nblk((long long)g.Nx * g.Ny)), dim3(BS), 0, ctx->stream,
                       SRdB<MIX>{in, g.plane, inB}, ctx->trouble, ctx->tmask, ctx->tlist, g, ctx->cfg.Ck, kx_hk(g),
                       sc, ctx->sc);
    LAUNCHCHK(ctx);
    return 0;
  });
  if (rc) return rc;
  return launch_tvb_list(ctx, in, inB);
}

// KXRCF flags of every cell into kx[], TVB change bits of every cell into
// tmask[] (indicator statistics)
int launch_tvb_all(mhd2d_ctx* ctx, const double* in, int* kx) {
  const Geo& g = ctx->g;
  const dim3 gt(nblk((long long)g.Nx * g.Ny)), bt(BS);
  return with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_kxrcf<MIX>), gt, bt, 0, ctx->stream, SRd<MIX>{in, g.plane}, kx, g, ctx->cfg.Ck, kx_hk(g),
                       (DevScal*)nullptr);
    hipLaunchKernelGGL((k_tvb<MHD2D_IND_ALL, MIX>), gt, bt, 0, ctx->stream, SRd<MIX>{in, g.plane}, ctx->trouble,
                       ctx->tmask, ctx->trec, g, ctx->cfg.M, ctx->cfg.Ck, kx_hk(g), (DevScal*)nullptr);
    LAUNCHCHK(ctx);
    return 0;
  });
}

// limiter pass in -> out: detector + TVB decisions (k_tvb), then one pass
// over every cell that applies them, the positivity limiter and the
// divergence-free reconstruction (k_limit)
int launch_limit(mhd2d_ctx* ctx, const double* in, double* out, const double* F, bool tvb, bool pp, bool div,
                 DevScal* sc, int stage) {
  const Geo& g = ctx->g;
  const long long n = (long long)(g.Nx + 2) * (g.Ny + 2);
  const int ind = ctx->cfg.indicator;
  if (tvb && ind == MHD2D_IND_NONE) tvb = false;
  Timed tm(ctx, KC_LIMIT);
  if (tvb) {
    int rc = 0;
    if (ind == MHD2D_IND_ALL) {
      rc = with_mix(ctx, [&](auto Mc) -> int {
        constexpr bool MIX = decltype(Mc)::value;
        hipLaunchKernelGGL((k_tvb<MHD2D_IND_ALL, MIX>), dim3(nblk((long long)g.Nx * g.Ny)), dim3(BS), 0, ctx->stream,
                           SRd<MIX>{in, g.plane}, ctx->trouble, ctx->tmask, ctx->trec, g, ctx->cfg.M, ctx->cfg.Ck,
                           kx_hk(g), sc);
        LAUNCHCHK(ctx);
        return 0;
      });
      if (rc) return rc;
    } else if (ind == MHD2D_IND_MLP) {
      // learned screen (once per step), then KXRCF + TVB on its candidates
      if (stage <= 1 && (rc = launch_mlp_screen(ctx, in))) return rc;
      if ((rc = launch_candidates_tvb(ctx, in, sc))) return rc;
    } else {
      if ((rc = launch_detect_tvb(ctx, in, sc))) return rc;
    }
  }
  const dim3 gr(nblk(n)), bl(BS);
  return with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    const SRd<MIX> vi{in, g.plane};
    const SWr<MIX> vo{out, g.plane};
#define LIM(a, b, c) \
    hipLaunchKernelGGL((k_limit<a, b, c, MIX>), gr, bl, 0, ctx->stream, vi, vo, F, ctx->tmask, ctx->trec, g)
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
  });
}

int launch_detection(mhd2d_ctx* ctx, const double* in, DevScal* sc, int stage, const double* inB) {
  const int ind = ctx->cfg.indicator;
  if (ind != MHD2D_IND_MLP && ind != MHD2D_IND_KXRCF) return 0;
  Timed tm(ctx, KC_LIMIT);
  int rc;
  if (ind == MHD2D_IND_MLP) {
    if (stage <= 1 && (rc = launch_mlp_screen(ctx, in, inB))) return rc;
    return launch_candidates_tvb(ctx, in, sc, inB);
  }
  return launch_detect_tvb(ctx, in, sc, inB);
}

int launch_limit_relist(mhd2d_ctx* ctx, double* in, double* out, const double* F, bool pp, int stage,
                        bool lean) {
  const int ind = ctx->cfg.indicator;
  if (ind != MHD2D_IND_MLP && ind != MHD2D_IND_KXRCF) return 0;
  const Geo& g = ctx->g;
  Timed tm(ctx, KC_LIMIT);
  const unsigned nb = std::min<unsigned>(nblk((long long)g.Nx * g.Ny), 1024u);
  return with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    const SWr<MIX> vi{in, g.plane}, vo{out, g.plane};
    auto go = [&](auto Pc, auto Lc) {
      constexpr bool P = decltype(Pc)::value, L = decltype(Lc)::value;
      hipLaunchKernelGGL((k_limit_list<P, MIX, L>), dim3(nb), dim3(BS), 0, ctx->stream, vi, vo, F,
                         (const int*)ctx->tlist, (const int*)ctx->tmask, (const double*)ctx->trec, g,
                         (const DevScal*)ctx->sc, ctx->lflag, stage);
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    if (pp && lean) go(T_{}, T_{});
    else if (pp) go(T_{}, F_{});
    else if (lean) go(F_{}, T_{});
    else go(F_{}, F_{});
    LAUNCHCHK(ctx);
    return 0;
  });
}

int launch_pp_list(mhd2d_ctx* ctx, double* q, double* out, const double* F, int stage) {
  const Geo& g = ctx->g;
  Timed tm(ctx, KC_LIMIT);
  const unsigned nb = std::min<unsigned>(nblk((long long)g.Nx * g.Ny), 1024u);
  return with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_pp_list<MIX>), dim3(nb), dim3(BS), 0, ctx->stream, SWr<MIX>{q, g.plane},
                       SWr<MIX>{out, g.plane}, F, (const int*)ctx->pplist, g, (const DevScal*)ctx->sc, ctx->lflag,
                       stage);
    LAUNCHCHK(ctx);
    return 0;
  });
}

int launch_divfree(mhd2d_ctx* ctx, double* out, const double* F) {
  const Geo& g = ctx->g;
  Timed tm(ctx, KC_LIMIT);
  const long long n = (long long)(g.Nx + 2) * (g.Ny + 2);
  return with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_divfree<MIX>), dim3(nblk(n)), dim3(BS), 0, ctx->stream, SWr<MIX>{out, g.plane}, F, g);
    LAUNCHCHK(ctx);
    return 0;
  });
}


}  // namespace mhd2d_impl
