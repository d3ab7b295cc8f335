// mhd2d_cell.hip -- per-cell residual (volume and surface quadratures, mass
// inverse) with the SSP-RK3 combination as its epilogue, the constrained-
// transport update of the face B modes, and the elementwise stage update of
// the per-subroutine API.
#include "mhd2d_internal.hpp"
#include "mhd2d_limit_cell.hpp"


namespace mhd2d_impl {
namespace {

// ---------------------------------------------------------------------------
// cell residual (Lh.f90:11-68 volume, :268-304 surface + mass inverse) and
// the RK3 combination of RK3.f90 as the epilogue:
//   MODE 0: du                     (Lh alone)
//   MODE 1: out = S + dt du        (uI,  RK3.f90:80)
//   MODE 2: out = 3/4 U0 + 1/4 (UI + dt du)   (RK3.f90:106,110)
//   MODE 3: out = 1/3 U0 + 2/3 (UII + dt du)  (RK3.f90:131,133)
// The launch covers the halo ring too: there du = 0 (Lh.f90:5), so the
// combination carries the start-of-step halos exactly as the reference's
// whole-array statements do (the stale halos TVB later reads).
//
// T is the quadrature type (double: plain form; float: mixed form).  The
// cell-average (mode 1) residual is accumulated in double in both forms.
// The volume integral is sum-factorised row by row in eta: per row the
// modes collapse to 1D coefficients in xi.  The cell's 52 modes are read
// from HBM once into LDS (each lane owns a column of a [plane][lane] array,
// so no barrier is needed) and re-read from LDS for every row; holding them
// in registers next to the 52 accumulators would spill, and re-reading them
// from global memory misses L2 at this occupancy (measured: 2.5x the
// algorithmic traffic).  In the plain form the epilogue takes S from LDS too.
// mixed form (T = float) on square cells (SQ): the volume fluxes are not
// scaled by 1/hx1, 1/hy1 point by point; the volume sums are scaled once
// LIM (plain form): the epilogue also applies the limiter pass without TVB
// (positivity when PP, then the divergence-free reconstruction from the
// stage's new faces Fn) to the stage value as stored.  The stage value's
// planes other than Bx, By go once into lout (the limited value of every cell
// positivity leaves alone), its Bx, By planes into out, the limited Bx, By
// planes into lout; the cells positivity changes go to pplist and are limited
// in place afterwards (k_pp_list), as are the cells TVB changes (k_limit_list).
// This saves the limiter pass's re-read of the stage value and its faces.
// In the mixed form the epilogue is lean: a full positivity epilogue needs
// 254 VGPRs, which halves the kernel's waves per SIMD and costs more than
// the pass it saves (measured).  Instead the epilogue accumulates, plane by
// plane as it stores them, the sums of the positivity bound (pp_settled) and
// writes the planes other than Bx, By of the stage value once, into the
// limited value lout, which they are for every cell the bound settles; the
// stage value's Bx, By planes go to out.  The cells the bound cannot settle
// are appended to pplist and limited in place afterwards by a list kernel
// (k_pp_list), which keeps their unlimited planes in out.  The limited B
// planes come from the faces alone (div_free_Balsara), so a separate pass
// (k_divfree) writes them after the CT update, and this kernel neither waits
// for the CT nor runs beside it.
template <int MODE, typename T, int CBS, bool SQ = false, bool MIX = (sizeof(T) == 4), bool LIM = false,
          bool PP = false>
__global__ __launch_bounds__(CBS, (MIX && LIM && MODE > 0) ? 4 : 1) void k_cell(SRd<MIX> S, const T* __restrict__ FX,
                                              const T* __restrict__ FY, SRd<MIX> A,
                                              SRd<MIX> U0, SWr<MIX> out, SWr<MIX> lout,
                                              const double* __restrict__ Fn, Geo g,
                                              const DevScal* __restrict__ sc, double dt_arg,
                                              int* __restrict__ pplist = nullptr, DevScal* lst = nullptr,
                                              const int* __restrict__ lflag = nullptr) {
  if (step_off(g)) return;
  // The mixed form keeps NR planes of the cell in registers and the other 40
  // in LDS: 160 B per lane, four 256-lane blocks (16 waves) per CU where 52
  // planes (208 B per lane) allowed three, so a SIMD runs four waves instead
  // of three (1.20 -> 1.13 ms per launch at 2048^2, measured).  More planes in
  // registers would need a fifth wave's 102-VGPR limit, which spills.  The
  // register planes are Bx and By modes 1-6 (planes 30-41): the epilogue
  // re-reads the exact FP64 B planes of S or takes A's, so the registers are
  // free once the quadrature is done, while the FP32 planes the epilogue takes
  // from the copy (the stage value's other planes, shared with S) stay in LDS.
  constexpr int NR = MIX ? 12 : 0;
  constexpr int R0 = 30;   // first register plane
  auto in_rg = [](int p) { return NR > 0 && p >= R0 && p < R0 + NR; };
  auto lds_slot = [](int p) { return NR > 0 && p >= R0 + NR ? p - NR : p; };
  __shared__ T sm[(NPL - NR) * CBS];
  T rg[NR > 0 ? NR : 1];
  const int tid = threadIdx.x;
  const long long t = xcd_block() * CBS + tid;
  const int W = g.Nx + 2;
  if (t >= (long long)W * (g.Ny + 2)) return;
  const int i = (int)(t % W), j = (int)(t / W);
  const long long c = cix(g, i, j);
  const long long pl = g.plane;
  const double dt = sc ? sc->dt : dt_arg;
  const bool interior = i >= 1 && i <= g.Nx && j >= 1 && j <= g.Ny;
  double acc0[8];       // mode-1 (cell average) residuals
  T acc[NPL];           // higher modes (slots PL(0,n) unused)
#pragma unroll
  for (int n = 0; n < 8; n++) acc0[n] = 0.0;
#pragma unroll
  for (int p = 0; p < NPL; p++) acc[p] = T(0);
  if (interior) {
#pragma unroll
    for (int p = 0; p < NPL; p++) {
      if (in_rg(p)) rg[in_rg(p) ? p - R0 : 0] = T(S(p, c));
      else sm[lds_slot(p) * CBS + tid] = T(S(p, c));
    }
    const T ihx1 = T(1.0 / g.hx1), ihy1 = T(1.0 / g.hy1);
    // one quadrature row; the mixed form runs rows and columns as
    // compile-time loops (the coefficients and which of them vanish fold to
    // constants), the plain form as a runtime row loop (measured faster in FP64)
    auto row = [&](auto J1) {
      const int j1 = J1;
      const double y = lam(j1);
      const T ty = T(y), tP2y = T(P2(y)), tP3y = T(P3(y));
      int off = tid;
      asm volatile("" : "+v"(off));   // re-read the modes from LDS for every row
      const T* ms = sm + off;
      auto mode = [&](int p) -> T { return in_rg(p) ? rg[in_rg(p) ? p - R0 : 0] : ms[lds_slot(p) * CBS]; };
      // U(x, y_j) = c0 + c1 x + c2 P2(x) + c3 P3(x) per variable
      T c0[8], c1[8], c2[8], c3b = T(0);
#pragma unroll
      for (int n = 0; n < 8; n++) {
        const T m0 = mode(PL(0, n)), m1 = mode(PL(1, n));
        const T m2 = mode(PL(2, n)), m3 = mode(PL(3, n));
        const T m4 = mode(PL(4, n)), m5 = mode(PL(5, n));
        c0[n] = m0 + m2 * ty + m5 * tP2y;
        c1[n] = m1 + m4 * ty;
        c2[n] = m3;
        if (n == 5) {   // Bx: phi7 = P3(xi), phi9 = xi P2(eta)
          c1[n] += mode(PL(8, 5)) * tP2y;
          c3b = mode(PL(6, 5));
        }
        if (n == 6) {   // By: phi8 = P2(xi) eta, phi10 = P3(eta)
          c2[n] += mode(PL(7, 6)) * ty;
          c0[n] += mode(PL(9, 6)) * tP3y;
        }
      }
      auto col = [&](auto I1) {
        const int i1 = I1;
        const double x = lam(i1);
        const T tx = T(x), tP2x = T(P2(x)), tP3x = T(P3(x));
        T Uq[8], Fx[8], Fy[8];
#pragma unroll
        for (int n = 0; n < 8; n++) Uq[n] = c0[n] + c1[n] * tx + c2[n] * tP2x;
        Uq[5] += c3b * tP3x;
        lh_flux_xy<T>(Uq, Fx, Fy);
        if (!SQ) {
#pragma unroll
          for (int n = 0; n < 8; n++) { Fx[n] *= ihx1; Fy[n] *= ihy1; }
        }
        const double ww = 0.25 * wgt(i1) * wgt(j1);
#pragma unroll
        for (int n = 0; n < 8; n++) {
#pragma unroll
          for (int b = 1; b < 10; b++) {
            const int p = PL(b, n);
            if (p < 0) continue;
            const double cx = ww * dphix(b, x, y), cy = ww * dphiy(b, x, y);
            if (cx != 0.0) acc[p] += T(cx) * Fx[n];
            if (cy != 0.0) acc[p] += T(cy) * Fy[n];
          }
        }
      };
      if constexpr (MIX) {
        static_for<0, 5>(col);
      } else {
#pragma unroll
        for (int i1 = 0; i1 < 5; i1++) col(i1);
      }
    };
    if constexpr (MIX) {
      static_for<0, 5>(row);
    } else {
#pragma unroll 1
      for (int j1 = 0; j1 < 5; j1++) row(j1);
    }
    if (SQ) {
#pragma unroll
      for (int p = 0; p < NPL; p++) acc[p] *= ihx1;
    }
    // surface integrals
    const long long cL = cix(g, i - 1, j), cD = cix(g, i, j - 1);
    if constexpr (MIX) {
      // face moments (FM_*): the trace of mode b on a face is trk(side, b, k)
      // L_k(s) for one k, so its Gauss sum is trk * M_k.  trk(1, b, k) =
      // +-trk(0, b, k) (the parity of b in xi), so each mode takes
      // c (M_R - M_L) or c (M_R + M_L).
      const T sx = T(0.5 / g.hx), sy = T(0.5 / g.hy);
      const double sxd = 0.5 / g.hx, syd = 0.5 / g.hy;
#pragma unroll
      for (int n = 0; n < 8; n++) {
        T mr[4], ml[4], mu[4], md[4];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          mr[k] = FX[(FM_N * n + k) * pl + c];
          ml[k] = FX[(FM_N * n + k) * pl + cL];
          mu[k] = FY[(FM_N * n + k) * pl + c];
          md[k] = FY[(FM_N * n + k) * pl + cD];
        }
        mr[3] = ml[3] = mu[3] = md[3] = T(0);
        if (n == 6) { mr[3] = FX[FM_P3 * pl + c]; ml[3] = FX[FM_P3 * pl + cL]; }
        if (n == 5) { mu[3] = FY[FM_P3 * pl + c]; md[3] = FY[FM_P3 * pl + cD]; }
        acc0[n] -= sxd * ((double)mr[0] - (double)ml[0]);
        acc0[n] -= syd * ((double)mu[0] - (double)md[0]);
        T xd[4], xs[4], yd[4], ys[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          xd[k] = (mr[k] - ml[k]) * sx; xs[k] = (mr[k] + ml[k]) * sx;
          yd[k] = (mu[k] - md[k]) * sy; ys[k] = (mu[k] + md[k]) * sy;
        }
#pragma unroll
        for (int b = 1; b < 10; b++) {
          const int p = PL(b, n);
          if (p < 0) continue;
          const int kx = trace_k(0, b), ky = trace_k(2, b);
          const double cr = trk(0, b, kx), cu = trk(2, b, ky);
          acc[p] -= T(cr) * (trk(1, b, kx) == -cr ? xs[kx] : xd[kx]);
          acc[p] -= T(cu) * (trk(3, b, ky) == -cu ? ys[ky] : yd[ky]);
        }
      }
    } else {
      const T sx = T(0.5 / g.hx), sy = T(0.5 / g.hy);
      const double sxd = 0.5 / g.hx, syd = 0.5 / g.hy;
#pragma unroll
      for (int q = 0; q < 5; q++) {
        const double s = lam(q), p2 = P2(s), p3 = P3(s), t23 = T23(s);
        const double wq = wgt(q);
#pragma unroll
        for (int n = 0; n < 8; n++) {
          const T fr = FX[(q * 8 + n) * pl + c], fl = FX[(q * 8 + n) * pl + cL];
          const T fu = FY[(q * 8 + n) * pl + c], fd = FY[(q * 8 + n) * pl + cD];
          acc0[n] -= sxd * wq * ((double)fr - (double)fl);
          acc0[n] -= syd * wq * ((double)fu - (double)fd);
#pragma unroll
          for (int b = 1; b < 10; b++) {
            const int p = PL(b, n);
            if (p < 0) continue;
            acc[p] -= sx * T(wq) * (fr * T(phiTr(0, b, s, p2, p3, t23)) - fl * T(phiTr(1, b, s, p2, p3, t23)));
            acc[p] -= sy * T(wq) * (fu * T(phiTr(2, b, s, p2, p3, t23)) - fd * T(phiTr(3, b, s, p2, p3, t23)));
          }
        }
      }
    }
  }
  // epilogue: mass inverse + RK combination, in double
  constexpr bool LEAN = LIM && MIX;   // the mixed form's lean limiter epilogue
  constexpr bool SHARE = LIM;         // the stage value shares its planes other than Bx, By with lout
  double m[LIM && !LEAN ? NPL : 1];   // LIM: the stage value as stored (FP32 planes rounded)
  double pa[LEAN ? 8 : 1], pav[LEAN ? 8 : 1];   // LEAN: positivity bound sums and the averages
#pragma unroll
  for (int n = 0; n < (LEAN ? 8 : 1); n++) { pa[n] = 0.0; pav[n] = 0.0; }
  // LEAN: the stage value's planes other than Bx, By are stored once, in the
  // limited value lout (for every cell the list kernels do not limit, the
  // two agree); out, its companion, holds its Bx, By planes, all planes of a
  // halo cell (set_bc later refreshes lout's halos) and, written by the list
  // kernels, the unlimited planes of the cells they limited in place (lflag).
  // So the previous stage value A is read from out's predecessor only where
  // aq; elsewhere its other planes are S's: the LDS copy (exact for the FP32
  // planes) and a re-read of the FP64 cell averages.  An aq lane (rare)
  // refills its LDS column with A's planes first, so that one code path
  // serves both.
  if constexpr (SHARE && MODE >= 2) {
    const int tag = (int)(4 * sc->epoch + (MODE - 1));
    if (interior && lflag[(long long)(j - 1) * g.Nx + (i - 1)] == tag) {
      static_for<0, NPL>([&](auto P) {
        constexpr int p = P;
        if constexpr (!is_bplane(p) && (sizeof(T) == 8 || !pl_f64(p))) {
          const T v = T(A(p, c));
          sm[lds_slot(p) * CBS + tid] = v;   // no register plane holds one of these
        }
      });
    }
  }
  auto plane = [&](auto P) {
    const int p = P;
    const bool avg = p < 48 && (p % 6) == 0;
    const int b = p < 48 ? p % 6 : (p == 48 ? 6 : p == 49 ? 8 : p == 50 ? 7 : 9);
    const double du = avg ? acc0[avg ? p / 6 : 0] : (double)acc[p] * (1.0 / mmass(b));
    // the LDS copy is exact for the planes stored in the quadrature type
    auto s_exact = [&]() -> double {
      return interior && (sizeof(T) == 8 || !pl_f64(p)) ? (double)(in_rg(p) ? rg[in_rg(p) ? p - R0 : 0] : sm[lds_slot(p) * CBS + tid])
                                                         : ((SHARE && MODE >= 2 && !interior) ? A(p, c) : S(p, c));
    };
    double a = 0.0;
    if (MODE >= 2) a = (SHARE && !is_bplane(p)) ? s_exact() : A(p, c);
    double r;
    if (MODE == 0) r = du;
    else if (MODE == 1) r = s_exact() + dt * du;
    else if (MODE == 2) r = (3.0 / 4.0) * U0(p, c) + (1.0 / 4.0) * (a + dt * du);
    else r = (1.0 / 3.0) * U0(p, c) + (2.0 / 3.0) * (a + dt * du);
    if (!SHARE || is_bplane(p) || !interior) out.set(p, c, r);
    if constexpr (LIM && !LEAN) {
      m[p] = r;
      if (!is_bplane(p)) lout.set(p, c, r);   // the stage value, also the limited value of a cell pp leaves alone
    }
    if constexpr (LEAN) {
      const double v = pl_f64(p) ? r : (double)(float)r;   // the value as stored
      if (!is_bplane(p)) lout.set(p, c, v);                  // the stage value, also the limited value of a settled cell
      const int n = p < 48 ? p / 6 : (p < 50 ? 5 : 6);
      if (avg) pav[n] = v;
      else pa[n] += fabs(v) * pp_pmax(b);                    // b in increasing order for every n
    }
  };
  // the mixed storage map needs the plane as a compile-time constant (a
  // partially unrolled loop would index acc[] dynamically, i.e. in scratch)
  if constexpr (MIX) {
    static_for<0, NPL>(plane);
  } else {
#pragma unroll
    for (int p = 0; p < NPL; p++) plane(p);
  }
  if constexpr (LEAN) {
    if constexpr (PP) {
      const bool need = interior && !pp_settled_sums(pa, pav, 1e-13);
      pp_append(need, interior ? (j - 1) * g.Nx + (i - 1) : 0, pplist, lst);
    }
  } else if constexpr (LIM) {
    // the limiter pass of k_limit<false, PP, true> on the stored stage value:
    // the cells positivity changes go to pplist (k_pp_list limits them in
    // place after the detection has read the stage value); the limited Bx, By
    // planes come from the faces (div_free_Balsara)
    bool changed = false;
    if (interior) {
      if (PP) changed = pp_cell(m);
      divfree_cell(m, Fn, g, i, j);
#pragma unroll
      for (int p = 0; p < NPL; p++)
        if (is_bplane(p)) lout.set(p, c, m[p]);
    } else {
#pragma unroll
      for (int p = 0; p < NPL; p++)
        if (is_bplane(p)) lout.set(p, c, 0.0);
    }
    if constexpr (PP) pp_append(changed, interior ? (j - 1) * g.Nx + (i - 1) : 0, pplist, lst);
  }
}

// constrained transport: LhEz.f90:71-96 (+ the face combinations of
// RK3.f90:82-83, 107-112, 129-135 for MODE 1..3 as in k_cell).  Face planes: 0..2 Bx modes
// on x faces (i = 0..Nx, j = 1..Ny), 3..5 By modes on y faces (i = 1..Nx,
// j = 0..Ny).
template <int MODE, typename T>
__global__ __launch_bounds__(BS) void k_ct(const T* __restrict__ EZX, const T* __restrict__ EZY,
                                           const double* __restrict__ EZV, const double* __restrict__ F,
                                           const double* __restrict__ A, const double* __restrict__ F0,
                                           double* __restrict__ out, Geo g, const DevScal* __restrict__ sc,
                                           double dt_arg) {
  if (step_off(g)) return;
  const long long t = xcd_block() * BS + threadIdx.x;
  const long long nx = (long long)(g.Nx + 1) * g.Ny, ny = (long long)g.Nx * (g.Ny + 1);
  if (t >= nx + ny) return;
  const double dt = sc ? sc->dt : dt_arg;
  const long long pl = g.plane;
  const double mmE[3] = {1.0, 1.0 / 3.0, 4.0 / 45.0};
  const double EzP[3] = {1.0, 1.0, 2.0 / 3.0};    // EzR = EzU
  const double EzM[3] = {1.0, -1.0, 2.0 / 3.0};   // EzL = EzD
  double d[3];
  long long o;
  int base;
  double va, vb, ch;   // mode-1 increment dt * (va - vb) / h of the face (mixed form)
  if (t < nx) {
    const int i = (int)(t % (g.Nx + 1)), j = 1 + (int)(t / (g.Nx + 1));
    o = cix(g, i, j);
    base = 0;
    const double iy1 = 1.0 / g.hy1;
    double a[3] = {0.0, 0.0, 0.0};
    if (sizeof(T) == 4) {   // mixed form: the face kernel stored sum_q w_q e_q and sum_q w_q s_q e_q
      a[1] = 0.5 * iy1 * (double)EZX[o];
      a[2] = iy1 * (double)EZX[pl + o];
    } else {
#pragma unroll
      for (int q = 0; q < 5; q++) {
        const double e = 0.5 * wgt(q) * (double)EZX[q * pl + o];
        a[1] += e * iy1;
        a[2] += e * (2 * lam(q) * iy1);
      }
    }
    const double vU = EZV[cix(g, i, j)], vD = EZV[cix(g, i, j - 1)];
#pragma unroll
    for (int k = 0; k < 3; k++) d[k] = (a[k] - vU * EzP[k] / g.hy + vD * EzM[k] / g.hy) / mmE[k];
    va = vD; vb = vU; ch = g.hy;
  } else {
    const long long u = t - nx;
    const int i = 1 + (int)(u % g.Nx), j = (int)(u / g.Nx);
    o = cix(g, i, j);
    base = 3;
    const double ix1 = 1.0 / g.hx1;
    double a[3] = {0.0, 0.0, 0.0};
    if (sizeof(T) == 4) {
      a[1] = -0.5 * ix1 * (double)EZY[o];
      a[2] = -ix1 * (double)EZY[pl + o];
    } else {
#pragma unroll
      for (int q = 0; q < 5; q++) {
        const double e = 0.5 * wgt(q) * (double)EZY[q * pl + o];
        a[1] -= e * ix1;
        a[2] -= e * (2 * lam(q) * ix1);
      }
    }
    const double vR = EZV[cix(g, i, j)], vL = EZV[cix(g, i - 1, j)];
#pragma unroll
    for (int k = 0; k < 3; k++) d[k] = (a[k] + vR * EzP[k] / g.hx - vL * EzM[k] / g.hx) / mmE[k];
    va = vR; vb = vL; ch = g.hx;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const long long oo = (base + k) * pl + o;
    double r;
    if (MODE == 0) r = d[k];
    else if (MODE == 1) r = F[oo] + dt * d[k];
    else if (MODE == 2) r = (3.0 / 4.0) * F0[oo] + (1.0 / 4.0) * (A[oo] + dt * d[k]);
    else r = (1.0 / 3.0) * F0[oo] + (2.0 / 3.0) * (A[oo] + dt * d[k]);
    if (sizeof(T) == 4 && MODE > 0 && k == 0) continue;   // mixed form: mode 1 below
    out[oo] = r;
  }
  if (sizeof(T) == 4 && MODE > 0) {
    // mixed form, mode 1 in double-double: the increment dt/h (va - vb) with
    // the vertex difference taken exactly (-dt/hy (Ez_U - Ez_D) on x faces,
    // dt/hx (Ez_R - Ez_L) on y faces), so the four increments of a cell
    // cancel in its discrete divergence and the RK combinations keep the
    // cancellation; the stored high parts carry one rounding, not one per stage
    const long long oo = base * pl + o, ol = (base == 0 ? FACE_LO_X : FACE_LO_Y) * pl + o;
    const dd inc = dd_scale(dt / ch, two_sum(va, -vb));
    dd r;
    if (MODE == 1) {
      r = dd_add(dd{F[oo], F[ol]}, inc);
    } else {
      const dd x = dd_add(dd{A[oo], A[ol]}, inc), f0{F0[oo], F0[ol]};
      r = MODE == 2 ? dd_add(dd_scale(3.0 / 4.0, f0), dd_scale(1.0 / 4.0, x))
                    : dd_add(dd_scale(1.0 / 3.0, f0), dd_scale(2.0 / 3.0, x));
    }
    out[oo] = r.hi;
    out[ol] = r.lo;
  }
}

__global__ void k_reset_npp(DevScal* lst, const int* gate) {
  if (step_off(gate)) return;
  if (threadIdx.x == 0) lst->npp = 0u;
}

// the FP32 planes of the mixed storage: combined in double, stored in float
__global__ __launch_bounds__(BS) void k_axpy_f(float* __restrict__ out, const float* __restrict__ X, double a,
                                               const float* __restrict__ Y, double b, const float* __restrict__ Z,
                                               double dt, long long n) {
  const long long k = (long long)blockIdx.x * BS + threadIdx.x;
  if (k >= n) return;
  const double y = (double)Y[k] + dt * (double)Z[k];
  out[k] = (float)(X ? a * (double)X[k] + b * y : y);
}

// whole-array combinations of RK3.f90 for the per-subroutine API
// (mhd2d_rk_stage): out = a*X + b*(Y + dt*Z), elementwise
__global__ __launch_bounds__(BS) void k_axpy(double* __restrict__ out, const double* __restrict__ X, double a,
                                             const double* __restrict__ Y, double b, const double* __restrict__ Z,
                                             double dt, long long n) {
  const long long k = (long long)blockIdx.x * BS + threadIdx.x;
  if (k >= n) return;
  const double y = Y[k] + dt * Z[k];
  out[k] = X ? a * X[k] + b * y : y;
}

}  // namespace

// block sizes of k_cell: its LDS copy of the state is 52 values per lane
// (plain form: 128 lanes x 416 B = 53 KB per block, 3 blocks per CU)
constexpr int CBS_F64 = 128, CBS_F32 = 256;

template <int MODE, bool LIM, bool PP>
void launch_cell_t(mhd2d_ctx* ctx, const double* S, const double* A, const double* U0, double* out,
                   double* lout, const double* Fn, const DevScal* sc, double dt) {
  const Geo& g = ctx->g;
  const long long n = (long long)(g.Nx + 2) * (g.Ny + 2);
  const long long pl = g.plane;
  if constexpr (LIM) {
    if (ctx->cfg.precision == MHD2D_MIXED) {
      // the lean epilogue's positivity list starts empty
      hipLaunchKernelGGL(k_reset_npp, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, g.gate);
      if (g.hx1 == g.hy1)
        hipLaunchKernelGGL((k_cell<MODE, float, CBS_F32, true, true, true, PP>),
                           dim3((unsigned)((n + CBS_F32 - 1) / CBS_F32)), dim3(CBS_F32), 0, ctx->stream,
                           SRd<true>{S, pl}, (const float*)ctx->FX, (const float*)ctx->FY, SRd<true>{A, pl},
                           SRd<true>{U0, pl}, SWr<true>{out, pl}, SWr<true>{lout, pl}, Fn, g, sc, dt, ctx->pplist,
                           ctx->sc, (const int*)ctx->lflag);
      else
        hipLaunchKernelGGL((k_cell<MODE, float, CBS_F32, false, true, true, PP>),
                           dim3((unsigned)((n + CBS_F32 - 1) / CBS_F32)), dim3(CBS_F32), 0, ctx->stream,
                           SRd<true>{S, pl}, (const float*)ctx->FX, (const float*)ctx->FY, SRd<true>{A, pl},
                           SRd<true>{U0, pl}, SWr<true>{out, pl}, SWr<true>{lout, pl}, Fn, g, sc, dt, ctx->pplist,
                           ctx->sc, (const int*)ctx->lflag);
      return;
    }
    hipLaunchKernelGGL(k_reset_npp, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, g.gate);
    hipLaunchKernelGGL((k_cell<MODE, double, CBS_F64, false, false, true, PP>),
                       dim3((unsigned)((n + CBS_F64 - 1) / CBS_F64)), dim3(CBS_F64), 0, ctx->stream, SRd<false>{S, pl},
                       (const double*)ctx->FX, (const double*)ctx->FY, SRd<false>{A, pl}, SRd<false>{U0, pl},
                       SWr<false>{out, pl}, SWr<false>{lout, pl}, Fn, g, sc, dt, ctx->pplist, ctx->sc,
                       (const int*)ctx->lflag);
    return;
  }
  if (ctx->cfg.precision == MHD2D_MIXED && g.hx1 == g.hy1)
    hipLaunchKernelGGL((k_cell<MODE, float, CBS_F32, true, true, false, false>), dim3((unsigned)((n + CBS_F32 - 1) / CBS_F32)),
                       dim3(CBS_F32), 0, ctx->stream, SRd<true>{S, pl}, (const float*)ctx->FX, (const float*)ctx->FY,
                       SRd<true>{A, pl}, SRd<true>{U0, pl}, SWr<true>{out, pl}, SWr<true>{lout, pl}, Fn, g, sc, dt);
  else if (ctx->cfg.precision == MHD2D_MIXED)
    hipLaunchKernelGGL((k_cell<MODE, float, CBS_F32, false, true, false, false>), dim3((unsigned)((n + CBS_F32 - 1) / CBS_F32)),
                       dim3(CBS_F32), 0, ctx->stream, SRd<true>{S, pl}, (const float*)ctx->FX, (const float*)ctx->FY,
                       SRd<true>{A, pl}, SRd<true>{U0, pl}, SWr<true>{out, pl}, SWr<true>{lout, pl}, Fn, g, sc, dt);
  else
    hipLaunchKernelGGL((k_cell<MODE, double, CBS_F64, false, false, false, false>), dim3((unsigned)((n + CBS_F64 - 1) / CBS_F64)),
                       dim3(CBS_F64), 0, ctx->stream, SRd<false>{S, pl}, (const double*)ctx->FX,
                       (const double*)ctx->FY, SRd<false>{A, pl}, SRd<false>{U0, pl}, SWr<false>{out, pl},
                       SWr<false>{lout, pl}, Fn, g, sc, dt);
}

template <bool LIM, bool PP>
void launch_cell_mode(mhd2d_ctx* ctx, int mode, const double* S, const double* A, const double* U0, double* out,
                      double* lout, const double* Fn, const DevScal* sc, double dt) {
  switch (mode) {
    case 0: if constexpr (!LIM) launch_cell_t<0, false, false>(ctx, S, A, U0, out, lout, Fn, sc, dt); break;
    case 1: launch_cell_t<1, LIM, PP>(ctx, S, A, U0, out, lout, Fn, sc, dt); break;
    case 2: launch_cell_t<2, LIM, PP>(ctx, S, A, U0, out, lout, Fn, sc, dt); break;
    default: launch_cell_t<3, LIM, PP>(ctx, S, A, U0, out, lout, Fn, sc, dt); break;
  }
}

int launch_cell(mhd2d_ctx* ctx, int mode, const double* S, const double* A, const double* U0, double* out,
                const DevScal* sc, double dt, double* lout, const double* Fn, bool pp) {
  if (lout && mode == 0) return fail(ctx, -1, "launch_cell: the fused limiter is for the RK stages");
  Timed tm(ctx, KC_CELL);
  if (!lout) launch_cell_mode<false, false>(ctx, mode, S, A, U0, out, nullptr, nullptr, sc, dt);
  else if (pp) launch_cell_mode<true, true>(ctx, mode, S, A, U0, out, lout, Fn, sc, dt);
  else launch_cell_mode<true, false>(ctx, mode, S, A, U0, out, lout, Fn, sc, dt);
  LAUNCHCHK(ctx);
  return 0;
}

template <int MODE>
void launch_ct_t(mhd2d_ctx* ctx, const double* F, const double* A, const double* F0, double* out,
                 const DevScal* sc, double dt) {
  const Geo& g = ctx->g;
  const long long n = (long long)(g.Nx + 1) * g.Ny + (long long)g.Nx * (g.Ny + 1);
  if (ctx->cfg.precision == MHD2D_MIXED)
    hipLaunchKernelGGL((k_ct<MODE, float>), dim3(nblk(n)), dim3(BS), 0, ctx->stream, (const float*)ctx->EZX,
                       (const float*)ctx->EZY, ctx->EZV, F, A, F0, out, g, sc, dt);
  else
    hipLaunchKernelGGL((k_ct<MODE, double>), dim3(nblk(n)), dim3(BS), 0, ctx->stream, (const double*)ctx->EZX,
                       (const double*)ctx->EZY, ctx->EZV, F, A, F0, out, g, sc, dt);
}

int launch_ct(mhd2d_ctx* ctx, int mode, const double* F, const double* A, const double* F0, double* out,
              const DevScal* sc, double dt) {
  Timed tm(ctx, KC_CT);
  switch (mode) {
    case 0: launch_ct_t<0>(ctx, F, A, F0, out, sc, dt); break;
    case 1: launch_ct_t<1>(ctx, F, A, F0, out, sc, dt); break;
    case 2: launch_ct_t<2>(ctx, F, A, F0, out, sc, dt); break;
    default: launch_ct_t<3>(ctx, F, A, F0, out, sc, dt); break;
  }
  LAUNCHCHK(ctx);
  return 0;
}

int launch_axpy(mhd2d_ctx* ctx, double* out, const double* X, double a, const double* Y, double b,
                const double* Z, double dt, long long n) {
  hipLaunchKernelGGL(k_axpy, dim3(nblk(n)), dim3(BS), 0, ctx->stream, out, X, a, Y, b, Z, dt, n);
  LAUNCHCHK(ctx);
  return 0;
}

// the same over state buffers: in the mixed storage the FP64 planes, then
// the FP32 planes (combined in double, rounded once)
int launch_axpy_state(mhd2d_ctx* ctx, double* out, const double* X, double a, const double* Y, double b,
                      const double* Z, double dt) {
  const long long pl = ctx->g.plane;
  if (ctx->cfg.precision != MHD2D_MIXED) return launch_axpy(ctx, out, X, a, Y, b, Z, dt, (long long)NPL * pl);
  int rc = launch_axpy(ctx, out, X, a, Y, b, Z, dt, (long long)NPL_D * pl);
  if (rc) return rc;
  auto f = [&](const double* p) { return p ? reinterpret_cast<const float*>(p + NPL_D * pl) : nullptr; };
  const long long nf = (long long)NPL_F * pl;
  hipLaunchKernelGGL(k_axpy_f, dim3(nblk(nf)), dim3(BS), 0, ctx->stream, reinterpret_cast<float*>(out + NPL_D * pl),
                     f(X), a, f(Y), b, f(Z), dt, nf);
  LAUNCHCHK(ctx);
  return 0;
}

}  // namespace mhd2d_impl
