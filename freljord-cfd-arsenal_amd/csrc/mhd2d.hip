// mhd2d.hip -- gfx950 kernels and the C ABI (include/mhd2d.h) for one time
// step of the 2D ideal-MHD P2 globally divergence-free DG solver of
// GDF_Balsara_P2_0316.
//
// Device state: the 52 active (mode, variable) planes of uh, each an
// (Nx+2) x (Ny+2) array with halos (so set_bc and the stale-halo reads of
// the TVB limiter keep the reference's exact meaning), plus 6 face planes
// (Bx modes 1-3 on x faces, By modes 1-3 on y faces).  Cell (i, j) of any
// plane sits at j*P + OFF + i (mhd2d_tables.hpp).
//
// Kernels per RK stage (fused path, mhd2d_step):
//   k_xface / k_yface  one lane per (face, Gauss point): traces, physical
//                      fluxes, Davis bounds, 1D Riemann flux, face EMF
//   k_vertex           one lane per vertex: 2D Riemann (corner) EMF
//   k_cell<MODE>       one lane per cell: volume + surface integrals, mass
//                      inverse, SSP-RK3 combination (RK3.f90:80-135)
//   k_ct<MODE>         one lane per face: CT right-hand side + combination
//   k_limit            one lane per cell: TVB (+ positivity) + Balsara
//                      divergence-free reconstruction
//   k_setbc            halo ring//
// This file: halo ring, reductions, initial conditions and the C ABI.
// Kernels of the stage live in mhd2d_faces.hip, mhd2d_cell.hip and
// mhd2d_limit.hip; shared declarations in mhd2d_internal.hpp.
#include <atomic>
#include <mutex>

#include "mhd2d_internal.hpp"

using namespace mhd;
using namespace mhd2d_impl;

namespace {

// set_bc.f90:5-27 over the halo ring: periodic (1) or outflow (2); the
// x-then-y order of the reference makes the corners (map_x(i), map_y(j)).
template <bool MIX>
__global__ __launch_bounds__(BS) void k_setbc(SWr<MIX> U, Geo g, int nplanes, const DevScal* __restrict__ sc) {
  if (step_off(g)) return;
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  const int W = g.Nx + 2, H = g.Ny + 2;
  const long long ring = 2LL * W + 2LL * g.Ny;
  if (t >= ring * nplanes) return;
  const int p = (int)(t / ring);
  const long long r = t % ring;
  int i, j;
  if (r < W) { i = (int)r; j = 0; }
  else if (r < 2 * W) { i = (int)(r - W); j = H - 1; }
  else if (r < 2 * W + g.Ny) { i = 0; j = 1 + (int)(r - 2 * W); }
  else { i = W - 1; j = 1 + (int)(r - 2 * W - g.Ny); }
  auto mapx = [&](int a) {
    if (a == 0) return g.bcL == 1 ? g.Nx : g.bcL == 2 ? 1 : 0;
    if (a == g.Nx + 1) return g.bcR == 1 ? 1 : g.bcR == 2 ? g.Nx : g.Nx + 1;
    return a;
  };
  auto mapy = [&](int b) {
    if (b == 0) return g.bcD == 1 ? g.Ny : g.bcD == 2 ? 1 : 0;
    if (b == g.Ny + 1) return g.bcU == 1 ? 1 : g.bcU == 2 ? g.Ny : g.Ny + 1;
    return b;
  };
  const int si = mapx(i), sj = mapy(j);
  if (si == i && sj == j) return;
  U.set(p, cix(g, i, j), U(p, cix(g, si, sj)));
}

// ---------------------------------------------------------------------------
// reductions
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_down(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__global__ void k_set_tend(DevScal* sc, double tend) {
  if (threadIdx.x == 0) sc->tend = tend;
}

__global__ void k_scal_reset(DevScal* sc, int which) {
  if (threadIdx.x != 0) return;
  if (which == 0) { sc->ax_bits = 0ull; sc->ay_bits = 0ull; }
  else { sc->umax_bits = 0ull; sc->pmin_key = dkey(1000000.0); sc->nflag = 0; }
}

// calculate_dt.f90:12-24: max |u_n +- c_f| at the cell averages
template <bool MIX>
__global__ __launch_bounds__(BS) void k_dt(SRd<MIX> U, Geo g, DevScal* sc) {
  double ax = 0.0, ay = 0.0;
  // grid-stride: a bounded grid keeps the same-address atomics below few thousand
  for (long long t = (long long)blockIdx.x * BS + threadIdx.x; t < (long long)g.Nx * g.Ny;
       t += (long long)gridDim.x * BS) {
    const int i = 1 + (int)(t % g.Nx), j = 1 + (int)(t / g.Nx);
    double u[8];
#pragma unroll
    for (int n = 0; n < 8; n++) u[n] = U(PL(0, n), cix(g, i, j));
    double a1, a2;
    eigenvalueMm<double, 1>(u, a1, a2);
    const double mx = fmax(fabs(a1), fabs(a2));
    if (!isfinite(mx)) atomicOr(&sc->health, MHD2D_FLAG_NONFINITE);
    if (mx > ax) ax = mx;      // NaN never exceeds the running max
    eigenvalueMm<double, 2>(u, a1, a2);
    const double my = fmax(fabs(a1), fabs(a2));
    if (my > ay) ay = my;
  }
  __shared__ double sx[BS / 64], sy[BS / 64];
  ax = wave_max(ax);
  ay = wave_max(ay);
  if ((threadIdx.x & 63) == 0) { sx[threadIdx.x >> 6] = ax; sy[threadIdx.x >> 6] = ay; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BS / 64; w++) { ax = fmax(ax, sx[w]); ay = fmax(ay, sy[w]); }
    atomicMax(&sc->ax_bits, (unsigned long long)__double_as_longlong(ax));
    atomicMax(&sc->ay_bits, (unsigned long long)__double_as_longlong(ay));
  }
}

// RK3.f90:60-71: dt from the maxima, clipped to tend; steps that start at
// t >= tend are switched off (active = 0) like the reference's loop test.
__global__ void k_dt_final(DevScal* sc, double hx, double hy, int advance) {
  if (threadIdx.x != 0) return;
  if (advance && !(sc->t < sc->tend)) { sc->active = 0; return; }
  const double ax = __longlong_as_double((long long)sc->ax_bits);
  const double ay = __longlong_as_double((long long)sc->ay_bits);
  sc->alphax = ax;
  sc->alphay = ay;
  double dt = sc->CFL / (ax / hx + ay / hy);
  if (!isfinite(dt)) sc->health |= MHD2D_FLAG_NONFINITE;
  if (!advance) { sc->dt = dt; return; }
  sc->active = 1;
  sc->epoch += 1;
  if (dt < 1e-7) sc->health |= MHD2D_FLAG_DT_SMALL;   // the CFL step itself collapsed
  if (sc->t + dt >= sc->tend) { dt = sc->tend - sc->t; sc->t = sc->tend; }
  else sc->t = sc->t + dt;
  if (dt < 1e-7) sc->t = sc->tend;
  sc->dt = dt;
  sc->count += 1;
}

// largest mismatch between a face's normal-B polynomial and the trace of the
// in-cell B of the interior cells on either side, over the face's Gauss
// points: zero up to rounding after div_free_Balsara (div_free_Balsara.f90:
// 43-58 makes each trace equal its face polynomial); halo cells are skipped
// (their B planes are zeroed by the limiter pass)
template <bool MIX>
__global__ __launch_bounds__(BS) void k_facejump(SRd<MIX> U, const double* __restrict__ F, Geo g,
                                                 unsigned long long* __restrict__ out) {
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  const long long nx = (long long)(g.Nx + 1) * g.Ny, ny = (long long)g.Nx * (g.Ny + 1);
  double mx = 0.0;
  if (t < nx + ny) {
    const bool xf = t < nx;
    const long long u = xf ? t : t - nx;
    const int i = xf ? (int)(u % (g.Nx + 1)) : 1 + (int)(u % g.Nx);
    const int j = xf ? 1 + (int)(u / (g.Nx + 1)) : (int)(u / g.Nx);
    const int n = xf ? 5 : 6, base = xf ? 0 : 3;
    const long long o = cix(g, i, j), cB = xf ? cix(g, i + 1, j) : cix(g, i, j + 1);
    const bool useA = xf ? i >= 1 : j >= 1, useB = xf ? i < g.Nx : j < g.Ny;
    const double a0 = F[base * g.plane + o], a1 = F[(base + 1) * g.plane + o], a2 = F[(base + 2) * g.plane + o];
#pragma unroll
    for (int q = 0; q < 5; q++) {
      const double sq = lam(q), p2 = P2(sq), p3 = P3(sq), t23 = T23(sq);
      const double face = a0 + a1 * sq + a2 * p2;
      double ta = 0.0, tb = 0.0;
#pragma unroll
      for (int b = 0; b < 10; b++) {
        const int p = PL(b, n);
        if (p < 0) continue;
        ta += U(p, o) * phiTr(xf ? 0 : 2, b, sq, p2, p3, t23);
        tb += U(p, cB) * phiTr(xf ? 1 : 3, b, sq, p2, p3, t23);
      }
      if (useA) mx = fmax(mx, fabs(ta - face));
      if (useB) mx = fmax(mx, fabs(tb - face));
    }
  }
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(mx));
}

// a step that started at t >= tend ran none of its kernels: its state and
// face B move to the buffers that hold them after the step (the host rotates
// the buffer roles every step)
__global__ __launch_bounds__(BS) void k_carry(const double* __restrict__ src, double* __restrict__ dst, long long n,
                                              const int* gate) {
  if (!gate || *gate != 0) return;
  for (long long t = (long long)blockIdx.x * BS + threadIdx.x; t < n; t += (long long)gridDim.x * BS) dst[t] = src[t];
}

// calculate_umax.f90, calculate_pmin.f90 and per-block partial sums of
// calculate_totaldiv.f90 (summed in a fixed order by k_div_final).  V is the
// state view: SRd<MIX> over a context's planes, or RefView over an array in
// the reference layout (the allocation-free diagnostics)
template <class V>
__global__ __launch_bounds__(BS) void k_diag(V U, Geo g, DevScal* sc, double* partial) {
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  double um = 0.0, pm = 1000000.0, dv = 0.0;
  if (t < (long long)g.Nx * g.Ny) {
    const int i = 1 + (int)(t % g.Nx), j = 1 + (int)(t / g.Nx);
    const long long c = cix(g, i, j);
    double m[NPL];
#pragma unroll
    for (int p = 0; p < NPL; p++) m[p] = U(p, c);
    um = fabs(m[PL(0, 0)]);
    const double ihx1 = 1.0 / g.hx1, ihy1 = 1.0 / g.hy1;
#pragma unroll
    for (int i1 = 0; i1 < 5; i1++)
#pragma unroll
      for (int j1 = 0; j1 < 5; j1++) {
        const double x = lam(i1), y = lam(j1);
        double uq[8];
#pragma unroll
        for (int n = 0; n < 8; n++) {
          double a = 0.0;
#pragma unroll
          for (int b = 0; b < 10; b++) {
            const int p = PL(b, n);
            if (p >= 0) a += m[p] * phi(b, x, y);
          }
          uq[n] = a;
        }
        const double p1 = pressure<double>(uq);
        if (p1 < pm) pm = p1;
        double dvg = 0.0;
#pragma unroll
        for (int b = 0; b < 10; b++) {
          const int p6 = PL(b, 5), p7 = PL(b, 6);
          if (p6 >= 0 && dphix(b, x, y) != 0.0) dvg += m[p6] * (dphix(b, x, y) * ihx1);
          if (p7 >= 0 && dphiy(b, x, y) != 0.0) dvg += m[p7] * (dphiy(b, x, y) * ihy1);
        }
        dv += wgt(i1) * wgt(j1) * (dvg * dvg);
      }
  }
  __shared__ double s1[BS / 64], s2[BS / 64], s3[BS / 64];
  um = wave_max(um);
  pm = wave_min(pm);
  dv = wave_sum(dv);
  if ((threadIdx.x & 63) == 0) { s1[threadIdx.x >> 6] = um; s2[threadIdx.x >> 6] = pm; s3[threadIdx.x >> 6] = dv; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BS / 64; w++) { um = fmax(um, s1[w]); pm = fmin(pm, s2[w]); dv += s3[w]; }
    atomicMax(&sc->umax_bits, (unsigned long long)__double_as_longlong(um));
    atomicMin(&sc->pmin_key, dkey(pm));
    partial[blockIdx.x] = dv;
  }
}

__global__ void k_div_final(DevScal* sc, const double* partial, long long nblocks, double hx1, double hy1,
                            double* out = nullptr) {
  __shared__ double s[BS];
  double a = 0.0;
  for (long long k = threadIdx.x; k < nblocks; k += BS) a += partial[k];
  s[threadIdx.x] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int k = 0; k < BS; k++) tot += s[k];
    sc->totaldiv = sqrt(tot * hx1 * hy1);
    sc->umax = __longlong_as_double((long long)sc->umax_bits);
    sc->pmin = dunkey(sc->pmin_key);
    if (!(isfinite(sc->totaldiv) && isfinite(sc->umax) && isfinite(sc->pmin))) sc->health |= MHD2D_FLAG_NONFINITE;
    if (sc->pmin < 0.0) sc->health |= MHD2D_FLAG_NEG_PRESSURE;
    if (out) { out[0] = sc->umax; out[1] = sc->totaldiv; out[2] = sc->pmin; }
  }
}

__device__ void ic_eval(const ICPar& c, double x, double y, double U[8]) {
  MHD_NOCONTRACT
  const double pi = 4 * atan(1.0);
  double rho = 1, v1 = 0, v2 = 0, v3 = 0, B1 = 0, B2 = 0, B3 = 0, p = 1;
  const double gam = 5.0 / 3.0;
  if (c.ic == MHD2D_IC_ORSZAG_TANG || c.ic == MHD2D_IC_OT_PERTURBED) {
    p = gam;
    rho = gam * gam;
    v1 = -sin(y);
    v2 = sin(x);
    B1 = -sin(y);
    B2 = sin(2.0 * x);
    if (c.ic == MHD2D_IC_OT_PERTURBED) {
      const double amp = c.par[0] / 16.0;
      double du = 0, dv = 0;
      for (int m = 0; m < 16; m++) {
        du += sin(c.kx[m] * x + c.ky[m] * y + c.phu[m]);
        dv += sin(c.kx[m] * x + c.ky[m] * y + c.phv[m]);
      }
      v1 += amp * du;
      v2 += amp * dv;
    }
  } else if (c.ic == MHD2D_IC_VORTEX) {
    const double r2 = x * x + y * y;
    p = 1 - r2 / (8.0 * (pi * pi)) * exp(1 - r2);
    rho = 1;
    v1 = 1 - 1.0 / (2.0 * pi) * y * exp(0.5 * (1 - r2));
    v2 = 1 + 1.0 / (2.0 * pi) * x * exp(0.5 * (1 - r2));
    B1 = -1.0 / (2.0 * pi) * y * exp(0.5 * (1 - r2));
    B2 = 1.0 / (2.0 * pi) * x * exp(0.5 * (1 - r2));
  } else if (c.ic == MHD2D_IC_ROTOR) {
    const double r0 = c.par[0], r1 = c.par[1], spin = c.par[2], xc = c.par[6], yc = c.par[7];
    const double r = sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc));
    const double f = (r1 - r) / (r1 - r0);
    const double rin = c.par[5];
    if (r < r0) { rho = rin; v1 = -spin * (y - yc) / r0; v2 = spin * (x - xc) / r0; }
    else if (r < r1) { rho = 1.0 + (rin - 1.0) * f; v1 = -spin * f * (y - yc) / r; v2 = spin * f * (x - xc) / r; }
    else { rho = 1.0; v1 = 0.0; v2 = 0.0; }
    p = c.par[3];
    B1 = c.par[4];
  } else if (c.ic == MHD2D_IC_BLAST) {                   // init5.txt:3-18, p0.f90:1-14
    p = sqrt(x * x + y * y) <= 0.1 ? 1000.0 : 0.1;
    B1 = 100.0 / sqrt(4 * pi);
  } else if (c.ic == MHD2D_IC_CLOUD_SHOCK) {            // Cloud_Shock_functions.f90:1-117
    const double r = sqrt((x - 1.4) * (x - 1.4) + (y - 0.5) * (y - 0.5));
    if (x <= 1.2) { rho = 3.88968; v3 = -0.05234; p = 14.2641; B3 = 3.9353; }
    else { rho = r >= 0.18 ? 1.0 : 5.0; v1 = -3.3156; p = 0.04; B3 = 1.0; }
    B1 = 1.0;
  } else if (c.ic == MHD2D_IC_CLOUD_SHOCK_2) {          // Cloud_Shock_functions2.f90:1-116
    const double r = sqrt((x - 0.25) * (x - 0.25) + (y - 0.5) * (y - 0.5));
    if (x <= 0.05) { rho = 3.88968; v1 = 11.2536; p = 167.345; B2 = 0.21826182; B3 = -0.21826182; }
    else { rho = r >= 0.15 ? 1.0 : 10.0; p = 1.0; B2 = 0.56418958; B3 = 0.56418958; }
  } else if (c.ic == MHD2D_IC_MAGNETIC_LOOP) {          // init7.txt:3-18, Loop_functions.f90:1-35
    const double r = sqrt(x * x + y * y);
    v1 = 2.0;
    v2 = 1.0;
    if (r < 0.3) { B1 = -1e-3 * y / r; B2 = 1e-3 * x / r; }
  } else if (c.ic == MHD2D_IC_BRIO_WU) {                // BrioWu_functions.f90:1-37
    rho = x < 0 ? 1.0 : 0.1225; p = x < 0 ? 1.0 : 0.1; B1 = 0.75; B2 = x < 0 ? 1.0 : -1.0;
  } else if (c.ic == MHD2D_IC_BRIO_WU_Y) {              // BrioWu_functions.f90:40-76
    rho = y < 0 ? 1.0 : 0.1225; p = y < 0 ? 1.0 : 0.1; B1 = y < 0 ? 1.0 : -1.0; B2 = 0.75;
  } else if (c.ic == MHD2D_IC_RYU_JONES) {              // RyuJones_functions.f90:1-49
    rho = x < 0 ? 1.08 : 1.0; p = x < 0 ? 0.95 : 1.0; v1 = x < 0 ? 1.2 : 0.0;
    B1 = 2.0 / sqrt(4 * pi); B2 = (x < 0 ? 3.6 : 4.0) / sqrt(4 * pi); B3 = 2.0 / sqrt(4 * pi);
  } else if (c.ic == MHD2D_IC_RYU_JONES_Y) {            // RyuJones_functions.f90:52-100
    rho = y < 0 ? 1.08 : 1.0; p = y < 0 ? 0.95 : 1.0; v2 = y < 0 ? 1.2 : 0.0;
    B1 = (y < 0 ? 3.6 : 4.0) / sqrt(4 * pi); B2 = 2.0 / sqrt(4 * pi); B3 = 2.0 / sqrt(4 * pi);
  } else if (c.ic == MHD2D_IC_RIEMANN4) {
    const int q = (x >= c.par[0] ? 0 : 1) + (y >= c.par[1] ? 0 : 2);
    const double* st = &c.par[2 + 6 * q];
    rho = st[0]; v1 = st[1]; v2 = st[2]; p = st[3]; B1 = st[4]; B2 = st[5];
  }
  U[0] = rho;
  U[1] = rho * v1;
  U[2] = rho * v2;
  U[3] = rho * v3;
  if (c.flags & IC_FULL_ENERGY)
    U[4] = p / (gam - 1.0) + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3) + 0.5 * (B1 * B1 + B2 * B2 + B3 * B3);
  else
    U[4] = p / (gam - 1.0) + 0.5 * rho * (v1 * v1 + v2 * v2) + 0.5 * (B1 * B1 + B2 * B2);
  U[5] = B1;
  U[6] = B2;
  U[7] = B3;
}

template <bool MIX>
__global__ __launch_bounds__(BS) void k_init_cells(SWr<MIX> S, Geo g, double xa, double ya,
                                                   const ICPar* __restrict__ icp) {
  MHD_NOCONTRACT
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  const int W = g.Nx + 2;
  if (t >= (long long)W * (g.Ny + 2)) return;
  const int i = (int)(t % W), j = (int)(t / W);
  const long long c = cix(g, i, j);
  if (!(i >= 1 && i <= g.Nx && j >= 1 && j <= g.Ny)) {
    for (int p = 0; p < NPL; p++) S.set(p, c, 0.0);
    return;
  }
  const ICPar& ic = *icp;
  const double Xc = xa + (i - 0.5) * g.hx, Yc = ya + (j - 0.5) * g.hy;
  if (ic.flags & IC_SHOCK) {   // init_data.f90:92-104: P0 state from the centre value
    double U[8];
    ic_eval(ic, Xc, Yc, U);
    for (int p = 0; p < NPL; p++) S.set(p, c, (p < 48 && p % 6 == 0) ? U[p / 6] : 0.0);
    return;
  }
  double acc[8][6];
  for (int n = 0; n < 8; n++)
    for (int b = 0; b < 6; b++) acc[n][b] = 0.0;
  for (int i1 = 0; i1 < 5; i1++)
    for (int j1 = 0; j1 < 5; j1++) {
      double U[8];
      ic_eval(ic, Xc + g.hx1 * lam(i1), Yc + g.hy1 * lam(j1), U);
      for (int b = 0; b < 6; b++) {
        const double ww = wgt(i1) * wgt(j1) * phi(b, lam(i1), lam(j1));
        for (int n = 0; n < 8; n++) acc[n][b] += ww * U[n];
      }
    }
  for (int n = 0; n < 8; n++)
    for (int b = 0; b < 10; b++) {
      const int p = PL(b, n);
      if (p < 0) continue;
      S.set(p, c, b < 6 ? (0.25 * acc[n][b]) / mmass(b) : 0.0);   // init_data.f90:43 leaves modes 7-10 at 0
    }
}

__global__ __launch_bounds__(BS) void k_init_faces(double* __restrict__ F, Geo g, double xa, double ya,
                                                   const ICPar* __restrict__ icp) {
  MHD_NOCONTRACT
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  const long long nx = (long long)(g.Nx + 1) * g.Ny, ny = (long long)g.Nx * (g.Ny + 1);
  if (t >= nx + ny) return;
  const ICPar& ic = *icp;
  const long long pl = g.plane;
  const double mmE[3] = {1.0, 1.0 / 3.0, 4.0 / 45.0};
  const bool shock = ic.flags & IC_SHOCK;   // init_data.f90:106-118: face-centre values
  if (t < nx) {
    const int i = (int)(t % (g.Nx + 1)), j = 1 + (int)(t / (g.Nx + 1));
    const double Yc = ya + (j - 0.5) * g.hy;
    double a[3] = {0, 0, 0};
    if (shock) {
      double U[8];
      ic_eval(ic, xa + i * g.hx, Yc, U);
      a[0] = U[5];
    } else {
      for (int q = 0; q < 5; q++) {
        double U[8];
        ic_eval(ic, xa + i * g.hx, Yc + g.hy1 * lam(q), U);
        const double e[3] = {1.0, lam(q), P2(lam(q))};
        for (int k = 0; k < 3; k++) a[k] += 0.5 * wgt(q) * e[k] * U[5];
      }
      for (int k = 0; k < 3; k++) a[k] /= mmE[k];
    }
    for (int k = 0; k < 3; k++) F[k * pl + cix(g, i, j)] = a[k];
  } else {
    const long long u = t - nx;
    const int i = 1 + (int)(u % g.Nx), j = (int)(u / g.Nx);
    const double Xc = xa + (i - 0.5) * g.hx;
    double a[3] = {0, 0, 0};
    if (shock) {
      double U[8];
      ic_eval(ic, Xc, ya + j * g.hy, U);
      a[0] = U[6];
    } else {
      for (int q = 0; q < 5; q++) {
        double U[8];
        ic_eval(ic, Xc + g.hx1 * lam(q), ya + j * g.hy, U);
        const double e[3] = {1.0, lam(q), P2(lam(q))};
        for (int k = 0; k < 3; k++) a[k] += 0.5 * wgt(q) * e[k] * U[6];
      }
      for (int k = 0; k < 3; k++) a[k] /= mmE[k];
    }
    for (int k = 0; k < 3; k++) F[(3 + k) * pl + cix(g, i, j)] = a[k];
  }
}

// calculate_Az.f90:13-26 with Loop_functions.f90:37-51 (Az of the magnetic loop)
__device__ double loop_az(double x, double y) {
  MHD_NOCONTRACT
  const double r = sqrt(x * x + y * y);
  return r < 0.3 ? 1e-3 * (0.3 - r) : 0.0;
}

__global__ __launch_bounds__(BS) void k_calc_az(double* __restrict__ F, Geo g, double xa, double ya) {
  MHD_NOCONTRACT
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  const long long nx = (long long)(g.Nx + 1) * g.Ny, ny = (long long)g.Nx * (g.Ny + 1);
  if (t >= nx + ny) return;
  if (t < nx) {
    const int i = (int)(t % (g.Nx + 1)), j = 1 + (int)(t / (g.Nx + 1));
    const double X = xa + i * g.hx;
    F[cix(g, i, j)] = -(1.0 / g.hy) * (loop_az(X, ya + j * g.hy) - loop_az(X, ya + (j - 1) * g.hy));
  } else {
    const long long u = t - nx;
    const int i = 1 + (int)(u % g.Nx), j = (int)(u / g.Nx);
    const double Y = ya + j * g.hy;
    F[3 * g.plane + cix(g, i, j)] = (1.0 / g.hx) * (loop_az(xa + i * g.hx, Y) - loop_az(xa + (i - 1) * g.hx, Y));
  }
}

// calculate_L2error.f90:7-33: per-block partial sums of the squared error at
// the 5x5 Gauss points against the vortex (init1.txt) translated by (t20, t20)
template <bool MIX>
__global__ __launch_bounds__(BS) void k_l2err(SRd<MIX> U, Geo g, double xa, double ya, double t20,
                                              const ICPar* __restrict__ icp, double* __restrict__ partial) {
  MHD_NOCONTRACT
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  double e2[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (t < (long long)g.Nx * g.Ny) {
    const int i = 1 + (int)(t % g.Nx), j = 1 + (int)(t / g.Nx);
    const long long c = cix(g, i, j);
    const double Xc = xa + (i - 0.5) * g.hx, Yc = ya + (j - 0.5) * g.hy;
    for (int i1 = 0; i1 < 5; i1++)
      for (int j1 = 0; j1 < 5; j1++) {
        double ex[8];
        ic_eval(*icp, Xc + g.hx1 * lam(i1) - t20, Yc + g.hy1 * lam(j1) - t20, ex);
#pragma unroll
        for (int n = 0; n < 8; n++) {
          double u = 0.0;
#pragma unroll
          for (int b = 0; b < 10; b++) {
            const int p = PL(b, n);
            if (p >= 0) u = u + U(p, c) * phi(b, lam(i1), lam(j1));
          }
          const double e = u - ex[n];
          e2[n] = e2[n] + wgt(i1) * wgt(j1) * (e * e);
        }
      }
  }
  __shared__ double sm[8][BS / 64];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    const double v = wave_sum(e2[n]);
    if ((threadIdx.x & 63) == 0) sm[n][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    double a = 0.0;
    for (int w = 0; w < BS / 64; w++) a += sm[threadIdx.x][w];
    partial[(long long)threadIdx.x * gridDim.x + blockIdx.x] = a;
  }
}

// fixed-order sum of the per-block partials, then (L2 hx1 hy1)^(1/2)
__global__ void k_l2err_final(const double* __restrict__ partial, long long nblocks, double hx1, double hy1,
                              double* __restrict__ out) {
  __shared__ double s[BS];
  for (int n = 0; n < 8; n++) {
    double a = 0.0;
    for (long long k = threadIdx.x; k < nblocks; k += BS) a += partial[n * nblocks + k];
    s[threadIdx.x] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = 0.0;
      for (int k = 0; k < BS; k++) tot += s[k];
      out[n] = sqrt(tot * hx1 * hy1);
    }
    __syncthreads();
  }
}

// mixed-form storage <-> 52 FP64 planes (the FP32 planes round to nearest)
template <bool TO_MIX>
__global__ __launch_bounds__(BS) void k_pack(double* __restrict__ full, SWr<true> v) {
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  if (t >= (long long)NPL * v.pl) return;
  const int p = (int)(t / v.pl);
  const long long c = t % v.pl;
  if (TO_MIX) v.set(p, c, full[t]);
  else full[t] = v(p, c);
}

}  // namespace

namespace {

int launch_setbc(mhd2d_ctx* ctx, double* U, const DevScal* sc) {
  const Geo& g = ctx->g;
  const long long ring = 2LL * (g.Nx + 2) + 2LL * g.Ny;
  Timed tm(ctx, KC_SETBC);
  return with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_setbc<MIX>), dim3(nblk(ring * NPL)), dim3(BS), 0, ctx->stream, SWr<MIX>{U, g.plane}, g, NPL,
                       sc);
    LAUNCHCHK(ctx);
    return 0;
  });
}

// the mixed form's upload/download staging buffer (52 FP64 planes, zeroed once)
int mix_stage(mhd2d_ctx* ctx) {
  if (ctx->stage) return 0;
  const size_t n = (size_t)NPL * ctx->g.plane * sizeof(double);
  HIPCHK(ctx, hipMalloc((void**)&ctx->stage, n));
  HIPCHK(ctx, hipMemsetAsync(ctx->stage, 0, n, ctx->stream));
  return 0;
}

int launch_pack(mhd2d_ctx* ctx, bool to_mix, double* mixbuf) {
  const long long n = (long long)NPL * ctx->g.plane;
  const SWr<true> v{mixbuf, ctx->g.plane};
  if (to_mix) hipLaunchKernelGGL(k_pack<true>, dim3(nblk(n)), dim3(BS), 0, ctx->stream, ctx->stage, v);
  else hipLaunchKernelGGL(k_pack<false>, dim3(nblk(n)), dim3(BS), 0, ctx->stream, ctx->stage, v);
  LAUNCHCHK(ctx);
  return 0;
}

long long geo_plane(const Geo& g) { return (long long)(g.Ny + 2) * g.P; }

int ctx_check(mhd2d_ctx* ctx) {
  if (!ctx) return -1;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* mhd2d_last_error(const mhd2d_ctx* ctx) { return ctx ? ctx->err : "null context"; }

int mhd2d_create(const mhd2d_config* cfg, mhd2d_ctx** out) {
  if (!cfg || !out) return -1;
  *out = nullptr;
  if (cfg->Nx < 2 || cfg->Ny < 2 || cfg->flux_type < 1 || cfg->flux_type > 4) return -1;
  if (!(cfg->xb > cfg->xa) || !(cfg->yb > cfg->ya)) return -1;
  const int bcs[4] = {cfg->bcR, cfg->bcL, cfg->bcU, cfg->bcD};
  for (int b : bcs)
    if (b != 1 && b != 2) return -1;
  if (cfg->precision != MHD2D_FP64 && cfg->precision != MHD2D_MIXED) return -1;
  if (cfg->indicator < MHD2D_IND_NONE || cfg->indicator > MHD2D_IND_MLP) return -1;
  mhd2d_ctx* ctx = (mhd2d_ctx*)calloc(1, sizeof(mhd2d_ctx));
  if (!ctx) return -1;
  new (&ctx->prof_ev) decltype(ctx->prof_ev)();
  new (&ctx->ev_pool) decltype(ctx->ev_pool)();
  new (&ctx->graphs) decltype(ctx->graphs)();
  ctx->cfg = *cfg;
  if (ctx->cfg.CFL <= 0) ctx->cfg.CFL = 0.15;
  Geo& g = ctx->g;
  g.Nx = cfg->Nx;
  g.Ny = cfg->Ny;
  g.P = ((cfg->Nx + 2 + OFF) + 15) / 16 * 16;
  g.bcR = cfg->bcR; g.bcL = cfg->bcL; g.bcU = cfg->bcU; g.bcD = cfg->bcD;
  g.plane = geo_plane(g);
  g.hx = (cfg->xb - cfg->xa) / cfg->Nx;      // init_data.f90:21-24
  g.hy = (cfg->yb - cfg->ya) / cfg->Ny;
  g.hx1 = 0.5 * g.hx;
  g.hy1 = 0.5 * g.hy;
  hipError_t e = hipGetDevice(&ctx->device);
  if (e != hipSuccess) { free(ctx); return -2; }
  const size_t sbytes = (size_t)state_bytes(cfg->precision == MHD2D_MIXED, g.plane);
  const size_t fbytes = (size_t)face_planes(ctx) * g.plane * sizeof(double);
  const size_t xbytes = (size_t)40 * g.plane * sizeof(double);
  const size_t ebytes = (size_t)5 * g.plane * sizeof(double);
  ctx->nblk_cells = nblk((long long)g.Nx * g.Ny);
  struct { void** p; size_t n; } al[] = {
      {(void**)&ctx->S, sbytes},   {(void**)&ctx->U0, sbytes},  {(void**)&ctx->UI, sbytes},
      {(void**)&ctx->UII, sbytes}, {(void**)&ctx->T, sbytes},   {(void**)&ctx->du, sbytes},
      {(void**)&ctx->F, fbytes},   {(void**)&ctx->F0, fbytes},  {(void**)&ctx->FI, fbytes},
      {(void**)&ctx->FII, fbytes}, {(void**)&ctx->Ft, fbytes},  {(void**)&ctx->dEz, fbytes},
      {(void**)&ctx->FX, xbytes},  {(void**)&ctx->FY, xbytes},  {(void**)&ctx->EZX, ebytes},
      {(void**)&ctx->EZY, ebytes}, {(void**)&ctx->EZV, (size_t)g.plane * sizeof(double)},
      {(void**)&ctx->trouble, (size_t)g.Nx * g.Ny * sizeof(int)},
      {(void**)&ctx->tmask, (size_t)g.Nx * g.Ny * sizeof(int)},
      {(void**)&ctx->tlist, (size_t)g.Nx * g.Ny * sizeof(int)},
      {(void**)&ctx->blist, (size_t)g.Nx * g.Ny * sizeof(int)},
      {(void**)&ctx->screen, (size_t)g.Nx * g.Ny * sizeof(int)},
      {(void**)&ctx->pplist, (size_t)g.Nx * g.Ny * sizeof(int)},
      {(void**)&ctx->lflag, (size_t)g.Nx * g.Ny * sizeof(int)},
      {(void**)&ctx->trec, (size_t)32 * g.Nx * g.Ny * sizeof(double)},
      {(void**)&ctx->prob, (size_t)g.Nx * g.Ny * sizeof(float)},
      {(void**)&ctx->sc, sizeof(DevScal)},
      {(void**)&ctx->partial, (size_t)ctx->nblk_cells * sizeof(double)},
      {(void**)&ctx->icp, sizeof(ICPar)},
  };
  for (auto& a : al) {
    e = hipMalloc(a.p, a.n);
    if (e != hipSuccess) {
      snprintf(ctx->err, sizeof ctx->err, "hipMalloc(%zu) failed: %s", a.n, hipGetErrorString(e));
      mhd2d_destroy(ctx);
      return -2;
    }
  }
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking);
  for (hipEvent_t* ev : {&ctx->ev_fork, &ctx->ev_faces, &ctx->ev_join, &ctx->ev_cell, &ctx->ev_det})
    if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
  if (e != hipSuccess) { mhd2d_destroy(ctx); return -2; }
  // zero everything once (halo and padding included)
  for (auto& a : al) (void)hipMemsetAsync(*a.p, 0, a.n, ctx->stream);
  DevScal h{};
  h.CFL = ctx->cfg.CFL;
  h.pmin = 1e6;
  h.active = 1;
  (void)hipMemcpyAsync(ctx->sc, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream);
  e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) { mhd2d_destroy(ctx); return -2; }
  *out = ctx;
  return 0;
}

int mhd2d_destroy(mhd2d_ctx* ctx) {
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  for (auto& pe : ctx->prof_ev) { (void)hipEventDestroy(pe.second.first); (void)hipEventDestroy(pe.second.second); }
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  for (auto& sg : ctx->graphs) (void)hipGraphExecDestroy(sg.exec);
  ctx->graphs.~vector();
  ctx->prof_ev.~vector();
  ctx->ev_pool.~vector();
  void* ps[] = {ctx->stage, ctx->S, ctx->U0, ctx->UI, ctx->UII, ctx->T, ctx->du, ctx->F, ctx->F0, ctx->FI, ctx->FII,
                ctx->Ft, ctx->dEz, ctx->FX, ctx->FY, ctx->EZX, ctx->EZY, ctx->EZV, ctx->trouble, ctx->tmask,
                ctx->tlist, ctx->blist, ctx->screen, ctx->pplist, ctx->lflag, ctx->trec, ctx->prob, ctx->mlpw, ctx->sc,
                ctx->partial, ctx->icp};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  for (hipEvent_t ev : {ctx->ev_fork, ctx->ev_faces, ctx->ev_join, ctx->ev_cell, ctx->ev_det})
    if (ev) (void)hipEventDestroy(ev);
  free(ctx);
  return 0;
}

int mhd2d_sync(mhd2d_ctx* ctx) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

// host uh(0:Nx+1,0:Ny+1,10,8) <-> device planes; Bx(0:Nx,1:Ny,3), By(1:Nx,0:Ny,3)
// mixed form: faces written from outside the CT update start their
// double-double low parts again at zero
static int clear_face_lo(mhd2d_ctx* ctx, double* F, bool x, bool y) {
  if (ctx->cfg.precision != MHD2D_MIXED) return 0;
  const size_t pb = (size_t)ctx->g.plane * sizeof(double);
  if (x) HIPCHK(ctx, hipMemsetAsync(F + FACE_LO_X * ctx->g.plane, 0, pb, ctx->stream));
  if (y) HIPCHK(ctx, hipMemsetAsync(F + FACE_LO_Y * ctx->g.plane, 0, pb, ctx->stream));
  return 0;
}

int mhd2d_upload(mhd2d_ctx* ctx, const double* uh, const double* Bx, const double* By) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  const Geo& g = ctx->g;
  const size_t W = g.Nx + 2, H = g.Ny + 2, P = g.P;
  const bool mix = ctx->cfg.precision == MHD2D_MIXED;
  if (uh && mix && (rc = mix_stage(ctx))) return rc;
  double* dst = mix ? ctx->stage : ctx->S;
  if (uh) {
    for (int n = 0; n < 8; n++)
      for (int b = 0; b < 10; b++) {
        const int p = PL(b, n);
        if (p < 0) continue;
        const double* src = uh + W * H * (b + 10 * n);
        HIPCHK(ctx, hipMemcpy2DAsync(dst + p * g.plane + OFF, P * 8, src, W * 8, W * 8, H, hipMemcpyHostToDevice,
                                     ctx->stream));
      }
    if (mix && (rc = launch_pack(ctx, true, ctx->S))) return rc;
  }
  if (Bx)
    for (int k = 0; k < 3; k++)
      HIPCHK(ctx, hipMemcpy2DAsync(ctx->F + k * g.plane + cix(g, 0, 1), P * 8, Bx + (size_t)(g.Nx + 1) * g.Ny * k,
                                   (g.Nx + 1) * 8, (g.Nx + 1) * 8, g.Ny, hipMemcpyHostToDevice, ctx->stream));
  if (By)
    for (int k = 0; k < 3; k++)
      HIPCHK(ctx, hipMemcpy2DAsync(ctx->F + (3 + k) * g.plane + cix(g, 1, 0), P * 8, By + (size_t)g.Nx * (g.Ny + 1) * k,
                                   g.Nx * 8, g.Nx * 8, g.Ny + 1, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = clear_face_lo(ctx, ctx->F, Bx != nullptr, By != nullptr))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->lh_done = false;
  return 0;
}

static int download_planes(mhd2d_ctx* ctx, double* dev, double* uh) {
  const Geo& g = ctx->g;
  const size_t W = g.Nx + 2, H = g.Ny + 2, P = g.P;
  if (ctx->cfg.precision == MHD2D_MIXED) {
    int rc = mix_stage(ctx);
    if (rc || (rc = launch_pack(ctx, false, dev))) return rc;
    dev = ctx->stage;
  }
  for (int n = 0; n < 8; n++)
    for (int b = 0; b < 10; b++) {
      double* dst = uh + W * H * (b + 10 * n);
      const int p = PL(b, n);
      if (p < 0) { memset(dst, 0, W * H * sizeof(double)); continue; }
      HIPCHK(ctx, hipMemcpy2DAsync(dst, W * 8, dev + p * g.plane + OFF, P * 8, W * 8, H, hipMemcpyDeviceToHost,
                                   ctx->stream));
    }
  return 0;
}

static int download_faces(mhd2d_ctx* ctx, const double* dev, double* Bx, double* By) {
  const Geo& g = ctx->g;
  const size_t P = g.P;
  if (Bx)
    for (int k = 0; k < 3; k++)
      HIPCHK(ctx, hipMemcpy2DAsync(Bx + (size_t)(g.Nx + 1) * g.Ny * k, (g.Nx + 1) * 8, dev + k * g.plane + cix(g, 0, 1),
                                   P * 8, (g.Nx + 1) * 8, g.Ny, hipMemcpyDeviceToHost, ctx->stream));
  if (By)
    for (int k = 0; k < 3; k++)
      HIPCHK(ctx, hipMemcpy2DAsync(By + (size_t)g.Nx * (g.Ny + 1) * k, g.Nx * 8, dev + (3 + k) * g.plane + cix(g, 1, 0),
                                   P * 8, g.Nx * 8, g.Ny + 1, hipMemcpyDeviceToHost, ctx->stream));
  return 0;
}

int mhd2d_download(mhd2d_ctx* ctx, double* uh, double* Bx, double* By) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if (uh && (rc = download_planes(ctx, ctx->S, uh))) return rc;
  if ((rc = download_faces(ctx, ctx->F, Bx, By))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int mhd2d_debug_get(mhd2d_ctx* ctx, int32_t which, double* host) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  const Geo& g = ctx->g;
  if (!host) return fail(ctx, -1, "debug_get: null host pointer");
  switch (which) {
    case MHD2D_DU: rc = download_planes(ctx, ctx->du, host); break;
    case MHD2D_DEZ1: rc = download_faces(ctx, ctx->dEz, host, nullptr); break;
    case MHD2D_DEZ2: rc = download_faces(ctx, ctx->dEz, nullptr, host); break;
    case MHD2D_EZVERTEX:
      HIPCHK(ctx, hipMemcpy2DAsync(host, (g.Nx + 1) * 8, ctx->EZV + cix(g, 0, 0), g.P * 8, (g.Nx + 1) * 8, g.Ny + 1,
                                   hipMemcpyDeviceToHost, ctx->stream));
      break;
    case MHD2D_TROUBLE: {
      std::vector<int> tmp((size_t)g.Nx * g.Ny);
      HIPCHK(ctx, hipMemcpyAsync(tmp.data(), ctx->trouble, tmp.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      for (size_t k = 0; k < tmp.size(); k++) host[k] = tmp[k];
      return 0;
    }
    default: return fail(ctx, -1, "debug_get: unknown array %d", which);
  }
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

static uint64_t splitmix64(uint64_t* x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int mhd2d_init_data(mhd2d_ctx* ctx, int32_t ic, const double* extra, int32_t nextra) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  ICPar h{};
  h.ic = ic;
  const double pi = 4 * atan(1.0);
  if (ic == MHD2D_IC_OT_PERTURBED) {
    h.par[0] = nextra > 0 ? extra[0] : 1e-2;
    uint64_t st = nextra > 1 ? (uint64_t)extra[1] : 20261203ull;
    for (int m = 0; m < 16; m++) {
      h.kx[m] = 1 + (m % 4);
      h.ky[m] = 1 + (m / 4);
      h.phu[m] = 2 * pi * ((double)(splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0));
      h.phv[m] = 2 * pi * ((double)(splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0));
    }
  } else if (ic == MHD2D_IC_ROTOR) {
    const double def[8] = {0.1, 0.115, 1.0, 0.5, 2.4 / sqrt(4 * pi), 10.0, 0.5, 0.5};
    for (int k = 0; k < 8; k++) h.par[k] = k < nextra ? extra[k] : def[k];
  } else if (ic == MHD2D_IC_RIEMANN4) {
    for (int k = 0; k < 26 && k < nextra; k++) h.par[k] = extra[k];
  } else {
    mhd2d_ic_info info;
    if (mhd2d_ic_defaults(ic, &info)) return fail(ctx, -1, "init_data: unknown initial condition %d", ic);
    if (info.shock) h.flags |= IC_SHOCK;
    if (ic != MHD2D_IC_ORSZAG_TANG && ic != MHD2D_IC_VORTEX) h.flags |= IC_FULL_ENERGY;
  }
  const Geo& g = ctx->g;
  HIPCHK(ctx, hipMemcpyAsync(ctx->icp, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
  const long long nc = (long long)(g.Nx + 2) * (g.Ny + 2);
  const long long nf = (long long)(g.Nx + 1) * g.Ny + (long long)g.Nx * (g.Ny + 1);
  rc = with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_init_cells<MIX>), dim3(nblk(nc)), dim3(BS), 0, ctx->stream, SWr<MIX>{ctx->S, g.plane}, g,
                       ctx->cfg.xa, ctx->cfg.ya, (const ICPar*)ctx->icp);
    return 0;
  });
  hipLaunchKernelGGL(k_init_faces, dim3(nblk(nf)), dim3(BS), 0, ctx->stream, ctx->F, g, ctx->cfg.xa, ctx->cfg.ya,
                     (const ICPar*)ctx->icp);
  LAUNCHCHK(ctx);
  if ((rc = clear_face_lo(ctx, ctx->F, true, true))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->lh_done = false;
  return 0;
}

// ---------------------------------------------------------------------------
// checkpoint (format in include/mhd2d.h)
}  // extern "C"
namespace {

// the mixed form's face low parts (Bx mode 1 on x faces, then By mode 1 on
// y faces, Fortran order) between host and the current face buffer
int face_lo_io(mhd2d_ctx* ctx, double* host, bool to_device) {
  const Geo& g = ctx->g;
  const size_t P = g.P, nx = (size_t)(g.Nx + 1) * g.Ny;
  double* dx = ctx->F + FACE_LO_X * g.plane + cix(g, 0, 1);
  double* dy = ctx->F + FACE_LO_Y * g.plane + cix(g, 1, 0);
  if (to_device) {
    HIPCHK(ctx, hipMemcpy2DAsync(dx, P * 8, host, (g.Nx + 1) * 8, (g.Nx + 1) * 8, g.Ny, hipMemcpyHostToDevice,
                                 ctx->stream));
    HIPCHK(ctx, hipMemcpy2DAsync(dy, P * 8, host + nx, g.Nx * 8, g.Nx * 8, g.Ny + 1, hipMemcpyHostToDevice,
                                 ctx->stream));
  } else {
    HIPCHK(ctx, hipMemcpy2DAsync(host, (g.Nx + 1) * 8, dx, P * 8, (g.Nx + 1) * 8, g.Ny, hipMemcpyDeviceToHost,
                                 ctx->stream));
    HIPCHK(ctx, hipMemcpy2DAsync(host + nx, g.Nx * 8, dy, P * 8, g.Nx * 8, g.Ny + 1, hipMemcpyDeviceToHost,
                                 ctx->stream));
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

struct CkHeader {
  char magic[8];
  uint32_t version, header_bytes;
  int32_t Nx, Ny, degree, nplanes, flux_type, precision, bcR, bcL, bcU, bcD;
  double xa, xb, ya, yb, t, dt;
  int64_t count;
  uint64_t payload_bytes;
  char pad[128 - 8 - 8 - 40 - 48 - 16];
};
static_assert(sizeof(CkHeader) == 128, "checkpoint header is 128 bytes");

// CRC-32, reflected polynomial 0xEDB88320 (IEEE 802.3), slicing by 8
struct Crc32 {
  uint32_t t[8][256];
  Crc32() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      t[0][i] = c;
    }
    for (uint32_t i = 0; i < 256; i++)
      for (int s = 1; s < 8; s++) t[s][i] = (t[s - 1][i] >> 8) ^ t[0][t[s - 1][i] & 0xFF];
  }
  uint32_t update(uint32_t crc, const void* data, size_t n) const {
    const unsigned char* p = (const unsigned char*)data;
    uint32_t c = ~crc;
    while (n >= 8) {
      uint32_t lo, hi;
      memcpy(&lo, p, 4);
      memcpy(&hi, p + 4, 4);
      lo ^= c;
      c = t[7][lo & 0xFF] ^ t[6][(lo >> 8) & 0xFF] ^ t[5][(lo >> 16) & 0xFF] ^ t[4][lo >> 24] ^ t[3][hi & 0xFF] ^
          t[2][(hi >> 8) & 0xFF] ^ t[1][(hi >> 16) & 0xFF] ^ t[0][hi >> 24];
      p += 8;
      n -= 8;
    }
    while (n--) c = t[0][(c ^ *p++) & 0xFF] ^ (c >> 8);
    return ~c;
  }
};
const Crc32& crc32_tables() {
  static const Crc32 tab;
  return tab;
}

// active planes in device plane order (the order PL() assigns)
void plane_list(int (&bn)[52][2]) {
  for (int n = 0; n < 8; n++)
    for (int b = 0; b < 10; b++) {
      const int p = PL(b, n);
      if (p >= 0) { bn[p][0] = b; bn[p][1] = n; }
    }
}

}  // namespace
extern "C" {

int mhd2d_checkpoint_save(mhd2d_ctx* ctx, const char* path) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if (!path) return fail(ctx, -1, "checkpoint_save: null path");
  const Geo& g = ctx->g;
  const size_t W = g.Nx + 2, H = g.Ny + 2;
  const size_t nbx = (size_t)(g.Nx + 1) * g.Ny * 3, nby = (size_t)g.Nx * (g.Ny + 1) * 3;
  const bool lo = ctx->cfg.precision == MHD2D_MIXED;   // version 2: the faces' low parts follow
  const size_t nlo = lo ? (nbx + nby) / 3 : 0;
  std::vector<double> uh(W * H * 80), pay(52 * W * H + nbx + nby + nlo);
  if ((rc = mhd2d_download(ctx, uh.data(), pay.data() + 52 * W * H, pay.data() + 52 * W * H + nbx))) return rc;
  if (lo && (rc = face_lo_io(ctx, pay.data() + 52 * W * H + nbx + nby, false))) return rc;
  int bn[52][2];
  plane_list(bn);
  for (int p = 0; p < 52; p++)
    memcpy(pay.data() + p * W * H, uh.data() + W * H * (bn[p][0] + 10 * bn[p][1]), W * H * sizeof(double));
  CkHeader h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, "MHD2DCK1", 8);
  h.version = lo ? 2 : 1;
  h.header_bytes = sizeof h;
  h.Nx = g.Nx; h.Ny = g.Ny; h.degree = 2; h.nplanes = 52;
  h.flux_type = ctx->cfg.flux_type; h.precision = ctx->cfg.precision;
  h.bcR = g.bcR; h.bcL = g.bcL; h.bcU = g.bcU; h.bcD = g.bcD;
  h.xa = ctx->cfg.xa; h.xb = ctx->cfg.xb; h.ya = ctx->cfg.ya; h.yb = ctx->cfg.yb;
  // a positive status only reports the sticky numerical flags: the state is
  // still saved (a flagged run is the one most worth keeping)
  if ((rc = mhd2d_get_time(ctx, &h.t, &h.dt, &h.count)) < 0) return rc;
  h.payload_bytes = pay.size() * sizeof(double);
  const Crc32& C = crc32_tables();
  uint32_t crc = C.update(0, &h, sizeof h);
  crc = C.update(crc, pay.data(), h.payload_bytes);
  FILE* f = fopen(path, "wb");
  if (!f) return fail(ctx, -4, "checkpoint_save: cannot open %s", path);
  bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(pay.data(), 1, h.payload_bytes, f) == h.payload_bytes &&
            fwrite(&crc, 4, 1, f) == 1;
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(ctx, -4, "checkpoint_save: write to %s failed", path);
  return 0;
}

int mhd2d_checkpoint_load(mhd2d_ctx* ctx, const char* path) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if (!path) return fail(ctx, -1, "checkpoint_load: null path");
  const Geo& g = ctx->g;
  const size_t W = g.Nx + 2, H = g.Ny + 2;
  const size_t nbx = (size_t)(g.Nx + 1) * g.Ny * 3, nby = (size_t)g.Nx * (g.Ny + 1) * 3;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(ctx, -4, "checkpoint_load: cannot open %s", path);
  CkHeader h;
  std::vector<double> pay;
  uint32_t crc_file = 0;
  const char* why = nullptr;
  if (fread(&h, sizeof h, 1, f) != 1) why = "short header";
  else if (memcmp(h.magic, "MHD2DCK1", 8) || (h.version != 1 && h.version != 2) || h.header_bytes != sizeof h)
    why = "not a version-1 or version-2 checkpoint";
  else if (h.Nx != g.Nx || h.Ny != g.Ny || h.degree != 2 || h.nplanes != 52) why = "grid differs from the context";
  else if (h.xa != ctx->cfg.xa || h.xb != ctx->cfg.xb || h.ya != ctx->cfg.ya || h.yb != ctx->cfg.yb)
    why = "domain differs from the context";
  else if (h.bcR != g.bcR || h.bcL != g.bcL || h.bcU != g.bcU || h.bcD != g.bcD)
    why = "boundary conditions differ from the context";
  else if (h.flux_type != ctx->cfg.flux_type) why = "flux_type differs from the context";
  else if (h.payload_bytes != (52 * W * H + nbx + nby + (h.version == 2 ? (nbx + nby) / 3 : 0)) * sizeof(double))
    why = "payload size differs";
  if (!why) {
    pay.resize(h.payload_bytes / sizeof(double));
    if (fread(pay.data(), 1, h.payload_bytes, f) != h.payload_bytes || fread(&crc_file, 4, 1, f) != 1)
      why = "truncated file";
  }
  fclose(f);
  if (why) return fail(ctx, -4, "checkpoint_load: %s: %s", path, why);
  const Crc32& C = crc32_tables();
  const uint32_t crc = C.update(C.update(0, &h, sizeof h), pay.data(), h.payload_bytes);
  if (crc != crc_file) return fail(ctx, -4, "checkpoint_load: %s: CRC-32 mismatch (file %08x, data %08x)", path, crc_file, crc);
  std::vector<double> uh(W * H * 80, 0.0);
  int bn[52][2];
  plane_list(bn);
  for (int p = 0; p < 52; p++)
    memcpy(uh.data() + W * H * (bn[p][0] + 10 * bn[p][1]), pay.data() + p * W * H, W * H * sizeof(double));
  if ((rc = mhd2d_upload(ctx, uh.data(), pay.data() + 52 * W * H, pay.data() + 52 * W * H + nbx))) return rc;
  if (h.version == 2 && ctx->cfg.precision == MHD2D_MIXED &&
      (rc = face_lo_io(ctx, pay.data() + 52 * W * H + nbx + nby, true)))
    return rc;
  return mhd2d_set_time(ctx, h.t, h.count);
}

int mhd2d_ic_defaults(int32_t ic, mhd2d_ic_info* out) {
  if (!out) return -1;
  const double pi = 4 * atan(1.0);
  struct Row { int ic; double xa, xb, ya, yb, tend, M; int bc, shock; };
  // initN.txt: domain (xa..yb), tend, M, boundary type (all four sides) and shock
  const Row rows[] = {
      {MHD2D_IC_VORTEX, -10, 10, -10, 10, 0, 1000000, 1, 0},                  // init1.txt:41-53
      {MHD2D_IC_ORSZAG_TANG, 0, 2 * pi, 0, 2 * pi, 3, 10, 1, 0},               // init3.txt
      {MHD2D_IC_ROTOR, 0, 1, 0, 1, 0, 10, 1, 0},                               // init4.txt
      {MHD2D_IC_BLAST, -0.5, 0.5, -0.5, 0.5, 0.01, 10, 2, 1},                  // init5.txt
      {MHD2D_IC_CLOUD_SHOCK, 0, 2, 0, 1, 0.6, 10, 2, 1},                       // init6.txt
      {MHD2D_IC_CLOUD_SHOCK_2, 0, 1, 0, 1, 0.06, 10, 2, 1},                    // init6b.txt
      {MHD2D_IC_MAGNETIC_LOOP, -1, 1, -0.5, 0.5, 2, 10, 1, 0},                 // init7.txt
      {MHD2D_IC_BRIO_WU, -0.5, 0.8, 0, 0.1, 0, 1, 2, 0},                       // init8.txt
      {MHD2D_IC_BRIO_WU_Y, 0, 0.1, -0.5, 0.8, 0.2, 1, 2, 1},                   // init8b.txt
      {MHD2D_IC_RYU_JONES, -0.5, 0.5, 0, 0.1, 0.2, 1, 2, 0},                   // init9.txt
      {MHD2D_IC_RYU_JONES_Y, 0, 0.1, -0.5, 0.5, 0.2, 1, 2, 0},                 // init9b.txt
      {MHD2D_IC_OT_PERTURBED, 0, 2 * pi, 0, 2 * pi, 3, 50, 1, 0},
      {MHD2D_IC_RIEMANN4, 0, 1, 0, 1, 0.2, 10, 2, 0},
  };
  for (const Row& r : rows)
    if (r.ic == ic) {
      *out = mhd2d_ic_info{r.xa, r.xb, r.ya, r.yb, r.tend, r.M, r.bc, r.bc, r.bc, r.bc, r.shock, 0};
      return 0;
    }
  return -1;
}

int mhd2d_calculate_az(mhd2d_ctx* ctx) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  const Geo& g = ctx->g;
  const long long nf = (long long)(g.Nx + 1) * g.Ny + (long long)g.Nx * (g.Ny + 1);
  hipLaunchKernelGGL(k_calc_az, dim3(nblk(nf)), dim3(BS), 0, ctx->stream, ctx->F, g, ctx->cfg.xa, ctx->cfg.ya);
  LAUNCHCHK(ctx);
  if ((rc = clear_face_lo(ctx, ctx->F, true, true))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->lh_done = false;
  return 0;
}

int mhd2d_calculate_l2error(mhd2d_ctx* ctx, double t20, double L2[8]) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if (!L2) return fail(ctx, -1, "calculate_l2error: null output");
  const Geo& g = ctx->g;
  ICPar h{};
  h.ic = MHD2D_IC_VORTEX;
  const long long nb = ctx->nblk_cells;
  double* buf = nullptr;
  HIPCHK(ctx, hipMalloc((void**)&buf, (size_t)(8 * nb + 8) * sizeof(double)));
  HIPCHK(ctx, hipMemcpyAsync(ctx->icp, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
  rc = with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_l2err<MIX>), dim3((unsigned)nb), dim3(BS), 0, ctx->stream, SRd<MIX>{ctx->S, g.plane}, g,
                       ctx->cfg.xa, ctx->cfg.ya, t20, (const ICPar*)ctx->icp, buf);
    return 0;
  });
  hipLaunchKernelGGL(k_l2err_final, dim3(1), dim3(BS), 0, ctx->stream, (const double*)buf, nb, g.hx1, g.hy1,
                     buf + 8 * nb);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(L2, buf + 8 * nb, 8 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(ctx, -2, "calculate_l2error: %s", hipGetErrorString(e));
  return 0;
}

int mhd2d_set_bc(mhd2d_ctx* ctx) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  return launch_setbc(ctx, ctx->S, nullptr);
}

int mhd2d_lh(mhd2d_ctx* ctx) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if ((rc = launch_fluxes(ctx, ctx->S, false, nullptr))) return rc;
  if ((rc = launch_cell(ctx, 0, ctx->S, nullptr, nullptr, ctx->du, nullptr, 0.0))) return rc;
  ctx->lh_done = true;
  return 0;
}

int mhd2d_lhez(mhd2d_ctx* ctx) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if (!ctx->lh_done) return fail(ctx, -5, "lhez: requires a preceding lh on the same state (LhEz.f90 reads Lh's traces)");
  if ((rc = launch_vertex(ctx, ctx->S, nullptr))) return rc;
  return launch_ct(ctx, 0, nullptr, nullptr, nullptr, ctx->dEz, nullptr, 0.0);
}

// RK3.f90:80-91 (stage 1), :106-116 (stage 2), :129-135 (stage 3)
int mhd2d_rk_stage(mhd2d_ctx* ctx, int32_t stage, double dt) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  const long long nf = 6LL * ctx->g.plane;
  const size_t bs = (size_t)state_bytes(ctx->cfg.precision == MHD2D_MIXED, ctx->g.plane), bf = nf * sizeof(double);
  const long long ns = -1;     // state buffers: launch_axpy_state
  hipStream_t st = ctx->stream;
  auto ax = [&](double* out, const double* X, double a, const double* Y, double b, const double* Z, double d, long long n) {
    if (n < 0) (void)launch_axpy_state(ctx, out, X, a, Y, b, Z, d);
    else (void)launch_axpy(ctx, out, X, a, Y, b, Z, d, n);
  };
  if (stage == 1) {
    ax(ctx->UI, nullptr, 0, ctx->S, 1, ctx->du, dt, ns);
    ax(ctx->FI, nullptr, 0, ctx->F, 1, ctx->dEz, dt, nf);
    HIPCHK(ctx, hipMemcpyAsync(ctx->U0, ctx->S, bs, hipMemcpyDeviceToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->F0, ctx->F, bf, hipMemcpyDeviceToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->S, ctx->UI, bs, hipMemcpyDeviceToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->F, ctx->FI, bf, hipMemcpyDeviceToDevice, st));
  } else if (stage == 2) {
    ax(ctx->S, nullptr, 0, ctx->UI, 1, ctx->du, dt, ns);
    ax(ctx->F, nullptr, 0, ctx->FI, 1, ctx->dEz, dt, nf);
    ax(ctx->UII, ctx->U0, 3.0 / 4.0, ctx->S, 1.0 / 4.0, ctx->S, 0.0, ns);
    ax(ctx->FII, ctx->F0, 3.0 / 4.0, ctx->F, 1.0 / 4.0, ctx->F, 0.0, nf);
    HIPCHK(ctx, hipMemcpyAsync(ctx->S, ctx->UII, bs, hipMemcpyDeviceToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->F, ctx->FII, bf, hipMemcpyDeviceToDevice, st));
  } else if (stage == 3) {
    ax(ctx->F, nullptr, 0, ctx->FII, 1, ctx->dEz, dt, nf);
    ax(ctx->S, nullptr, 0, ctx->UII, 1, ctx->du, dt, ns);
    ax(ctx->S, ctx->U0, 1.0 / 3.0, ctx->S, 2.0 / 3.0, ctx->S, 0.0, ns);
    ax(ctx->F, ctx->F0, 1.0 / 3.0, ctx->F, 2.0 / 3.0, ctx->F, 0.0, nf);
  } else {
    return fail(ctx, -1, "rk_stage: stage must be 1, 2 or 3");
  }
  LAUNCHCHK(ctx);
  // the subroutine path combines the faces in plain FP64 (as the reference);
  // the mixed form's low parts are not carried through it
  for (double* fb : {ctx->F, ctx->F0, ctx->FI, ctx->FII})
    if ((rc = clear_face_lo(ctx, fb, true, true))) return rc;
  ctx->lh_done = false;
  return 0;
}

int mhd2d_tvb_limiter_p2(mhd2d_ctx* ctx) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if ((rc = launch_limit(ctx, ctx->S, ctx->T, ctx->F, true, ctx->cfg.pp != 0, false, nullptr))) return rc;
  std::swap(ctx->S, ctx->T);
  ctx->lh_done = false;
  return launch_setbc(ctx, ctx->S, nullptr);   // TVB_Limiter_P2.f90:126
}

int mhd2d_pp_limiter(mhd2d_ctx* ctx) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if ((rc = launch_limit(ctx, ctx->S, ctx->T, ctx->F, false, true, false, nullptr))) return rc;
  std::swap(ctx->S, ctx->T);
  ctx->lh_done = false;
  return 0;
}

int mhd2d_div_free_balsara(mhd2d_ctx* ctx) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if ((rc = launch_limit(ctx, ctx->S, ctx->T, ctx->F, false, false, true, nullptr))) return rc;
  std::swap(ctx->S, ctx->T);
  ctx->lh_done = false;
  return 0;
}

int mhd2d_kxrcf_detector(mhd2d_ctx* ctx, int64_t* nflag) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_scal_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, 1);
  if ((rc = launch_kxrcf(ctx, ctx->S, ctx->sc))) return rc;
  if (nflag) {
    DevScal h;
    HIPCHK(ctx, hipMemcpyAsync(&h, ctx->sc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *nflag = h.nflag;
  }
  return 0;
}

static int launch_dt(mhd2d_ctx* ctx, int advance) {
  const Geo& g = ctx->g;
  Timed tm(ctx, KC_DT);
  hipLaunchKernelGGL(k_scal_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, 0);
  with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_dt<MIX>), dim3((unsigned)std::min<long long>(ctx->nblk_cells, 2048)), dim3(BS), 0,
                       ctx->stream, SRd<MIX>{ctx->S, g.plane}, g, ctx->sc);
    return 0;
  });
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, g.hx, g.hy, advance);
  LAUNCHCHK(ctx);
  return 0;
}

int mhd2d_calculate_dt(mhd2d_ctx* ctx, double* dt) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if ((rc = launch_dt(ctx, 0))) return rc;
  DevScal h;
  HIPCHK(ctx, hipMemcpyAsync(&h, ctx->sc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (dt) *dt = h.dt;
  return h.health;
}

static int launch_diag(mhd2d_ctx* ctx) {
  const Geo& g = ctx->g;
  Timed tm(ctx, KC_DIAG);
  hipLaunchKernelGGL(k_scal_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, 1);
  with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_diag<SRd<MIX>>), dim3(ctx->nblk_cells), dim3(BS), 0, ctx->stream, SRd<MIX>{ctx->S, g.plane}, g,
                       ctx->sc, ctx->partial);
    return 0;
  });
  hipLaunchKernelGGL(k_div_final, dim3(1), dim3(BS), 0, ctx->stream, ctx->sc, (const double*)ctx->partial,
                     ctx->nblk_cells, g.hx1, g.hy1);
  LAUNCHCHK(ctx);
  return 0;
}

int mhd2d_diagnostics(mhd2d_ctx* ctx, double* umax, double* totaldiv, double* pmin) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if ((rc = launch_diag(ctx))) return rc;
  DevScal h;
  HIPCHK(ctx, hipMemcpyAsync(&h, ctx->sc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (umax) *umax = h.umax;
  if (totaldiv) *totaldiv = h.totaldiv;
  if (pmin) *pmin = h.pmin;
  return h.health;
}

int mhd2d_face_jump(mhd2d_ctx* ctx, double* jump) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if (!jump) return fail(ctx, -1, "face_jump: null output");
  const Geo& g = ctx->g;
  const long long n = (long long)(g.Nx + 1) * g.Ny + (long long)g.Nx * (g.Ny + 1);
  unsigned long long* d = reinterpret_cast<unsigned long long*>(ctx->partial);
  HIPCHK(ctx, hipMemsetAsync(d, 0, sizeof *d, ctx->stream));
  with_mix(ctx, [&](auto Mc) -> int {
    constexpr bool MIX = decltype(Mc)::value;
    hipLaunchKernelGGL((k_facejump<MIX>), dim3(nblk(n)), dim3(BS), 0, ctx->stream, SRd<MIX>{ctx->S, g.plane},
                       (const double*)ctx->F, g, d);
    return 0;
  });
  LAUNCHCHK(ctx);
  unsigned long long h = 0;
  HIPCHK(ctx, hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *jump = __builtin_bit_cast(double, h);
  return 0;
}

int mhd2d_calculate_ez(int64_t n, const double* u1, const double* u2, const double* B1, const double* B2,
                       double* Ez) {
  if (n < 0 || (n > 0 && (!u1 || !u2 || !B1 || !B2 || !Ez))) return -1;
  for (int64_t k = 0; k < n; k++) Ez[k] = calculate_ez(u1[k], u2[k], B1[k], B2[k]);
  return 0;
}

int mhd2d_set_time(mhd2d_ctx* ctx, double t, int64_t count) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  DevScal h;
  HIPCHK(ctx, hipMemcpyAsync(&h, ctx->sc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  h.t = t;
  h.count = count;
  h.active = 1;
  h.health = 0;
  HIPCHK(ctx, hipMemcpyAsync(ctx->sc, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int mhd2d_get_time(mhd2d_ctx* ctx, double* t, double* dt, int64_t* count) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  DevScal h;
  HIPCHK(ctx, hipMemcpyAsync(&h, ctx->sc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (t) *t = h.t;
  if (dt) *dt = h.dt;
  if (count) *count = h.count;
  return h.health;
}

// RK3.f90:5-18
int mhd2d_rk3_prologue(mhd2d_ctx* ctx) {
  int rc;
  if ((rc = mhd2d_set_time(ctx, 0.0, 0))) return rc;
  if ((rc = mhd2d_set_bc(ctx))) return rc;
  if ((rc = mhd2d_tvb_limiter_p2(ctx))) return rc;
  if ((rc = mhd2d_div_free_balsara(ctx))) return rc;
  return launch_diag(ctx);
}

}  // extern "C"

namespace {

// the kernels of one pass of the RK3.f90:58-174 loop body, from dt to the
// new state (tend is set by the caller); rotates the buffer roles
int step_body(mhd2d_ctx* ctx) {
  struct GateOn {
    mhd2d_ctx* c;
    explicit GateOn(mhd2d_ctx* c_) : c(c_) { c->g.gate = &c->sc->active; }
    ~GateOn() { c->g.gate = nullptr; }
  } gate_on(ctx);
  int rc;
  if ((rc = launch_dt(ctx, 1))) return rc;
  DevScal* sc = ctx->sc;
  const bool pp = ctx->cfg.pp != 0;
  // one stage: fluxes of S; the CT update of the faces (Fa, Fb -> Fo); the
  // cell update (A, U0 -> out); the limiter pass (out -> lim, faces Fo).
  // Except with TVB on every cell, the limiter pass without TVB runs in the
  // cell kernel's epilogue (the CT goes first so the new faces are there),
  // and only the cells TVB changes are limited again; in the mixed form the
  // epilogue leaves the cells its positivity bound cannot settle to a list
  // kernel, and the limited B planes to a pass after the CT update, which
  // needs no CT before it: the CT update and that pass run after the cell
  // kernel, beside the detection.
  const bool mixed = ctx->cfg.precision == MHD2D_MIXED;
  const bool fused = ctx->cfg.indicator != MHD2D_IND_ALL;
  // The vertex EMFs (FP64 VALU-bound) run on a second stream beside the face
  // kernels and the CT update after them on that stream (beside the cell
  // kernel unfused, beside the detection in the mixed form), forked from and
  // joined back into the context stream by events inside each stage.
  // With the per-kernel event timing on, everything stays on one stream so
  // that each kernel is timed alone.  Graph capture also stays on one stream:
  // with this ROCm, a captured graph with parallel branches keeps the buffers
  // its kernels used from being released when the context frees them (about
  // 3 GB per context at 2048^2, measured), which a single-stream graph does not.
  const bool two = !ctx->prof && !ctx->capturing;
  auto on2 = [&](auto&& f) -> int {
    if (!two) return f();
    std::swap(ctx->stream, ctx->stream2);
    const int r = f();
    std::swap(ctx->stream, ctx->stream2);
    return r;
  };
  auto wait = [&](hipEvent_t ev, hipStream_t from, hipStream_t to) -> int {
    if (!two) return 0;
    HIPCHK(ctx, hipEventRecord(ev, from));
    HIPCHK(ctx, hipStreamWaitEvent(to, ev, 0));
    return 0;
  };
  auto stage = [&](int k, const double* Fa, const double* Fb, double* Fo, const double* A, double* out,
                   double* lim) -> int {
    int r;
    auto ct = [&] { return launch_ct(ctx, k, k == 1 ? Fa : nullptr, k == 1 ? nullptr : Fa, Fb, Fo, sc, 0.0); };
    auto vertex = [&] { return launch_vertex(ctx, ctx->S, sc); };
    const hipStream_t s1 = ctx->stream, s2 = ctx->stream2;
    if (fused && !mixed) {
      // vertex EMFs beside the face kernels; the cell kernel's limiter
      // epilogue reads the new faces, so it follows the CT.  As in the mixed
      // form below, the stage value's planes other than Bx, By live in lim
      // (its Bx, By planes in out) until the list kernels limit cells in
      // place, so the detection reads lim and out and runs first.
      if ((r = wait(ctx->ev_fork, s1, s2))) return r;
      if ((r = on2(vertex))) return r;
      if ((r = launch_fluxes(ctx, ctx->S, false, sc))) return r;
      if ((r = wait(ctx->ev_faces, s1, s2))) return r;
      if ((r = on2(ct))) return r;
      if ((r = wait(ctx->ev_join, s2, s1))) return r;
      if ((r = launch_cell(ctx, k, ctx->S, A, ctx->U0, out, sc, 0.0, lim, Fo, pp))) return r;
      if ((r = launch_detection(ctx, lim, sc, k, out))) return r;
      if (pp && (r = launch_pp_list(ctx, out, lim, Fo, k))) return r;
      return launch_limit_relist(ctx, out, lim, Fo, pp, k, true);
    }
    if (fused) {
      // mixed form: vertex EMFs beside the face kernels; the cell kernel's
      // lean epilogue writes the planes other than Bx, By and needs no
      // faces, so the CT update and the divergence-free B planes of the
      // limited value follow it, beside the detection (memory-bound work
      // beside the memory-bound cell kernel only stretched it: 8.29-8.30
      // against 8.11 ms per step at 2048^2, measured).  The stage value's
      // planes other than Bx, By live in lim (its Bx, By planes in out)
      // until the list kernels limit cells in place, so the detection reads
      // lim and out and runs first.
      if ((r = wait(ctx->ev_fork, s1, s2))) return r;
      if ((r = on2(vertex))) return r;
      if ((r = launch_fluxes(ctx, ctx->S, false, sc))) return r;
      if ((r = launch_cell(ctx, k, ctx->S, A, ctx->U0, out, sc, 0.0, lim, Fo, pp))) return r;
      // the CT update and the limited B planes (memory-bound, as the cell
      // kernel is) on the second stream after the cell kernel, beside the
      // detection chain; then the positivity list and the cells TVB changed,
      // limited again (after the list, so that their TVB-limited value is
      // the one that stays)
      if ((r = wait(ctx->ev_cell, s1, s2))) return r;
      if ((r = on2(ct)) || (r = on2([&] { return launch_divfree(ctx, lim, Fo); }))) return r;
      if (two) HIPCHK(ctx, hipEventRecord(ctx->ev_join, s2));       // new faces and limited B done
      if ((r = launch_detection(ctx, lim, sc, k, out))) return r;
      if (two) HIPCHK(ctx, hipStreamWaitEvent(s1, ctx->ev_join, 0));
      if (pp && (r = launch_pp_list(ctx, out, lim, Fo, k))) return r;
      return launch_limit_relist(ctx, out, lim, Fo, pp, k, true);
    }
    // vertex EMFs and CT beside the memory-bound cell kernel
    if ((r = launch_fluxes(ctx, ctx->S, false, sc))) return r;
    if ((r = wait(ctx->ev_faces, s1, s2))) return r;
    if ((r = on2(vertex)) || (r = on2(ct))) return r;
    if (two) HIPCHK(ctx, hipEventRecord(ctx->ev_join, s2));       // CT done
    if ((r = launch_cell(ctx, k, ctx->S, A, ctx->U0, out, sc, 0.0))) return r;
    const int ind = ctx->cfg.indicator;
    if (ind != MHD2D_IND_MLP && ind != MHD2D_IND_KXRCF) {
      if (two) HIPCHK(ctx, hipStreamWaitEvent(s1, ctx->ev_join, 0));
      return launch_limit(ctx, out, lim, Fo, true, pp, true, sc, k);
    }
    // detection and TVB on the second stream beside the limiter pass without
    // TVB; then the cells TVB changed are limited again
    if ((r = wait(ctx->ev_cell, s1, s2))) return r;
    if ((r = on2([&] { return launch_detection(ctx, out, sc, k); }))) return r;
    if (two) HIPCHK(ctx, hipEventRecord(ctx->ev_det, s2));
    if (two) HIPCHK(ctx, hipStreamWaitEvent(s1, ctx->ev_join, 0));
    if ((r = launch_limit(ctx, out, lim, Fo, false, pp, true, sc, k))) return r;
    if (two) HIPCHK(ctx, hipStreamWaitEvent(s1, ctx->ev_det, 0));
    return launch_limit_relist(ctx, out, lim, Fo, pp, k);
  };
  // stage 1 (RK3.f90:74-97): set_bc; UI = S + dt L(S); S1 = limit(UI);
  // U0 keeps S with its fresh halos
  if ((rc = launch_setbc(ctx, ctx->S, sc))) return rc;
  if ((rc = stage(1, ctx->F, nullptr, ctx->FI, nullptr, ctx->UI, ctx->T))) return rc;
  if ((rc = launch_setbc(ctx, ctx->T, sc))) return rc;          // RK3.f90:97
  std::swap(ctx->U0, ctx->S);   // U0 <- start-of-step state
  std::swap(ctx->S, ctx->T);    // S  <- limited stage-1 state
  std::swap(ctx->F0, ctx->F);   // F0 <- start-of-step faces; F is free
  // stage 2: UII = 3/4 U0 + 1/4 (UI + dt L(S)); S2 = limit(UII)
  if ((rc = stage(2, ctx->FI, ctx->F0, ctx->FII, ctx->UI, ctx->UII, ctx->T))) return rc;
  if ((rc = launch_setbc(ctx, ctx->T, sc))) return rc;          // RK3.f90:123 (stage 3 set_bc)
  std::swap(ctx->S, ctx->T);
  // stage 3: T = 1/3 U0 + 2/3 (UII + dt L(S)); S_new = limit(T)
  if ((rc = stage(3, ctx->FII, ctx->F0, ctx->F, ctx->UII, ctx->T, ctx->UI))) return rc;
  // a switched-off step: the start-of-step state and faces become the new ones
  const long long ns = (long long)(state_bytes(ctx->cfg.precision == MHD2D_MIXED, ctx->g.plane) / sizeof(double));
  const long long nf = (long long)face_planes(ctx) * ctx->g.plane;
  const unsigned nbc = std::min<unsigned>(nblk(ns), 1024u);
  hipLaunchKernelGGL(k_carry, dim3(nbc), dim3(BS), 0, ctx->stream, (const double*)ctx->U0, ctx->UI, ns, ctx->g.gate);
  hipLaunchKernelGGL(k_carry, dim3(nbc), dim3(BS), 0, ctx->stream, (const double*)ctx->F0, ctx->F, nf, ctx->g.gate);
  LAUNCHCHK(ctx);
  std::swap(ctx->S, ctx->UI);   // S <- new state (halos refreshed by the next step's set_bc); UI is free
  ctx->lh_done = false;
  return 0;
}

// the buffers whose roles rotate from step to step (period 6)
constexpr int NROT = 9;
void rot_get(const mhd2d_ctx* c, double* p[NROT]) {
  double* const v[NROT] = {c->S, c->U0, c->UI, c->UII, c->T, c->F, c->F0, c->FI, c->FII};
  for (int k = 0; k < NROT; k++) p[k] = v[k];
}
void rot_set(mhd2d_ctx* c, double* const p[NROT]) {
  c->S = p[0]; c->U0 = p[1]; c->UI = p[2]; c->UII = p[3]; c->T = p[4];
  c->F = p[5]; c->F0 = p[6]; c->FI = p[7]; c->FII = p[8];
}

// captures step_body for the current buffer roles; the roles are left as
// they were (nothing has run) and the ones after the step go to sg.after
int capture_step(mhd2d_ctx* ctx, StepGraph& sg) {
  rot_get(ctx, sg.before);
  hipGraph_t graph = nullptr;
  HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  const int rc = step_body(ctx);
  ctx->capturing = false;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  rot_get(ctx, sg.after);
  rot_set(ctx, sg.before);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(ctx, -2, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&sg.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(ctx, -2, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
  return 0;
}

int read_t(mhd2d_ctx* ctx, double* t_out) {
  if (!t_out) return 0;
  DevScal h;
  HIPCHK(ctx, hipMemcpyAsync(&h, ctx->sc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *t_out = h.t;
  return h.health;
}

}  // namespace

extern "C" {

// one pass of the RK3.f90:58-174 loop body with t and dt on the device
int mhd2d_step(mhd2d_ctx* ctx, double tend, double* t_out) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_tend, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, tend);
  if ((rc = step_body(ctx))) return rc;
  return read_t(ctx, t_out);
}

// nsteps passes of the loop body, each one launch of a HIP graph captured
// (once per buffer-role phase) from the same kernels as mhd2d_step
int mhd2d_step_graph(mhd2d_ctx* ctx, double tend, int64_t nsteps, double* t_out) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  if (nsteps < 0) return fail(ctx, -1, "mhd2d_step_graph: nsteps = %lld < 0", (long long)nsteps);
  // the graph of the current buffer roles (captured when missing)
  auto graph_now = [&](StepGraph*& sg) -> int {
    double* cur[NROT];
    rot_get(ctx, cur);
    sg = nullptr;
    for (auto& e : ctx->graphs)
      if (std::equal(cur, cur + NROT, e.before)) { sg = &e; return 0; }
    StepGraph n{};
    if (int r = capture_step(ctx, n)) return r;
    if (ctx->graphs.size() >= 8) {   // only six phases occur while stepping
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      (void)hipGraphExecDestroy(ctx->graphs.front().exec);
      ctx->graphs.erase(ctx->graphs.begin());
    }
    ctx->graphs.push_back(n);
    sg = &ctx->graphs.back();
    return 0;
  };
  StepGraph* sg = nullptr;
  if (nsteps == 0) {   // capture the six phases ahead, run nothing
    if (ctx->prof) return 0;
    double* start[NROT];
    rot_get(ctx, start);
    for (int k = 0; k < 6; k++) {
      if ((rc = graph_now(sg))) { rot_set(ctx, start); return rc; }
      rot_set(ctx, sg->after);
    }
    rot_set(ctx, start);
    return 0;
  }
  hipLaunchKernelGGL(k_set_tend, dim3(1), dim3(64), 0, ctx->stream, ctx->sc, tend);
  LAUNCHCHK(ctx);
  for (int64_t s = 0; s < nsteps; s++) {
    if (ctx->prof) {   // event-timed launches are not captured
      if ((rc = step_body(ctx))) return rc;
      continue;
    }
    if ((rc = graph_now(sg))) return rc;
    HIPCHK(ctx, hipGraphLaunch(sg->exec, ctx->stream));
    rot_set(ctx, sg->after);
    ctx->lh_done = false;
  }
  return read_t(ctx, t_out);
}

// live kernel timing with HIP events on the context stream (no counterpart
// in the reference): enable, run, then read per-class totals.
int mhd2d_profile_enable(mhd2d_ctx* ctx, int32_t on) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& pe : ctx->prof_ev) { ctx->ev_pool.push_back(pe.second.first); ctx->ev_pool.push_back(pe.second.second); }
  ctx->prof_ev.clear();
  ctx->prof = on != 0;
  return 0;
}

int mhd2d_profile_get(mhd2d_ctx* ctx, double* ms, int64_t* launches, int32_t nclass) {
  int rc = ctx_check(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < nclass; k++) { ms[k] = 0.0; launches[k] = 0; }
  for (auto& pe : ctx->prof_ev) {
    float e = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&e, pe.second.first, pe.second.second));
    if (pe.first < nclass) { ms[pe.first] += e; launches[pe.first] += 1; }
  }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// calculate_umax / _totaldiv / _pmin without a context and without any
// allocation.  Its two reductions need scratch memory (a DevScal of atomics
// and one partial divergence sum per 256-cell block); it lives in the
// file-scope device array below instead of a context's buffers.  The host
// only ever passes the array's device address, looked up once per device.
namespace {

constexpr long long NOALLOC_MAX_CELLS = 4096LL * 4096;           // Nx * Ny
constexpr long long NOALLOC_MAX_BLOCKS = NOALLOC_MAX_CELLS / BS;  // partial sums
constexpr long long NOALLOC_SC_DOUBLES = 64;                      // the DevScal
static_assert(sizeof(DevScal) <= NOALLOC_SC_DOUBLES * sizeof(double), "DevScal fits its scratch slot");
// 512 KiB of partial sums + 512 B of scalars, held per device until exit
static __device__ __align__(256) double scratch[NOALLOC_MAX_BLOCKS + NOALLOC_SC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::atomic<double*> noalloc_ptr[NOALLOC_MAX_DEVICES];
std::mutex noalloc_mu;

// the device address of `scratch` on the current device (hipGetSymbolAddress
// once per device, with that device current); nullptr on failure
double* noalloc_scratch(int dev) {
  double* p = noalloc_ptr[dev].load(std::memory_order_acquire);
  if (p) return p;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  p = noalloc_ptr[dev].load(std::memory_order_relaxed);
  if (p) return p;
  void* a = nullptr;
  if (hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)) != hipSuccess || !a) return nullptr;
  p = static_cast<double*>(a);
  noalloc_ptr[dev].store(p, std::memory_order_release);
  return p;
}

// uh(0:Nx+1, 0:Ny+1, 10, 8) in device memory, column-major (com.txt:22):
// plane p of PL() is the reference slot (mode b, variable n), b + 10 n, and
// cell index c of cix() is OFF + i + (Nx + 2) j when the row pitch is Nx + 2
struct RefView {
  const double* b;
  long long pl;   // (Nx + 2) (Ny + 2)
  __device__ __forceinline__ double operator()(int p, long long c) const {
    const int slot = p < 48 ? (p % 6) + 10 * (p / 6) : p == 48 ? 6 + 50 : p == 49 ? 8 + 50 : p == 50 ? 7 + 60 : 9 + 60;
    return b[(long long)slot * pl + (c - OFF)];
  }
};

}  // namespace

extern "C" int mhd2d_diagnostics_noalloc(int32_t Nx, int32_t Ny, double xa, double xb, double ya, double yb,
                                         const double* uh, double* out, void* stream) {
  if (!uh || !out || Nx < 2 || Ny < 2 || !(xb > xa) || !(yb > ya)) return -1;
  if ((long long)Nx * Ny > NOALLOC_MAX_CELLS) {
    fprintf(stderr, "mhd2d_diagnostics_noalloc: %d x %d cells exceed its scratch maximum of %lld cells\n", Nx, Ny,
            NOALLOC_MAX_CELLS);
    return -3;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return -2;
  double* base = noalloc_scratch(dev);
  if (!base) return -2;
  double* partial = base;
  DevScal* sc = reinterpret_cast<DevScal*>(base + NOALLOC_MAX_BLOCKS);
  hipStream_t st = static_cast<hipStream_t>(stream);
  Geo g{};
  g.Nx = Nx;
  g.Ny = Ny;
  g.P = Nx + 2;   // cix(g, i, j) - OFF = i + (Nx + 2) j: the reference's (i, j) offset
  g.plane = (long long)(Nx + 2) * (Ny + 2);
  g.hx = (xb - xa) / Nx;
  g.hy = (yb - ya) / Ny;
  g.hx1 = 0.5 * g.hx;
  g.hy1 = 0.5 * g.hy;
  const unsigned nb = nblk((long long)Nx * Ny);
  // the atomics' start values come from a kernel (no memset node under capture)
  hipLaunchKernelGGL(k_scal_reset, dim3(1), dim3(64), 0, st, sc, 1);
  hipLaunchKernelGGL((k_diag<RefView>), dim3(nb), dim3(BS), 0, st, RefView{uh, g.plane}, g, sc, partial);
  hipLaunchKernelGGL(k_div_final, dim3(1), dim3(BS), 0, st, sc, (const double*)partial, (long long)nb, g.hx1,
                     g.hy1, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
