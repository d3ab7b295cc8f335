// mhd2d_physics.hpp -- pointwise ideal-MHD physics of GDF_Balsara_P2_0316 as
// CDNA4 device functions, templated on the arithmetic type (double for the
// plain form; float for the flux/quadrature part of the mixed form).
//
// Every function follows the reference routine it cites; divisions by a
// common denominator are done once as a reciprocal (a rounding-level
// difference from the reference, far below the 1e-12 parity bar).
#pragma once
#include <hip/hip_runtime.h>

namespace mhd {

// com.txt:7-8: gamma = 5/3, gamma1 = gamma - 1 (both rounded in FP64)
template <typename T> __device__ __forceinline__ T GAM() { return T(5.0 / 3.0); }
template <typename T> __device__ __forceinline__ T GAM1() { return T(5.0 / 3.0 - 1.0); }

template <typename T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
// Fortran SIGN(1, x)
template <typename T> __device__ __forceinline__ T sign1(T x) { return copysign(T(1), x); }

// eigenvalueMm.f90:1-24 with n = (1,0,0) (DIR 1) or (0,1,0) (DIR 2)
template <typename T, int DIR>
__device__ __forceinline__ void eigenvalueMm(const T U[8], T& amax, T& amin) {
  const T rho = U[0], ir = T(1) / rho;
  const T u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
  const T B1 = U[5], B2 = U[6], B3 = U[7];
  const T BP = B1 * B1 + B2 * B2 + B3 * B3;
  const T Bn = DIR == 1 ? B1 : B2;
  const T un = DIR == 1 ? u : v;
  const T p = GAM1<T>() * (U[4] - T(0.5) * rho * (u * u + v * v + w * w) - T(0.5) * BP);
  const T c = sqrt(fabs(GAM<T>() * p * ir));
  const T c2 = c * c;
  const T q = c2 + BP * ir;
  const T cf = sqrt(fabs(T(0.5) * (q + sqrt(q * q - T(4) * c2 * (Bn * Bn) * ir))));
  amax = un + cf;
  amin = un - cf;
}

// physical flux of Lh.f90:22-53 / :89-202 in direction DIR.  YDOWN
// reproduces the pressure of Lh.f90:189 (-wMD**2) used for the lower-side
// traces of the y faces.
template <typename T, int DIR, bool YDOWN = false>
__device__ __forceinline__ void lh_flux(const T U[8], T F[8]) {
  const T rho = U[0], ir = T(1) / rho;
  const T u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
  const T E = U[4], B1 = U[5], B2 = U[6], B3 = U[7];
  const T Bs = B1 * B1 + B2 * B2 + B3 * B3;
  const T p = YDOWN ? GAM1<T>() * (E - T(0.5) * rho * (u * u + v * v - w * w) - T(0.5) * Bs)
                    : GAM1<T>() * (E - T(0.5) * rho * (u * u + v * v + w * w) - T(0.5) * Bs);
  const T S = p + T(0.5) * Bs;
  const T TT = E + S;
  const T K = u * B1 + v * B2 + w * B3;
  if (DIR == 1) {
    F[0] = U[1];
    F[1] = rho * (u * u) + S - B1 * B1;
    F[2] = rho * u * v - B1 * B2;
    F[3] = rho * u * w - B1 * B3;
    F[4] = TT * u - K * B1;
    F[5] = T(0);
    F[6] = u * B2 - v * B1;
    F[7] = u * B3 - w * B1;
  } else {
    F[0] = U[2];
    F[1] = rho * u * v - B1 * B2;
    F[2] = rho * (v * v) + S - B2 * B2;
    F[3] = rho * v * w - B2 * B3;
    F[4] = TT * v - K * B2;
    F[5] = v * B1 - u * B2;
    F[6] = T(0);
    F[7] = v * B3 - w * B2;
  }
}

// both volume fluxes at once (Lh.f90:22-53)
template <typename T>
__device__ __forceinline__ void lh_flux_xy(const T U[8], T Fx[8], T Fy[8]) {
  const T rho = U[0], ir = T(1) / rho;
  const T u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
  const T E = U[4], B1 = U[5], B2 = U[6], B3 = U[7];
  const T Bs = B1 * B1 + B2 * B2 + B3 * B3;
  const T p = GAM1<T>() * (E - T(0.5) * rho * (u * u + v * v + w * w) - T(0.5) * Bs);
  const T S = p + T(0.5) * Bs;
  const T TT = E + S;
  const T K = u * B1 + v * B2 + w * B3;
  const T ruv = rho * u * v - B1 * B2;
  Fx[0] = U[1];
  Fx[1] = rho * (u * u) + S - B1 * B1;
  Fx[2] = ruv;
  Fx[3] = rho * u * w - B1 * B3;
  Fx[4] = TT * u - K * B1;
  Fx[5] = T(0);
  Fx[6] = u * B2 - v * B1;
  Fx[7] = u * B3 - w * B1;
  Fy[0] = U[2];
  Fy[1] = ruv;
  Fy[2] = rho * (v * v) + S - B2 * B2;
  Fy[3] = rho * v * w - B2 * B3;
  Fy[4] = TT * v - K * B2;
  Fy[5] = v * B1 - u * B2;
  Fy[6] = T(0);
  Fy[7] = v * B3 - w * B2;
}

// MHD_flux.f90:7-40 (corner solvers)
template <typename T>
__device__ __forceinline__ void mhd_flux(const T U[8], T Fx[8], T Fy[8]) {
  const T rho = U[0], ir = T(1) / rho;
  const T u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
  const T E = U[4], B1 = U[5], B2 = U[6], B3 = U[7];
  const T Bs = B1 * B1 + B2 * B2 + B3 * B3;
  const T p = GAM1<T>() * (E - T(0.5) * rho * (u * u + v * v + w * w) - T(0.5) * Bs);
  const T S = p + T(0.5) * Bs;
  const T TT = E + S;
  const T K = u * B1 + v * B2 + w * B3;
  Fx[0] = U[1];
  Fx[1] = rho * (u * u) + S - B1 * B1;
  Fx[2] = rho * u * v - B1 * B2;
  Fx[3] = rho * u * w - B1 * B3;
  Fx[4] = TT * u - K * B1;
  Fx[5] = T(0);
  Fx[6] = u * B2 - v * B1;
  Fx[7] = u * B3 - w * B1;
  Fy[0] = U[2];
  Fy[1] = rho * u * v - B1 * B2;
  Fy[2] = rho * (v * v) + S - B2 * B2;
  Fy[3] = rho * w * v - B3 * B2;
  Fy[4] = TT * v - K * B2;
  Fy[5] = v * B1 - u * B2;
  Fy[6] = T(0);
  Fy[7] = v * B3 - w * B2;
}

// LF_Flux.f90:5
template <typename T>
__device__ __forceinline__ void lf_flux(T SR, T SL, const T UL[8], const T UR[8], const T FL[8],
                                        const T FR[8], T F[8]) {
  const T a = tmax(fabs(SR), fabs(SL));
#pragma unroll
  for (int n = 0; n < 8; n++) F[n] = T(0.5) * (FR[n] + FL[n] - a * (UR[n] - UL[n]));
}

// HLL_Flux.f90:5-11
template <typename T>
__device__ __forceinline__ void hll_flux(T SR, T SL, const T UL[8], const T UR[8], const T FL[8],
                                         const T FR[8], T F[8]) {
  if (SR < T(0)) {
#pragma unroll
    for (int n = 0; n < 8; n++) F[n] = FR[n];
  } else if (SL > T(0)) {
#pragma unroll
    for (int n = 0; n < 8; n++) F[n] = FL[n];
  } else {
    const T id = T(1) / (SR - SL);
#pragma unroll
    for (int n = 0; n < 8; n++) F[n] = (SR * FL[n] - SL * FR[n] + SL * SR * (UR[n] - UL[n])) * id;
  }
}

// HLLC_Flux.f90:9-106 and HLLD_Flux.f90:14-189 share their first stage;
// HLLD adds the Alfven (double-star) states.  DIR is the reference's
// 'direction'.  Ustar (optional) receives the star state as the reference
// leaves it in COMMON /flux6/.
template <typename T, int DIR, bool HLLD>
__device__ __forceinline__ void hllcd_flux(T SR, T SL, const T UL1[8], const T UR1[8],
                                           const T FL1[8], const T FR1[8], T F[8], T* Ustar) {
  if (SR < T(0)) {
#pragma unroll
    for (int n = 0; n < 8; n++) { F[n] = FR1[n]; if (Ustar) Ustar[n] = UR1[n]; }
    return;
  }
  if (SL > T(0)) {
#pragma unroll
    for (int n = 0; n < 8; n++) { F[n] = FL1[n]; if (Ustar) Ustar[n] = UL1[n]; }
    return;
  }
  T Us[8], URs[8], ULs[8];
  const T idS = T(1) / (SR - SL);
#pragma unroll
  for (int n = 0; n < 8; n++) Us[n] = (SR * UR1[n] - SL * UL1[n] + FL1[n] - FR1[n]) * idS;
  const T rhoR = UR1[0], irR = T(1) / rhoR;
  const T u1R = UR1[1] * irR, u2R = UR1[2] * irR, u3R = UR1[3] * irR, ER = UR1[4];
  const T B1R = UR1[5], B2R = UR1[6], B3R = UR1[7];
  const T rhoL = UL1[0], irL = T(1) / rhoL;
  const T u1L = UL1[1] * irL, u2L = UL1[2] * irL, u3L = UL1[3] * irL, EL = UL1[4];
  const T B1L = UL1[5], B2L = UL1[6], B3L = UL1[7];
  const T Ss = (DIR == 1 ? Us[1] : Us[2]) / Us[0];
  const T uRb = DIR == 1 ? u1R : u2R, uLb = DIR == 1 ? u1L : u2L;
  const T BRb = DIR == 1 ? B1R : B2R, BLb = DIR == 1 ? B1L : B2L;
  const T Bsb = DIR == 1 ? Us[5] : Us[6];
  const T iRS = T(1) / (SR - Ss), iLS = T(1) / (SL - Ss);
  const T rRu = rhoR * (SR - uRb), rLu = rhoL * (SL - uLb);
  const T rhoRs = rRu * iRS;
  const T rhoLs = rLu * iLS;
  const T PtotR = GAM1<T>() * (ER - T(0.5) * rhoR * (u1R * u1R + u2R * u2R + u3R * u3R));
  const T PtotL = GAM1<T>() * (EL - T(0.5) * rhoL * (u1L * u1L + u2L * u2L + u3L * u3L));
  const T PtotRs = PtotR + rRu * (Ss - uRb) + Bsb * Bsb - BRb * BRb;
  const T PtotLs = PtotL + rLu * (Ss - uLb) + Bsb * Bsb - BLb * BLb;
  T B1Rs, B1Ls, B2Rs, B2Ls, B3Rs, B3Ls;
  if (HLLD) {
    const T AR = rhoR * ((SR - uRb) * (SR - uRb)) - Bsb * BRb;
    const T AL = rhoL * ((SL - uLb) * (SL - uLb)) - Bsb * BLb;
    const T CR = rRu * (BRb - Bsb);
    const T CL = rLu * (BLb - Bsb);
    const T iDR = T(1) / (rRu * (SR - Ss) - Bsb * Bsb + T(1e-15));
    const T iDL = T(1) / (rLu * (SL - Ss) - Bsb * Bsb + T(1e-15));
    if (DIR == 1) {
      B1Rs = Bsb; B1Ls = Bsb;
      B2Rs = (AR * B2R + CR * u2R) * iDR;
      B2Ls = (AL * B2L + CL * u2L) * iDL;
    } else {
      B1Rs = (AR * B1R + CR * u1R) * iDR;
      B1Ls = (AL * B1L + CL * u1L) * iDL;
      B2Rs = Bsb; B2Ls = Bsb;
    }
    B3Rs = (AR * B3R + CR * u3R) * iDR;
    B3Ls = (AL * B3L + CL * u3L) * iDL;
  } else {
    B1Rs = B1Ls = Us[5];
    B2Rs = B2Ls = Us[6];
    B3Rs = B3Ls = Us[7];
  }
  if (DIR == 1) {
    URs[1] = (UR1[1] * (SR - uRb) + PtotRs - PtotR + BRb * B1R - Bsb * B1Rs) * iRS;
    URs[2] = (UR1[2] * (SR - uRb) + BRb * B2R - Bsb * B2Rs) * iRS;
    ULs[1] = (UL1[1] * (SL - uLb) + PtotLs - PtotL + BLb * B1L - Bsb * B1Ls) * iLS;
    ULs[2] = (UL1[2] * (SL - uLb) + BLb * B2L - Bsb * B2Ls) * iLS;
  } else {
    URs[1] = (UR1[1] * (SR - uRb) + BRb * B1R - Bsb * B1Rs) * iRS;
    URs[2] = (UR1[2] * (SR - uRb) + PtotRs - PtotR + BRb * B2R - Bsb * B2Rs) * iRS;
    ULs[1] = (UL1[1] * (SL - uLb) + BLb * B1L - Bsb * B1Ls) * iLS;
    ULs[2] = (UL1[2] * (SL - uLb) + PtotLs - PtotL + BLb * B2L - Bsb * B2Ls) * iLS;
  }
  URs[3] = (UR1[3] * (SR - uRb) + BRb * B3R - Bsb * B3Rs) * iRS;
  ULs[3] = (UL1[3] * (SL - uLb) + BLb * B3L - Bsb * B3Ls) * iLS;
  const T irRs = T(1) / rhoRs, irLs = T(1) / rhoLs;
  const T u1Rs = URs[1] * irRs, u2Rs = URs[2] * irRs, u3Rs = URs[3] * irRs;
  const T u1Ls = ULs[1] * irLs, u2Ls = ULs[2] * irLs, u3Ls = ULs[3] * irLs;
  URs[4] = ((SR - uRb) * ER - PtotR * uRb + PtotRs * Ss + BRb * (u1R * B1R + u2R * B2R + u3R * B3R) -
            Bsb * (u1Rs * B1Rs + u2Rs * B2Rs + u3Rs * B3Rs)) * iRS;
  ULs[4] = ((SL - uLb) * EL - PtotL * uLb + PtotLs * Ss + BLb * (u1L * B1L + u2L * B2L + u3L * B3L) -
            Bsb * (u1Ls * B1Ls + u2Ls * B2Ls + u3Ls * B3Ls)) * iLS;
  URs[0] = rhoRs; URs[5] = Us[5]; URs[6] = Us[6]; URs[7] = Us[7];
  ULs[0] = rhoLs; ULs[5] = Us[5]; ULs[6] = Us[6]; ULs[7] = Us[7];
  bool dstar = false;
  T SRs = T(0), SLs = T(0);
  if (HLLD) {
    SRs = Ss + sqrt(Bsb * Bsb * irRs);
    SLs = Ss - sqrt(Bsb * Bsb * irLs);
    dstar = fabs(Bsb) > T(1e-4) && SRs - SLs > T(1e-4);
  }
  if (dstar) {
    if (SLs > T(0)) {
#pragma unroll
      for (int n = 0; n < 8; n++) { F[n] = FL1[n] + SL * (ULs[n] - UL1[n]); if (Ustar) Ustar[n] = ULs[n]; }
    } else if (SRs < T(0)) {
#pragma unroll
      for (int n = 0; n < 8; n++) { F[n] = FR1[n] + SR * (URs[n] - UR1[n]); if (Ustar) Ustar[n] = URs[n]; }
    } else {
      const T sg = sign1(Bsb);
      const T sL = sqrt(fabs(rhoLs)), sR = sqrt(fabs(rhoRs)), sLR = sqrt(fabs(rhoLs * rhoRs));
      const T isum = T(1) / (sL + sR);
      const T u1ss = (sL * u1Ls + sR * u1Rs + sg * (B1Rs - B1Ls)) * isum;
      const T u2ss = (sL * u2Ls + sR * u2Rs + sg * (B2Rs - B2Ls)) * isum;
      const T u3ss = (sL * u3Ls + sR * u3Rs + sg * (B3Rs - B3Ls)) * isum;
      const T B1ss = (sL * B1Ls + sR * B1Rs + sg * sLR * (u1Rs - u1Ls)) * isum;
      const T B2ss = (sL * B2Ls + sR * B2Rs + sg * sLR * (u2Rs - u2Ls)) * isum;
      const T B3ss = (sL * B3Ls + sR * B3Rs + sg * sLR * (u3Rs - u3Ls)) * isum;
      const T dss = u1ss * B1ss + u2ss * B2ss + u3ss * B3ss;
      if (Ss > T(0)) {
        const T ELss = ULs[4] - sL * (u1Ls * B1Ls + u2Ls * B2Ls + u3Ls * B3Ls - dss) * sg;
        const T ULss[8] = {rhoLs, rhoLs * u1ss, rhoLs * u2ss, rhoLs * u3ss, ELss, B1ss, B2ss, B3ss};
#pragma unroll
        for (int n = 0; n < 8; n++) {
          const T FLm = FL1[n] + SL * (ULs[n] - UL1[n]);
          F[n] = FLm + SLs * (ULss[n] - ULs[n]);
          if (Ustar) Ustar[n] = ULss[n];
        }
      } else {
        const T ERss = URs[4] + sR * (u1Rs * B1Rs + u2Rs * B2Rs + u3Rs * B3Rs - dss) * sg;
        const T URss[8] = {rhoRs, rhoRs * u1ss, rhoRs * u2ss, rhoRs * u3ss, ERss, B1ss, B2ss, B3ss};
#pragma unroll
        for (int n = 0; n < 8; n++) {
          const T FRm = FR1[n] + SR * (URs[n] - UR1[n]);
          F[n] = FRm + SRs * (URss[n] - URs[n]);
          if (Ustar) Ustar[n] = URss[n];
        }
      }
    }
  } else {
    if (Ss > T(0)) {
#pragma unroll
      for (int n = 0; n < 8; n++) { F[n] = FL1[n] + SL * (ULs[n] - UL1[n]); if (Ustar) Ustar[n] = ULs[n]; }
    } else {
#pragma unroll
      for (int n = 0; n < 8; n++) { F[n] = FR1[n] + SR * (URs[n] - UR1[n]); if (Ustar) Ustar[n] = URs[n]; }
    }
  }
}

// flux_type dispatch (Lh.f90:220-230): 1 LF, 2 HLL, 3 HLLC, 4 HLLD
template <typename T, int DIR, int FT>
__device__ __forceinline__ void riemann(T SR, T SL, const T UL[8], const T UR[8], const T FL[8],
                                        const T FR[8], T F[8], T* Ustar = nullptr) {
  if (FT == 1) lf_flux(SR, SL, UL, UR, FL, FR, F);
  else if (FT == 2) hll_flux(SR, SL, UL, UR, FL, FR, F);
  else hllcd_flux<T, DIR, FT == 4>(SR, SL, UL, UR, FL, FR, F, Ustar);
}

// calculate_Ez.f90:1-7: the corner electric field Ez = u2*B1 - u1*B2
template <typename T> __host__ __device__ __forceinline__ T calculate_ez(T u1, T u2, T B1, T B2) {
  return u2 * B1 - u1 * B2;
}
// the same from a conserved state (velocities through one reciprocal)
template <typename T> __device__ __forceinline__ T corner_ez(const T U[8]) {
  const T ir = T(1) / U[0];
  return calculate_ez(U[1] * ir, U[2] * ir, U[5], U[6]);
}

// vertex EMF: LF_Flux_2D.f90:13-60, HLL_Flux_2D.f90:17-109 (uninitialised
// corner fields resolved as corner_ez), HLLC_Flux_2D.f90 / HLLD_Flux_2D.f90:17-114
template <typename T, int FT>
__device__ __forceinline__ T vertex_emf(const T URU[8], const T ULU[8], const T URD[8], const T ULD[8]) {
  if (FT == 1) {
    const T ax = tmax(tmax(tmax(fabs(URU[1] / URU[0]), fabs(ULU[1] / ULU[0])), fabs(URD[1] / URD[0])), fabs(ULD[1] / ULD[0]));
    const T ay = tmax(tmax(tmax(fabs(URU[2] / URU[0]), fabs(ULU[2] / ULU[0])), fabs(URD[2] / URD[0])), fabs(ULD[2] / ULD[0]));
    const T EzRU = corner_ez(URU), EzLU = corner_ez(ULU), EzRD = corner_ez(URD), EzLD = corner_ez(ULD);
    return T(0.25) * (EzRU + EzLU + EzRD + EzLD) - T(0.25) * ay * (T(0.5) * (URU[5] + ULU[5]) - T(0.5) * (URD[5] + ULD[5])) +
           T(0.25) * ax * (T(0.5) * (URU[6] + URD[6]) - T(0.5) * (ULU[6] + ULD[6]));
  }
  T aRU, bRU, aLU, bLU, aRD, bRD, aLD, bLD, cRU, dRU, cLU, dLU, cRD, dRD, cLD, dLD;
  eigenvalueMm<T, 1>(URU, aRU, bRU);
  eigenvalueMm<T, 2>(URU, cRU, dRU);
  eigenvalueMm<T, 1>(ULU, aLU, bLU);
  eigenvalueMm<T, 2>(ULU, cLU, dLU);
  eigenvalueMm<T, 1>(URD, aRD, bRD);
  eigenvalueMm<T, 2>(URD, cRD, dRD);
  eigenvalueMm<T, 1>(ULD, aLD, bLD);
  eigenvalueMm<T, 2>(ULD, cLD, dLD);
  const T SR = tmax(tmax(tmax(aRU, aLU), aRD), aLD);
  const T SL = tmin(tmin(tmin(bRU, bLU), bRD), bLD);
  const T SU = tmax(tmax(tmax(cRU, cLU), cRD), cLD);
  const T SD = tmin(tmin(tmin(dRU, dLU), dRD), dLD);
  // the selection of HLL_Flux_2D.f90:97-107 (same in HLLC/HLLD)
  T FxA[8], FyA[8], FxB[8], FyB[8];
  T EzLs, EzRs, EzDs, EzUs;
  if (FT == 2) {
    T FxRU[8], FyRU[8], FxLU[8], FyLU[8], FxRD[8], FyRD[8], FxLD[8], FyLD[8];
    mhd_flux(URU, FxRU, FyRU);
    mhd_flux(ULU, FxLU, FyLU);
    mhd_flux(URD, FxRD, FyRD);
    mhd_flux(ULD, FxLD, FyLD);
    // HLL1.f90:5-14 for the four pairs; only components 6 and 7 are used
    auto hll1 = [](T sr, T sl, const T* Ua, const T* Ub, const T* Fa, const T* Fb, int n, T& Fo, T& Uo) {
      if (sr < T(0)) { Fo = Fa[n]; Uo = Ua[n]; }
      else if (sl > T(0)) { Fo = Fb[n]; Uo = Ub[n]; }
      else {
        const T id = T(1) / (sr - sl);
        Fo = (sr * Fb[n] - sl * Fa[n] + sl * sr * (Ua[n] - Ub[n])) * id;
        Uo = (sr * Ua[n] - sl * Ub[n] - (Fa[n] - Fb[n])) * id;
      }
    };
    T FU7, UU6d, FD7, UD6d, FR6, UR7d, FL6, UL7d, dummyF, BxUs, BxDs, ByRs, ByLs;
    hll1(SR, SL, URU, ULU, FxRU, FxLU, 6, FU7, UU6d);
    hll1(SR, SL, URU, ULU, FxRU, FxLU, 5, dummyF, BxUs);
    hll1(SR, SL, URD, ULD, FxRD, FxLD, 6, FD7, UD6d);
    hll1(SR, SL, URD, ULD, FxRD, FxLD, 5, dummyF, BxDs);
    hll1(SU, SD, URU, URD, FyRU, FyRD, 5, FR6, UR7d);
    hll1(SU, SD, URU, URD, FyRU, FyRD, 6, dummyF, ByRs);
    hll1(SU, SD, ULU, ULD, FyLU, FyLD, 5, FL6, UL7d);
    hll1(SU, SD, ULU, ULD, FyLU, FyLD, 6, dummyF, ByLs);
    (void)UU6d; (void)UD6d; (void)UR7d; (void)UL7d; (void)dummyF;
    const T EzRU = corner_ez(URU), EzLU = corner_ez(ULU), EzRD = corner_ez(URD), EzLD = corner_ez(ULD);
    EzRs = FR6; EzLs = FL6; EzUs = -FU7; EzDs = -FD7;
    const T den = T(1) / (T(2) * (SR - SL) * (SU - SD));
    const T BxSS = (T(2) * SR * SU * URU[5] - T(2) * SL * SU * ULU[5] + T(2) * SL * SD * ULD[5] - T(2) * SR * SD * URD[5] -
                    SR * (EzRU - EzRD) + SL * (EzLU - EzLD) - (SR - SL) * (EzUs - EzDs)) * den;
    const T BySS = (T(2) * SR * SU * URU[6] - T(2) * SL * SU * ULU[6] + T(2) * SL * SD * ULD[6] - T(2) * SR * SD * URD[6] +
                    SU * (EzRU - EzLU) - SD * (EzRD - EzLD) + (SU - SD) * (EzRs - EzLs)) * den;
    const T EzSS = T(0.25) * (EzRs + EzLs + EzUs + EzDs) - T(0.25) * SU * (BxUs - BxSS) - T(0.25) * SD * (BxDs - BxSS) +
                   T(0.25) * SR * (ByRs - BySS) + T(0.25) * SL * (ByLs - BySS);
    if (SL > T(0)) return EzLs;
    if (SR < T(0)) return EzRs;
    if (SD > T(0)) return EzDs;
    if (SU < T(0)) return EzUs;
    return EzSS;
  } else {
    // HLLC_Flux_2D / HLLD_Flux_2D: four 1D solves with the 2D bounds
    T F[8];
    mhd_flux(URU, FxA, FyA);   // RU
    mhd_flux(ULU, FxB, FyB);   // LU
    riemann<T, 1, FT>(SR, SL, ULU, URU, FxB, FxA, F);
    EzUs = -F[6];
    T FyLU[8];
#pragma unroll
    for (int n = 0; n < 8; n++) FyLU[n] = FyB[n];
    T FyRU[8];
#pragma unroll
    for (int n = 0; n < 8; n++) FyRU[n] = FyA[n];
    mhd_flux(URD, FxA, FyA);   // RD
    mhd_flux(ULD, FxB, FyB);   // LD
    riemann<T, 1, FT>(SR, SL, ULD, URD, FxB, FxA, F);
    EzDs = -F[6];
    riemann<T, 2, FT>(SU, SD, URD, URU, FyA, FyRU, F);
    EzRs = F[5];
    riemann<T, 2, FT>(SU, SD, ULD, ULU, FyB, FyLU, F);
    EzLs = F[5];
    const T EzSS = T(0.25) * (EzRs + EzLs + EzUs + EzDs);
    if (SL > T(0)) return EzLs;
    if (SR < T(0)) return EzRs;
    if (SD > T(0)) return EzDs;
    if (SU < T(0)) return EzUs;
    return EzSS;
  }
}

// pressure.f90:6
template <typename T> __device__ __forceinline__ T pressure(const T U[8]) {
  return (GAM<T>() - T(1)) * (U[4] - T(0.5) * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0] -
                               T(0.5) * (U[5] * U[5] + U[6] * U[6] + U[7] * U[7]));
}

}  // namespace mhd
