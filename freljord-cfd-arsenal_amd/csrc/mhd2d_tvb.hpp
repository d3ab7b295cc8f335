// mhd2d_tvb.hpp -- characteristic TVB limiting of one cell in one direction
// (TVB_Limiter_P2.f90:35-103) with the conservative MHD eigenvectors of
// eigenmatrix.f90:1-407.
//
// The 8x8 matrices are never held whole: rows of L and columns of R are
// generated one at a time (each depends only on a handful of shared
// scalars), which keeps the limiter inside a register budget that leaves
// room for several waves per SIMD.  The identity fallback of
// eigenmatrix.f90:397-404 (sum |R| > 1e4 or sum |L| > 1e4) needs both sums
// before any product is used, so R's columns are generated twice.
//
// The eigen-system scalars are computed without contraction and with true
// divisions, in the reference's operation order: for B parallel to n the
// reference's alpha_f = sqrt|a^2 - c_s^2| / ... (eigenmatrix.f90:58) is the
// square root of a rounding residue, so a single ulp of difference in it
// changes alpha_f by O(1) relative.  The matrix entries built from those
// scalars are well conditioned and use reciprocals and FMA.
#pragma once
#include "mhd2d_physics.hpp"
#include "mhd2d_tables.hpp"

namespace mhd {

struct EigScal {
  double n1, n2, n3, t1, t2, t3;
  double rho, u1, u2, u3, B1, B2, B3, g1;
  // right eigenvectors (eigenmatrix.f90:21-180)
  double rhosqR, aR, a2R, sqg2, sq12, sqpr, sqpor, sq1og, bNR, bstR, cfR, alphafR, alphasR, beta1R;
  // left eigenvectors (eigenmatrix.f90:185-394)
  double rhosqL, a2L, aL, cfL, alphafL, alphasL, BN, BT, gu1sq, sqgam, nen, nen2, nen3, nen31, nen51, nen52, nen71;
  double Term51, Term52, Term53, Term54, Term55, Term56, Term57;
  double Term71, Term72, Term73, Term74, Term75, Term76, Term77;
  // reciprocals shared by the matrix entries
  double irho, inen2, inen3, inen31, inen51, inen52, inen71, irc;
};

// compute_t_vector (eigenmatrix.f90:411-433) + pick_orthogonal (:451-469)
__device__ __forceinline__ void t_vector(double n1, double n2, double n3, double B1, double B2, double B3,
                                         double& t1, double& t2, double& t3) {
  MHD_NOCONTRACT
  const double nB1 = n2 * B3 - n3 * B2, nB2 = n3 * B1 - n1 * B3, nB3 = n1 * B2 - n2 * B1;
  double tp1, tp2, tp3;
  if (nB1 * nB1 + nB2 * nB2 + nB3 * nB3 < 1e-7) {
    if (fabs(n1) >= fabs(n2) && fabs(n1) >= fabs(n3)) { tp2 = 1.0; tp3 = 1.0; tp1 = (-n2 - n3) / n1; }
    else if (fabs(n2) >= fabs(n1) && fabs(n2) >= fabs(n3)) { tp1 = 1.0; tp3 = 1.0; tp2 = (-n1 - n3) / n2; }
    else { tp1 = 1.0; tp2 = 1.0; tp3 = (-n1 - n2) / n3; }
  } else {
    tp1 = nB2 * n3 - nB3 * n2;
    tp2 = nB3 * n1 - nB1 * n3;
    tp3 = nB1 * n2 - nB2 * n1;
  }
  const double tn = sqrt(tp1 * tp1 + tp2 * tp2 + tp3 * tp3);
  t1 = tp1 / tn; t2 = tp2 / tn; t3 = tp3 / tn;
}

template <int DIR>
__device__ __forceinline__ void eig_scalars(const double U[8], EigScal& s) {
  MHD_NOCONTRACT
  const double gamma = 5.0 / 3.0;
  s.n1 = DIR == 1 ? 1.0 : 0.0; s.n2 = DIR == 1 ? 0.0 : 1.0; s.n3 = 0.0;
  const double n1 = s.n1, n2 = s.n2, n3 = s.n3;
  s.rho = U[0];
  // TVB_Limiter_P2.f90:26-33
  s.u1 = U[1] / s.rho; s.u2 = U[2] / s.rho; s.u3 = U[3] / s.rho;
  s.B1 = U[5]; s.B2 = U[6]; s.B3 = U[7];
  const double rho = s.rho, B1 = s.B1, B2 = s.B2, B3 = s.B3, u1 = s.u1, u2 = s.u2, u3 = s.u3;
  const double p = (gamma - 1.0) * (U[4] - 0.5 * rho * (u1 * u1 + u2 * u2 + u3 * u3) - 0.5 * (B1 * B1 + B2 * B2 + B3 * B3));
  t_vector(n1, n2, n3, B1, B2, B3, s.t1, s.t2, s.t3);
  const double t1 = s.t1, t2 = s.t2, t3 = s.t3;
  s.g1 = gamma - 1.0;
  // R part (eigenmatrix.f90:27-60)
  s.rhosqR = sqrt(fabs(rho));
  s.a2R = fabs(gamma * p / rho);
  s.aR = sqrt(s.a2R);
  s.sqg2 = sqrt(1.0 / (2.0 * gamma));
  s.sq12 = sqrt(0.5);
  s.sqpr = sqrt(fabs(p)) / rho;
  s.sqpor = sqrt(fabs(p / rho));
  s.sq1og = sqrt(1.0 / gamma);
  {
    const double b1s = B1 / s.rhosqR, b2s = B2 / s.rhosqR, b3s = B3 / s.rhosqR;
    s.bNR = b1s * n1 + b2s * n2 + b3s * n3;
    const double d = s.a2R + (b1s * b1s + b2s * b2s + b3s * b3s);
    s.cfR = sqrt(0.5 * fabs(d + sqrt(fabs(d * d - 4.0 * s.a2R * (s.bNR * s.bNR)))));
    const double cs = sqrt(0.5 * fabs(d - sqrt(fabs(d * d - 4.0 * s.a2R * (s.bNR * s.bNR)))));
    s.beta1R = copysign(1.0, (b1s * n1 + b2s * n2 + b3s * n3) * 1.0);
    if (fabs(s.cfR * s.cfR - cs * cs) <= 1e-12) {
      s.alphafR = 0.70710678118654746;   // sin(atan(1)) in IEEE double
      s.alphasR = 0.70710678118654757;   // cos(atan(1))
    } else {
      s.alphafR = sqrt(fabs(s.a2R - cs * cs)) / sqrt(fabs(s.cfR * s.cfR - cs * cs));
      s.alphasR = sqrt(fabs(s.cfR * s.cfR - s.a2R)) / sqrt(fabs(s.cfR * s.cfR - cs * cs));
    }
    s.bstR = (b1s * t1 + b2s * t2 + b3s * t3);
  }
  // L part (eigenmatrix.f90:187-227)
  s.rhosqL = sqrt(rho);
  s.a2L = fabs(gamma * p / rho);
  s.aL = sqrt(s.a2L);
  const double b1s = B1 / s.rhosqL, b2s = B2 / s.rhosqL, b3s = B3 / s.rhosqL;
  const double BNs = b1s * n1 + b2s * n2 + b3s * n3;
  s.BN = (B1 * n1 + B2 * n2 + B3 * n3);
  const double d = s.a2L + (b1s * b1s + b2s * b2s + b3s * b3s);
  const double cf = sqrt(0.5 * fabs(d + sqrt(fabs(d * d - 4.0 * s.a2L * (BNs * BNs)))));
  const double cs = sqrt(0.5 * fabs(d - sqrt(fabs(d * d - 4.0 * s.a2L * (BNs * BNs)))));
  s.cfL = cf;
  double alphaf, alphas;
  if (fabs(cf * cf - cs * cs) <= 1e-12) {
    alphaf = 0.70710678118654746;
    alphas = 0.70710678118654757;
  } else {
    alphaf = sqrt(fabs(s.a2L - cs * cs)) / sqrt(fabs(cf * cf - cs * cs));
    alphas = sqrt(fabs(cf * cf - s.a2L)) / sqrt(fabs(cf * cf - cs * cs));
  }
  s.alphafL = alphaf; s.alphasL = alphas;
  const double twosq = sqrt(2.0);
  const double beta1 = copysign(1.0, (BNs) * 1.0);
  s.sqgam = sqrt(s.g1 / gamma);
  const double TxN1 = n3 * t2 - n2 * t3, TxN2 = n1 * t3 - n3 * t1, TxN3 = n2 * t1 - n1 * t2;
  s.BT = (B1 * t1 + B2 * t2 + B3 * t3);
  s.gu1sq = sqrt(1.0 / gamma);
  const double a = s.aL, rhosq = s.rhosqL, BN = s.BN, BT = s.BT, gu1sq = s.gu1sq;
  s.nen = (n3 * n3 * (t1 * t1 + t2 * t2) - 2.0 * n1 * n3 * t1 * t3 - 2.0 * n2 * t2 * (n1 * t1 + n3 * t3) +
           n2 * n2 * (t1 * t1 + t3 * t3) + n1 * n1 * (t2 * t2 + t3 * t3));
  const double nen = s.nen;
  s.nen2 = a * gu1sq * nen;
  s.nen3 = sqrt(2.0) * sqrt(fabs(p)) * nen;
  s.nen31 = s.nen3 / sqrt(fabs(rho));
  s.Term51 = rho * cf * ((n2 * t1 * t2 + n3 * t1 * t3 - n1 * (t2 * t2 + t3 * t3)) * rhosq * (cf * cf) * alphaf -
                         a * BN * alphas * (-(n2 * n2) * t1 + n1 * n2 * t2 + n3 * TxN2));
  s.Term52 = rho * cf * ((-t2 * (n1 * t1 + n3 * t3) + n2 * (t1 * t1 + t3 * t3)) * rhosq * (cf * cf) * alphaf -
                         a * BN * alphas * (-n1 * n2 * t1 + (n1 * n1) * t2 + n3 * TxN1));
  s.Term53 = rho * cf * ((n3 * (t1 * t1 + t2 * t2) - (n1 * t1 + n2 * t2) * t3) * rhosq * (cf * cf) * alphaf -
                         a * (-n1 * n3 * t1 + (n1 * n1) * t3 + n2 * (-n3 * t2 + n2 * t3)) * BN * alphas);
  s.Term54 = alphaf / (twosq * (a * a) * gu1sq * rhosq * (alphaf * alphaf + alphas * alphas));
  s.Term55 = ((n2 * n2) * t1 - n1 * n2 * t2 + n3 * (n3 * t1 - n1 * t3)) * alphas;
  s.Term56 = alphas * (-n1 * n2 * t1 + (n1 * n1) * t2 + n3 * TxN1);
  s.Term57 = (-n1 * n3 * t1 + (n1 * n1) * t3 + n2 * (-n3 * t2 + n2 * t3)) * alphas;
  s.nen51 = twosq * a * nen * gu1sq *
            (a * (BN * BN) * (alphas * alphas) + rhosq * (cf * cf) * alphaf * (a * rhosq * alphaf + BT * alphas));
  s.nen52 = twosq * a * gu1sq * (alphaf * alphaf + alphas * alphas) * nen;
  s.Term71 = rho * cf * (a * ((n2 * n2) * t1 - n1 * n2 * t2 + n3 * (n3 * t1 - n1 * t3)) * rhosq * alphaf +
                         ((B3 * n2 * t1 - B2 * n3 * t1 - B3 * n1 * t2 + B2 * n1 * t3) * (-n3 * t2 + n2 * t3) +
                          B1 * ((n2 * n2) * (t1 * t1) + (n3 * n3) * (t1 * t1) - 2.0 * n1 * n2 * t1 * t2 -
                                2.0 * n1 * n3 * t1 * t3 + (n1 * n1) * (t2 * t2 + t3 * t3))) * alphas);
  s.Term72 = rho * cf * (((n3 * t1 - n1 * t3) * (B3 * n2 * t1 - B3 * n1 * t2 + B1 * n3 * t2 - B1 * n2 * t3) +
                          B2 * (((n1 * n1) + (n3 * n3)) * (t2 * t2) - 2.0 * n2 * t2 * (n1 * t1 + n3 * t3) +
                                (n2 * n2) * (t1 * t1 + t3 * t3))) * alphas +
                         a * rhosq * alphaf * (-n1 * n2 * t1 + (n1 * n1) * t2 + n3 * TxN1));
  s.Term73 = rho * cf * (a * (-n1 * n3 * t1 + (n1 * n1) * t3 + n2 * (-n3 * t2 + n2 * t3)) * rhosq * alphaf +
                         alphas * (B3 * ((n3 * n3) * (t1 * t1 + t2 * t2) - 2.0 * n3 * (n1 * t1 + n2 * t2) * t3 +
                                         ((n1 * n1) + (n2 * n2)) * (t3 * t3)) +
                                   B2 * (n3 * t1 - n1 * t3) * TxN3 + B1 * (-n3 * t2 + n2 * t3) * TxN3));
  s.Term74 = alphas / (twosq * (a * a) * gu1sq * rhosq * (alphaf * alphaf + alphas * alphas));
  s.Term75 = -alphaf * ((n2 * n2) * t1 - n1 * n2 * t2 + n3 * (n3 * t1 - n1 * t3));
  s.Term76 = -alphaf * (-n1 * n2 * t1 + (n1 * n1) * t2 + n3 * TxN1);
  s.Term77 = -alphaf * (-n1 * n3 * t1 + (n1 * n1) * t3 + n2 * (-n3 * t2 + n2 * t3));
  s.nen71 = twosq * beta1 * nen * gu1sq *
            (a * (BN * BN) * (alphas * alphas) + rhosq * (cf * cf) * alphaf * (a * rhosq * alphaf + BT * alphas));
  s.irho = 1.0 / rho;
  s.inen2 = 1.0 / s.nen2;
  s.inen3 = 1.0 / s.nen3;
  s.inen31 = 1.0 / s.nen31;
  s.inen51 = 1.0 / s.nen51;
  s.inen52 = 1.0 / s.nen52;
  s.inen71 = 1.0 / s.nen71;
  s.irc = 1.0 / (s.rhosqR * s.cfR);
}

// row m (1..8) of the conservative left eigenvectors evl(m, 1:8)
__device__ __forceinline__ void eig_lrow(const EigScal& s, int m, double e[8]) {
  const double n1 = s.n1, n2 = s.n2, n3 = s.n3, t1 = s.t1, t2 = s.t2, t3 = s.t3;
  const double TxN1 = n3 * t2 - n2 * t3, TxN2 = n1 * t3 - n3 * t1, TxN3 = n2 * t1 - n1 * t2;
  double l[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (m == 1) {
    l[0] = 1.0 / (s.sqgam * s.rhosqL);
    l[4] = -1.0 / (s.a2L * s.sqgam * s.rhosqL);
  } else if (m == 2) {
    l[5] = (-n2 * t1 * t2 - n3 * t1 * t3 + n1 * (t2 * t2 + t3 * t3)) * s.inen2;
    l[6] = (-t2 * (n1 * t1 + n3 * t3) + n2 * (t1 * t1 + t3 * t3)) * s.inen2;
    l[7] = (n3 * (t1 * t1 + t2 * t2) - (n1 * t1 + n2 * t2) * t3) * s.inen2;
  } else if (m == 3 || m == 4) {
    const double sg = m == 3 ? 1.0 : -1.0;
    l[1] = sg * (s.rho * (-TxN1) * s.inen3);
    l[2] = sg * (s.rho * (-TxN2) * s.inen3);
    l[3] = sg * (s.rho * (-TxN3) * s.inen3);
    l[5] = TxN1 * s.inen31; l[6] = TxN2 * s.inen31; l[7] = TxN3 * s.inen31;
  } else if (m == 5 || m == 6) {
    const double sg = m == 5 ? 1.0 : -1.0;
    l[1] = sg * (-s.Term51 * s.inen51);
    l[2] = sg * (s.Term52 * s.inen51);
    l[3] = sg * (s.Term53 * s.inen51);
    l[4] = s.Term54;
    l[5] = s.Term55 * s.inen52; l[6] = s.Term56 * s.inen52; l[7] = s.Term57 * s.inen52;
  } else {
    const double sg = m == 7 ? 1.0 : -1.0;
    l[1] = sg * (s.Term71 * s.inen71);
    l[2] = sg * (s.Term72 * s.inen71);
    l[3] = sg * (s.Term73 * s.inen71);
    l[4] = s.Term74;
    l[5] = s.Term75 * s.inen52; l[6] = s.Term76 * s.inen52; l[7] = s.Term77 * s.inen52;
  }
  const double g1 = s.g1, g12 = g1 * g1, gr = g1 * s.irho;
  const double u1 = s.u1, u2 = s.u2, u3 = s.u3;
  e[0] = l[0] * g1 - (l[1] * u1 + l[2] * u2 + l[3] * u3) * gr + l[4] * g12 * (u1 * u1 + u2 * u2 + u3 * u3) * .5;
  e[1] = -l[4] * u1 * g12 + l[1] * gr;
  e[2] = -l[4] * u2 * g12 + l[2] * gr;
  e[3] = -l[4] * u3 * g12 + l[3] * gr;
  e[4] = l[4] * g12;
  e[5] = l[5] * g1 - s.B1 * l[4] * g12;
  e[6] = l[6] * g1 - s.B2 * l[4] * g12;
  e[7] = l[7] * g1 - s.B3 * l[4] * g12;
}

// column m (1..8) of the conservative right eigenvectors evr(1:8, m)
__device__ __forceinline__ void eig_rcol(const EigScal& s, int m, double e[8]) {
  const double n1 = s.n1, n2 = s.n2, n3 = s.n3, t1 = s.t1, t2 = s.t2, t3 = s.t3;
  const double TxN1 = n3 * t2 - n2 * t3, TxN2 = n1 * t3 - n3 * t1, TxN3 = n2 * t1 - n1 * t2;
  const double a = s.aR, a2 = s.a2R, rhosq = s.rhosqR, sqg2 = s.sqg2, cf = s.cfR;
  const double alphaf = s.alphafR, alphas = s.alphasR, bN = s.bNR, bst = s.bstR;
  double r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (m == 1) {
    r[0] = sqrt(s.g1 / (5.0 / 3.0)) * rhosq;
  } else if (m == 2) {
    r[5] = s.sq1og * a * n1; r[6] = s.sq1og * a * n2; r[7] = s.sq1og * a * n3;
  } else if (m == 3 || m == 4) {
    const double sg = m == 3 ? 1.0 : -1.0;
    r[1] = sg * (-s.sq12 * (s.sqpr * TxN1));
    r[2] = sg * (-s.sq12 * (s.sqpr * TxN2));
    r[3] = sg * (-s.sq12 * (s.sqpr * TxN3));
    r[5] = s.sq12 * s.sqpor * TxN1; r[6] = s.sq12 * s.sqpor * TxN2; r[7] = s.sq12 * s.sqpor * TxN3;
  } else if (m == 5 || m == 6) {
    const double sg = m == 5 ? 1.0 : -1.0;
    const double c = sqg2 * s.irc;
    r[0] = sqg2 * alphaf * rhosq;
    r[1] = sg * (c * (alphaf * a2 * n1 + alphas * a * (bst * n1 - bN * t1)));
    r[2] = sg * (c * (alphaf * a2 * n2 + alphas * a * (bst * n2 - bN * t2)));
    r[3] = sg * (c * (alphaf * a2 * n3 + alphas * a * (bst * n3 - bN * t3)));
    r[4] = sqg2 * alphaf * rhosq * a2;
    r[5] = sqg2 * alphas * a * t1; r[6] = sqg2 * alphas * a * t2; r[7] = sqg2 * alphas * a * t3;
  } else {
    const double sg = m == 7 ? 1.0 : -1.0;
    const double c = s.beta1R * sqg2 * s.irc;
    r[0] = sqg2 * alphas * rhosq;
    r[1] = sg * (c * (alphaf * (cf * cf) * t1 + a * n1 * alphas * bN));
    r[2] = sg * (c * (alphaf * (cf * cf) * t2 + alphas * a * bN * n2));
    r[3] = sg * (c * (alphaf * (cf * cf) * t3 + alphas * a * bN * n3));
    r[4] = (a * a) * sqg2 * alphas * rhosq;
    r[5] = -sqg2 * alphaf * a * t1; r[6] = -sqg2 * alphaf * a * t2; r[7] = -sqg2 * alphaf * a * t3;
  }
  constexpr double ig1 = 1.0 / (5.0 / 3.0 - 1.0);
  const double rho = s.rho, u1 = s.u1, u2 = s.u2, u3 = s.u3;
  e[0] = r[0] * ig1;
  e[1] = (r[0] * u1 + r[1] * rho) * ig1;
  e[2] = (r[0] * u2 + r[2] * rho) * ig1;
  e[3] = (r[0] * u3 + r[3] * rho) * ig1;
  e[4] = (r[4] * ig1 + s.B1 * r[5] + s.B2 * r[6] + s.B3 * r[7] + 0.5 * r[0] * (u1 * u1 + u2 * u2 + u3 * u3) +
          r[1] * u1 * rho + r[2] * u2 * rho + r[3] * u3 * rho) * ig1;
  e[5] = r[5] * ig1;
  e[6] = r[6] * ig1;
  e[7] = r[7] * ig1;
}

// minmod.f90:5-27 for one characteristic field
__device__ __forceinline__ double minmod1(double du, double dr, double dl, double Mh2) {
  if (fabs(du) <= Mh2) return du;
  const double a = copysign(1.0, du), b = copysign(1.0, dr), c = copysign(1.0, dl);
  const double s = (a + b + c) / 3.0;
  if (fabs(s) == 1.0) return s * fmin(fmin(fabs(du), fabs(dr)), fabs(dl));
  return 0.0;
}

// one direction of TVB_Limiter_P2.f90 (:36-68 for x, :71-103 for y):
// modified edge deviations R1mod (= DeltaUR1mod / DeltaUU1mod) and L1mod.
template <int DIR>
__device__ __forceinline__ void tvb_dir(const double U[8], const double dR[8], const double dL[8],
                                        const double r1[8], const double l1[8], double Mh2,
                                        double R1mod[8], double L1mod[8]) {
  EigScal s;
  eig_scalars<DIR>(U, s);
  double cR[8], cL[8], cr[8], cl[8];
  double sumL = 0.0, sumR = 0.0;
#pragma unroll
  for (int m = 1; m <= 8; m++) {
    double e[8];
    eig_lrow(s, m, e);
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      a = a + e[k] * dR[k];
      b = b + e[k] * dL[k];
      c = c + e[k] * r1[k];
      d = d + e[k] * l1[k];
      sumL += fabs(e[k]);
    }
    cR[m - 1] = a; cL[m - 1] = b; cr[m - 1] = c; cl[m - 1] = d;
  }
#pragma unroll
  for (int m = 1; m <= 8; m++) {
    double e[8];
    eig_rcol(s, m, e);
#pragma unroll
    for (int k = 0; k < 8; k++) sumR += fabs(e[k]);
  }
  const bool ident = sumR > 1e4 || sumL > 1e4;
  if (ident) {
#pragma unroll
    for (int k = 0; k < 8; k++) { cR[k] = dR[k]; cL[k] = dL[k]; cr[k] = r1[k]; cl[k] = l1[k]; }
  }
  double mr[8], ml[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    mr[k] = minmod1(cr[k], cR[k], cL[k], Mh2);
    ml[k] = minmod1(cl[k], cR[k], cL[k], Mh2);
  }
  if (ident) {
#pragma unroll
    for (int k = 0; k < 8; k++) { R1mod[k] = mr[k]; L1mod[k] = ml[k]; }
    return;
  }
#pragma unroll
  for (int k = 0; k < 8; k++) { R1mod[k] = 0.0; L1mod[k] = 0.0; }
#pragma unroll
  for (int m = 1; m <= 8; m++) {
    double e[8];
    eig_rcol(s, m, e);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      R1mod[k] = R1mod[k] + e[k] * mr[m - 1];
      L1mod[k] = L1mod[k] + e[k] * ml[m - 1];
    }
  }
}

}  // namespace mhd
