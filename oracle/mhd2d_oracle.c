/*
 * mhd2d_oracle.c -- CPU restatement of GDF_Balsara_P2_0316 (2D ideal MHD,
 * P2 modal DG, Balsara divergence-free reconstruction, constrained transport
 * with a corner EMF, TVB limiter, optional positivity limiter, SSP-RK3).
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker for the HIP path, never the
 * product: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.
 *
 * It follows the reference operation by operation (same summation order,
 * same association of products) so that, on x86-64 without FMA, it agrees
 * with a serial flang -fdefault-real-8 build of the reference to the last
 * bits in most places.  Each function cites the reference file:line it
 * restates.  Deliberate resolutions of reference defects:
 *   - Lh.f90:268-298 races under OpenMP; this sums each du element serially.
 *   - HLL_Flux_2D.f90:8,61-62 reads EzRU..EzLD uninitialised; here they are
 *     the corner electric fields v*B1 - u*B2 (calculate_Ez.f90:5), as
 *     LF_Flux_2D.f90:18-21 computes them.
 *   - KXRCF_Detector.f90 does not compile (uG undeclared) and never resets
 *     its flags; here uG is a per-cell local, flags are reset on every call
 *     and Ck is a parameter (the shipped Ck = 0 flags every cell).
 */
#include "mhd2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NG 5   /* NumGLP */
#define NP 10  /* dimPk  */
#define NE 8   /* NumEq  */

static const double GAMMA = 5.0 / 3.0;          /* com.txt:7 */
#define GAMMA1 (GAMMA - 1.0)                     /* com.txt:8 */

typedef struct {
  double lambda[NG], lambdaL[NG], weight[NG], mm[NP];
  double phiG[NP][NG][NG];          /* [d][i][j] = phiG(i,j,d) */
  double phiGLL[2][NP][NG][NG];     /* [fam][d][i][j] */
  double phixG[NP][NG][NG], phiyG[NP][NG][NG];
  double phiGR[NP][NG], phiGL[NP][NG], phiGU[NP][NG], phiGD[NP][NG];
  double phiRU[NP], phiLU[NP], phiRD[NP], phiLD[NP];
  double EzG[3][NG], EzxG[3][NG], EzyG[3][NG];
  double mmE[3], EzR[3], EzL[3], EzU[3], EzD[3];
} tables_t;

struct orc_state {
  orc_params p;
  double hx, hy, hx1, hy1;
  tables_t T;
  double t, dt, umax, totaldiv, pmin, alphax, alphay;
  long count;
  size_t nuh, nbx, nby;
  double *uh, *du, *uI, *uII, *uh0, *uhmod;
  double *Bx, *By, *dEz1, *dEz2, *BxI, *ByI, *BxII, *ByII, *Bx0, *By0;
  double *Fxhat, *Fyhat, *EzRL, *EzUD, *EzVertex;
  int32_t *trouble;
};

/* ---- index helpers: reference (Fortran) layouts --------------------- */
#define NXr (s->p.Nx)
#define NYr (s->p.Ny)
/* uh(0:Nx+1, 0:Ny+1, 1:10, 1:8) */
#define UHI(i, j, d, n) \
  ((size_t)(i) + (size_t)(NXr + 2) * ((size_t)(j) + (size_t)(NYr + 2) * ((size_t)((d)-1) + 10u * (size_t)((n)-1))))
/* Bx(0:Nx, 1:Ny, 1:3) */
#define BXI(i, j, d) ((size_t)(i) + (size_t)(NXr + 1) * ((size_t)((j)-1) + (size_t)NYr * (size_t)((d)-1)))
/* By(1:Nx, 0:Ny, 1:3) */
#define BYI(i, j, d) ((size_t)((i)-1) + (size_t)NXr * ((size_t)(j) + (size_t)(NYr + 1) * (size_t)((d)-1)))
/* EzVertex(0:Nx, 0:Ny) */
#define EVI(i, j) ((size_t)(i) + (size_t)(NXr + 1) * (size_t)(j))
/* private layouts for the face fluxes (not compared element-wise) */
#define FXI(i, j) (((size_t)(i) + (size_t)(NXr + 1) * (size_t)((j)-1)) * (NG * NE))
#define FYI(i, j) (((size_t)((i)-1) + (size_t)NXr * (size_t)(j)) * (NG * NE))
#define EXI(i, j) (((size_t)(i) + (size_t)(NXr + 1) * (size_t)((j)-1)) * NG)
#define EYI(i, j) (((size_t)((i)-1) + (size_t)NXr * (size_t)(j)) * NG)

/* active (mode, variable) pairs: Lh.f90:59 */
static inline int active(int d, int n) {
  return d <= 6 || (n == 6 && (d == 7 || d == 9)) || (n == 7 && (d == 8 || d == 10));
}

static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fmin2(double a, double b) { return a < b ? a : b; }
/* Fortran SIGN(1d0, x) */
static inline double fsign1(double x) { return copysign(1.0, x); }

/* ---------------------------------------------------------------------- */
/* get_GLP.f90:38-56 (NumGLP = 5)                                          */
static void get_GLP(tables_t *T) {
  T->lambda[0] = -0.9061798459386639927976269;
  T->lambda[1] = -0.5384693101056830910363144;
  T->lambda[2] = 0.0;
  T->lambda[3] = 0.5384693101056830910363144;
  T->lambda[4] = 0.9061798459386639927976269;
  T->weight[0] = 0.2369268850561890875142640;
  T->weight[1] = 0.4786286704993664680412915;
  T->weight[2] = 0.5688888888888888888888889;
  T->weight[3] = 0.4786286704993664680412915;
  T->weight[4] = 0.2369268850561890875142640;
  /* lambdaL is not declared in com.txt, so implicit typing makes it an
     INTEGER array: get_GLP.f90:52-56 stores the Gauss-Lobatto nodes
     truncated to -1, 0, 0, 0, 1.  They are used only by pp_Limiter
     (phiGLL), which therefore checks positivity at those points. */
  T->lambdaL[0] = -1.0;
  T->lambdaL[1] = 0.0;
  T->lambdaL[2] = 0.0;
  T->lambdaL[3] = 0.0;
  T->lambdaL[4] = 1.0;
}

/* get_basis.f90:5-189: Legendre products on [-1,1]^2 */
static void get_basis(tables_t *T, double hx1, double hy1) {
  const double *l = T->lambda, *lL = T->lambdaL;
  const double t3 = 1.0 / 3.0, t23 = 2.0 / 3.0, t25 = 2.0 / 5.0;
  const double RU[NP] = {1, 1, 1, t23, 1, t23, t25, t23, t23, t25};
  const double LU[NP] = {1, -1, 1, t23, -1, t23, -t25, t23, -t23, t25};
  const double RD[NP] = {1, 1, -1, t23, -1, t23, t25, -t23, t23, -t25};
  const double LD[NP] = {1, -1, -1, t23, 1, t23, -t25, -t23, -t23, -t25};
  memcpy(T->phiRU, RU, sizeof RU);
  memcpy(T->phiLU, LU, sizeof LU);
  memcpy(T->phiRD, RD, sizeof RD);
  memcpy(T->phiLD, LD, sizeof LD);
  const double mm[NP] = {1.0, 1.0 / 3.0, 1.0 / 3.0, 4.0 / 45.0, 1.0 / 9.0,
                         4.0 / 45.0, 4.0 / 175.0, 4.0 / 135.0, 4.0 / 135.0, 4.0 / 175.0};
  memcpy(T->mm, mm, sizeof mm);
  for (int i = 0; i < NG; i++) {
    for (int j = 0; j < NG; j++) {
      double li = l[i], lj = l[j], Li = lL[i], Lj = lL[j];
      double P2i = li * li - t3, P2j = lj * lj - t3;
      double P3i = li * li * li - 3.0 * li / 5.0, P3j = lj * lj * lj - 3.0 * lj / 5.0;
      double v[NP] = {1.0, li, lj, P2i, li * lj, P2j, P3i, P2i * (lj), (li) * P2j, P3j};
      double g1[NP] = {1.0, Li, lj, Li * Li - t3, Li * lj, P2j,
                       Li * Li * Li - 3.0 * Li / 5.0, (Li * Li - t3) * (lj), (Li) * P2j, P3j};
      double g2[NP] = {1.0, li, Lj, P2i, li * Lj, Lj * Lj - t3,
                       P3i, P2i * (Lj), (li) * (Lj * Lj - t3), Lj * Lj * Lj - 3.0 * Lj / 5.0};
      double dx[NP] = {0.0, 1.0 / hx1, 0.0, 2.0 * li / hx1, lj / hx1, 0.0,
                       (3 * (li * li) - 3.0 / 5.0) / hx1, 2.0 * li * lj / hx1, P2j / hx1, 0.0};
      double dy[NP] = {0.0, 0.0, 1.0 / hy1, 0.0, li / hy1, 2.0 * lj / hy1,
                       0.0, P2i / hy1, 2.0 * li * lj / hy1, (3 * (lj * lj) - 3.0 / 5.0) / hy1};
      for (int d = 0; d < NP; d++) {
        T->phiG[d][i][j] = v[d];
        T->phiGLL[0][d][i][j] = g1[d];
        T->phiGLL[1][d][i][j] = g2[d];
        T->phixG[d][i][j] = dx[d];
        T->phiyG[d][i][j] = dy[d];
      }
    }
    /* face traces, get_basis.f90:49-160 (index i plays j1 or i1) */
    double li = l[i], P2 = li * li - t3, P3 = li * li * li - 3.0 * li / 5.0;
    double R[NP] = {1, 1, li, t23, li, P2, t25, t23 * (li), P2, P3};
    double Lf[NP] = {1, -1, li, t23, -li, P2, -t25, t23 * (li), -P2, P3};
    double U[NP] = {1, li, 1, P2, li, t23, P3, P2, li * t23, t25};
    double D[NP] = {1, li, -1, P2, -li, t23, P3, -P2, li * t23, -t25};
    for (int d = 0; d < NP; d++) {
      T->phiGR[d][i] = R[d];
      T->phiGL[d][i] = Lf[d];
      T->phiGU[d][i] = U[d];
      T->phiGD[d][i] = D[d];
    }
    T->EzG[0][i] = 1.0;
    T->EzG[1][i] = li;
    T->EzG[2][i] = P2;
    T->EzxG[0][i] = 0.0;
    T->EzxG[1][i] = 1 / hx1;
    T->EzxG[2][i] = 2 * li / hx1;
    T->EzyG[0][i] = 0.0;
    T->EzyG[1][i] = 1 / hy1;
    T->EzyG[2][i] = 2 * li / hy1;
  }
  T->EzR[0] = 1; T->EzR[1] = 1; T->EzR[2] = t23;
  T->EzL[0] = 1; T->EzL[1] = -1; T->EzL[2] = t23;
  memcpy(T->EzU, T->EzR, sizeof T->EzR);
  memcpy(T->EzD, T->EzL, sizeof T->EzL);
  T->mmE[0] = 1.0; T->mmE[1] = 1.0 / 3.0; T->mmE[2] = 4.0 / 45.0;
}

/* ---------------------------------------------------------------------- */
/* pointwise physics                                                      */

/* eigenvalueMm.f90:1-24 */
void orc_eigenvalueMm(const double U[8], int n1, int n2, double *amax, double *amin) {
  double rho = U[0], E = U[4], B1 = U[5], B2 = U[6], B3 = U[7];
  double n3 = 0.0;
  double u = U[1] / rho, v = U[2] / rho, w = U[3] / rho;
  double BP = B1 * B1 + B2 * B2 + B3 * B3;
  double Bn = B1 * n1 + B2 * n2 + B3 * n3;
  double un = u * n1 + v * n2 + w * n3;
  double p = GAMMA1 * (E - 0.5 * rho * (u * u + v * v + w * w) - 0.5 * BP);
  double c = sqrt(fabs(GAMMA * p / rho));
  double c2 = c * c;
  double q = c2 + BP / rho;
  double cf = sqrt(fabs(0.5 * (c2 + BP / rho + sqrt(q * q - 4 * c2 * (Bn * Bn) / rho))));
  *amax = un + cf;
  *amin = un - cf;
}

/* physical flux as Lh.f90:22-53 (volume) and :89-202 (traces).  ydown=1
   reproduces the pMD sign defect of Lh.f90:189 (-wMD**2). */
static inline void lh_flux(const double U[8], double Fx[8], double Fy[8], int ydown) {
  double rho = U[0];
  double u = U[1] / rho, v = U[2] / rho, w = U[3] / rho;
  double E = U[4], B1 = U[5], B2 = U[6], B3 = U[7];
  double B2s = B1 * B1 + B2 * B2 + B3 * B3;
  double p = ydown ? GAMMA1 * (E - 0.5 * rho * (u * u + v * v - w * w) - 0.5 * B2s)
                   : GAMMA1 * (E - 0.5 * rho * (u * u + v * v + w * w) - 0.5 * B2s);
  double S = p + 0.5 * B2s;
  double T = E + S;
  double K = u * B1 + v * B2 + w * B3;
  if (Fx) {
    Fx[0] = U[1];
    Fx[1] = rho * (u * u) + S - B1 * B1;
    Fx[2] = rho * u * v - B1 * B2;
    Fx[3] = rho * u * w - B1 * B3;
    Fx[4] = T * u - K * B1;
    Fx[5] = 0.0;
    Fx[6] = u * B2 - v * B1;
    Fx[7] = u * B3 - w * B1;
  }
  if (Fy) {
    Fy[0] = U[2];
    Fy[1] = rho * u * v - B1 * B2;
    Fy[2] = rho * (v * v) + S - B2 * B2;
    Fy[3] = rho * v * w - B2 * B3;
    Fy[4] = T * v - K * B2;
    Fy[5] = v * B1 - u * B2;
    Fy[6] = 0.0;
    Fy[7] = v * B3 - w * B2;
  }
}

/* MHD_flux.f90:7-40 (used by the corner solvers) */
void orc_mhd_flux(const double U[8], double Fx[8], double Fy[8]) {
  double rho = U[0];
  double u = U[1] / rho, v = U[2] / rho, w = U[3] / rho;
  double E = U[4], B1 = U[5], B2 = U[6], B3 = U[7];
  double p = GAMMA1 * (E - 0.5 * rho * (u * u + v * v + w * w) - 0.5 * (B1 * B1 + B2 * B2 + B3 * B3));
  double S = p + 0.5 * (B1 * B1 + B2 * B2 + B3 * B3);
  double T = E + S;
  double K = u * B1 + v * B2 + w * B3;
  Fx[0] = U[1];
  Fx[1] = rho * (u * u) + S - B1 * B1;
  Fx[2] = rho * u * v - B1 * B2;
  Fx[3] = rho * u * w - B1 * B3;
  Fx[4] = T * u - K * B1;
  Fx[5] = 0.0;
  Fx[6] = u * B2 - v * B1;
  Fx[7] = u * B3 - w * B1;
  Fy[0] = U[2];
  Fy[1] = rho * u * v - B1 * B2;
  Fy[2] = rho * (v * v) + S - B2 * B2;
  Fy[3] = rho * w * v - B3 * B2;
  Fy[4] = T * v - K * B2;
  Fy[5] = v * B1 - u * B2;
  Fy[6] = 0.0;
  Fy[7] = v * B3 - w * B2;
}

/* LF_Flux.f90:5 */
static void lf_flux(double SR, double SL, const double UL[8], const double UR[8],
                    const double FL[8], const double FR[8], double F[8]) {
  double a = fmax2(fabs(SR), fabs(SL));
  for (int n = 0; n < 8; n++) F[n] = 0.5 * (FR[n] + FL[n] - a * (UR[n] - UL[n]));
}

/* HLL_Flux.f90:5-11 */
static void hll_flux(double SR, double SL, const double UL[8], const double UR[8],
                     const double FL[8], const double FR[8], double F[8]) {
  if (SR < 0) {
    for (int n = 0; n < 8; n++) F[n] = FR[n];
  } else if (SL > 0) {
    for (int n = 0; n < 8; n++) F[n] = FL[n];
  } else {
    for (int n = 0; n < 8; n++)
      F[n] = (SR * FL[n] - SL * FR[n] + SL * SR * (UR[n] - UL[n])) / (SR - SL);
  }
}

/* HLLC_Flux.f90:9-106 */
static void hllc_flux(int dir, double SR, double SL, const double UL1[8], const double UR1[8],
                      const double FL1[8], const double FR1[8], double F[8], double Ustar[8]) {
  if (SR < 0) {
    for (int n = 0; n < 8; n++) { F[n] = FR1[n]; Ustar[n] = UR1[n]; }
    return;
  }
  if (SL > 0) {
    for (int n = 0; n < 8; n++) { F[n] = FL1[n]; Ustar[n] = UL1[n]; }
    return;
  }
  double URs[8], ULs[8];
  for (int n = 0; n < 8; n++) Ustar[n] = (SR * UR1[n] - SL * UL1[n] + FL1[n] - FR1[n]) / (SR - SL);
  double rhoR = UR1[0], u1R = UR1[1] / rhoR, u2R = UR1[2] / rhoR, u3R = UR1[3] / rhoR, ER = UR1[4];
  double B1R = UR1[5], B2R = UR1[6], B3R = UR1[7];
  double rhoL = UL1[0], u1L = UL1[1] / rhoL, u2L = UL1[2] / rhoL, u3L = UL1[3] / rhoL, EL = UL1[4];
  double B1L = UL1[5], B2L = UL1[6], B3L = UL1[7];
  double B1s = Ustar[5], B2s = Ustar[6], B3s = Ustar[7];
  double Ss, uRb, uLb, BRb, BLb, Bsb;
  if (dir == 1) { Ss = Ustar[1] / Ustar[0]; uRb = u1R; uLb = u1L; BRb = B1R; BLb = B1L; Bsb = B1s; }
  else { Ss = Ustar[2] / Ustar[0]; uRb = u2R; uLb = u2L; BRb = B2R; BLb = B2L; Bsb = B2s; }
  double rhoRs = rhoR * (SR - uRb) / (SR - Ss);
  double rhoLs = rhoL * (SL - uLb) / (SL - Ss);
  double PtotR = GAMMA1 * (ER - 0.5 * rhoR * (u1R * u1R + u2R * u2R + u3R * u3R));
  double PtotL = GAMMA1 * (EL - 0.5 * rhoL * (u1L * u1L + u2L * u2L + u3L * u3L));
  double PtotRs = PtotR + rhoR * (SR - uRb) * (Ss - uRb) + Bsb * Bsb - BRb * BRb;
  double PtotLs = PtotL + rhoL * (SL - uLb) * (Ss - uLb) + Bsb * Bsb - BLb * BLb;
  if (dir == 1) {
    URs[1] = (UR1[1] * (SR - uRb) + PtotRs - PtotR + BRb * B1R - Bsb * B1s) / (SR - Ss);
    URs[2] = (UR1[2] * (SR - uRb) + BRb * B2R - Bsb * B2s) / (SR - Ss);
    ULs[1] = (UL1[1] * (SL - uLb) + PtotLs - PtotL + BLb * B1L - Bsb * B1s) / (SL - Ss);
    ULs[2] = (UL1[2] * (SL - uLb) + BLb * B2L - Bsb * B2s) / (SL - Ss);
  } else {
    URs[1] = (UR1[1] * (SR - uRb) + BRb * B1R - Bsb * B1s) / (SR - Ss);
    URs[2] = (UR1[2] * (SR - uRb) + PtotRs - PtotR + BRb * B2R - Bsb * B2s) / (SR - Ss);
    ULs[1] = (UL1[1] * (SL - uLb) + BLb * B1L - Bsb * B1s) / (SL - Ss);
    ULs[2] = (UL1[2] * (SL - uLb) + PtotLs - PtotL + BLb * B2L - Bsb * B2s) / (SL - Ss);
  }
  URs[3] = (UR1[3] * (SR - uRb) + BRb * B3R - Bsb * B3s) / (SR - Ss);
  ULs[3] = (UL1[3] * (SL - uLb) + BLb * B3L - Bsb * B3s) / (SL - Ss);
  double u1Rs = URs[1] / rhoRs, u2Rs = URs[2] / rhoRs, u3Rs = URs[3] / rhoRs;
  double u1Ls = ULs[1] / rhoLs, u2Ls = ULs[2] / rhoLs, u3Ls = ULs[3] / rhoLs;
  double ERs = ((SR - uRb) * ER - PtotR * uRb + PtotRs * Ss + BRb * (u1R * B1R + u2R * B2R + u3R * B3R) -
                Bsb * (u1Rs * B1s + u2Rs * B2s + u3Rs * B3s)) / (SR - Ss);
  double ELs = ((SL - uLb) * EL - PtotL * uLb + PtotLs * Ss + BLb * (u1L * B1L + u2L * B2L + u3L * B3L) -
                Bsb * (u1Ls * B1s + u2Ls * B2s + u3Ls * B3s)) / (SL - Ss);
  URs[0] = rhoRs; URs[4] = ERs; URs[5] = Ustar[5]; URs[6] = Ustar[6]; URs[7] = Ustar[7];
  ULs[0] = rhoLs; ULs[4] = ELs; ULs[5] = Ustar[5]; ULs[6] = Ustar[6]; ULs[7] = Ustar[7];
  if (Ss > 0) {
    for (int n = 0; n < 8; n++) { F[n] = FL1[n] + SL * (ULs[n] - UL1[n]); Ustar[n] = ULs[n]; }
  } else {
    for (int n = 0; n < 8; n++) { F[n] = FR1[n] + SR * (URs[n] - UR1[n]); Ustar[n] = URs[n]; }
  }
}

/* HLLD_Flux.f90:14-189, with its quirks: PtotR/L = gamma1*(E - rho|u|^2/2)
   (:64-65), +1e-15 in DR/DL (:76-77), double-star states only when
   |B*n| > 1e-4 and SR* - SL* > 1e-4 (:126). */
static void hlld_flux(int dir, double SR, double SL, const double UL1[8], const double UR1[8],
                      const double FL1in[8], const double FR1in[8], double F[8], double Ustar[8]) {
  if (SR < 0) {
    for (int n = 0; n < 8; n++) { F[n] = FR1in[n]; Ustar[n] = UR1[n]; }
    return;
  }
  if (SL > 0) {
    for (int n = 0; n < 8; n++) { F[n] = FL1in[n]; Ustar[n] = UL1[n]; }
    return;
  }
  double FR1[8], FL1[8], URs[8], ULs[8];
  for (int n = 0; n < 8; n++) { FR1[n] = FR1in[n]; FL1[n] = FL1in[n]; }
  for (int n = 0; n < 8; n++) Ustar[n] = (SR * UR1[n] - SL * UL1[n] + FL1[n] - FR1[n]) / (SR - SL);
  double rhoR = UR1[0], u1R = UR1[1] / rhoR, u2R = UR1[2] / rhoR, u3R = UR1[3] / rhoR, ER = UR1[4];
  double B1R = UR1[5], B2R = UR1[6], B3R = UR1[7];
  double rhoL = UL1[0], u1L = UL1[1] / rhoL, u2L = UL1[2] / rhoL, u3L = UL1[3] / rhoL, EL = UL1[4];
  double B1L = UL1[5], B2L = UL1[6], B3L = UL1[7];
  double Ss, uRb, uLb, BRb, BLb, Bsb;
  if (dir == 1) { Ss = Ustar[1] / Ustar[0]; uRb = u1R; uLb = u1L; BRb = B1R; BLb = B1L; Bsb = Ustar[5]; }
  else { Ss = Ustar[2] / Ustar[0]; uRb = u2R; uLb = u2L; BRb = B2R; BLb = B2L; Bsb = Ustar[6]; }
  double rhoRs = rhoR * (SR - uRb) / (SR - Ss);
  double rhoLs = rhoL * (SL - uLb) / (SL - Ss);
  double SRs = Ss + sqrt(Bsb * Bsb / rhoRs);
  double SLs = Ss - sqrt(Bsb * Bsb / rhoLs);
  double PtotR = GAMMA1 * (ER - 0.5 * rhoR * (u1R * u1R + u2R * u2R + u3R * u3R));
  double PtotL = GAMMA1 * (EL - 0.5 * rhoL * (u1L * u1L + u2L * u2L + u3L * u3L));
  double PtotRs = PtotR + rhoR * (SR - uRb) * (Ss - uRb) + Bsb * Bsb - BRb * BRb;
  double PtotLs = PtotL + rhoL * (SL - uLb) * (Ss - uLb) + Bsb * Bsb - BLb * BLb;
  double AR = rhoR * ((SR - uRb) * (SR - uRb)) - Bsb * BRb;
  double AL = rhoL * ((SL - uLb) * (SL - uLb)) - Bsb * BLb;
  double CR = rhoR * (SR - uRb) * (BRb - Bsb);
  double CL = rhoL * (SL - uLb) * (BLb - Bsb);
  double DR = rhoR * (SR - uRb) * (SR - Ss) - Bsb * Bsb + 1e-15;
  double DL = rhoL * (SL - uLb) * (SL - Ss) - Bsb * Bsb + 1e-15;
  double B1Rs, B1Ls, B2Rs, B2Ls;
  if (dir == 1) {
    B1Rs = Bsb; B1Ls = Bsb;
    B2Rs = (AR * B2R + CR * u2R) / DR;
    B2Ls = (AL * B2L + CL * u2L) / DL;
  } else {
    B1Rs = (AR * B1R + CR * u1R) / DR;
    B1Ls = (AL * B1L + CL * u1L) / DL;
    B2Rs = Bsb; B2Ls = Bsb;
  }
  double B3Rs = (AR * B3R + CR * u3R) / DR;
  double B3Ls = (AL * B3L + CL * u3L) / DL;
  if (dir == 1) {
    URs[1] = (UR1[1] * (SR - uRb) + PtotRs - PtotR + BRb * B1R - Bsb * B1Rs) / (SR - Ss);
    URs[2] = (UR1[2] * (SR - uRb) + BRb * B2R - Bsb * B2Rs) / (SR - Ss);
    ULs[1] = (UL1[1] * (SL - uLb) + PtotLs - PtotL + BLb * B1L - Bsb * B1Ls) / (SL - Ss);
    ULs[2] = (UL1[2] * (SL - uLb) + BLb * B2L - Bsb * B2Ls) / (SL - Ss);
  } else {
    URs[1] = (UR1[1] * (SR - uRb) + BRb * B1R - Bsb * B1Rs) / (SR - Ss);
    URs[2] = (UR1[2] * (SR - uRb) + PtotRs - PtotR + BRb * B2R - Bsb * B2Rs) / (SR - Ss);
    ULs[1] = (UL1[1] * (SL - uLb) + BLb * B1L - Bsb * B1Ls) / (SL - Ss);
    ULs[2] = (UL1[2] * (SL - uLb) + PtotLs - PtotL + BLb * B2L - Bsb * B2Ls) / (SL - Ss);
  }
  URs[3] = (UR1[3] * (SR - uRb) + BRb * B3R - Bsb * B3Rs) / (SR - Ss);
  ULs[3] = (UL1[3] * (SL - uLb) + BLb * B3L - Bsb * B3Ls) / (SL - Ss);
  double u1Rs = URs[1] / rhoRs, u2Rs = URs[2] / rhoRs, u3Rs = URs[3] / rhoRs;
  double u1Ls = ULs[1] / rhoLs, u2Ls = ULs[2] / rhoLs, u3Ls = ULs[3] / rhoLs;
  double ERs = ((SR - uRb) * ER - PtotR * uRb + PtotRs * Ss + BRb * (u1R * B1R + u2R * B2R + u3R * B3R) -
                Bsb * (u1Rs * B1Rs + u2Rs * B2Rs + u3Rs * B3Rs)) / (SR - Ss);
  double ELs = ((SL - uLb) * EL - PtotL * uLb + PtotLs * Ss + BLb * (u1L * B1L + u2L * B2L + u3L * B3L) -
                Bsb * (u1Ls * B1Ls + u2Ls * B2Ls + u3Ls * B3Ls)) / (SL - Ss);
  URs[0] = rhoRs; URs[4] = ERs; URs[5] = Ustar[5]; URs[6] = Ustar[6]; URs[7] = Ustar[7];
  ULs[0] = rhoLs; ULs[4] = ELs; ULs[5] = Ustar[5]; ULs[6] = Ustar[6]; ULs[7] = Ustar[7];
  if (fabs(Bsb) > 1e-4 && SRs - SLs > 1e-4) {
    if (SLs > 0) {
      for (int n = 0; n < 8; n++) { F[n] = FL1[n] + SL * (ULs[n] - UL1[n]); Ustar[n] = ULs[n]; }
    } else if (SRs < 0) {
      for (int n = 0; n < 8; n++) { F[n] = FR1[n] + SR * (URs[n] - UR1[n]); Ustar[n] = URs[n]; }
    } else {
      for (int n = 0; n < 8; n++) {
        FR1[n] = FR1[n] + SR * (URs[n] - UR1[n]);
        FL1[n] = FL1[n] + SL * (ULs[n] - UL1[n]);
      }
      double sgn = fsign1(Bsb);
      double sL = sqrt(fabs(rhoLs)), sR = sqrt(fabs(rhoRs)), sLR = sqrt(fabs(rhoLs * rhoRs));
      double u1ss = (sL * u1Ls + sR * u1Rs + sgn * (B1Rs - B1Ls)) / (sL + sR);
      double u2ss = (sL * u2Ls + sR * u2Rs + sgn * (B2Rs - B2Ls)) / (sL + sR);
      double u3ss = (sL * u3Ls + sR * u3Rs + sgn * (B3Rs - B3Ls)) / (sL + sR);
      double B1ss = (sL * B1Ls + sR * B1Rs + sgn * sLR * (u1Rs - u1Ls)) / (sL + sR);
      double B2ss = (sL * B2Ls + sR * B2Rs + sgn * sLR * (u2Rs - u2Ls)) / (sL + sR);
      double B3ss = (sL * B3Ls + sR * B3Rs + sgn * sLR * (u3Rs - u3Ls)) / (sL + sR);
      double dot = u1ss * B1ss;
      double ERss = ERs + sR * (u1Rs * B1Rs + u2Rs * B2Rs + u3Rs * B3Rs - u1ss * B1ss - u2ss * B2ss - u3ss * B3ss) * sgn;
      double ELss = ELs - sL * (u1Ls * B1Ls + u2Ls * B2Ls + u3Ls * B3Ls - u1ss * B1ss - u2ss * B2ss - u3ss * B3ss) * sgn;
      (void)dot;
      double URss[8] = {rhoRs, rhoRs * u1ss, rhoRs * u2ss, rhoRs * u3ss, ERss, B1ss, B2ss, B3ss};
      double ULss[8] = {rhoLs, rhoLs * u1ss, rhoLs * u2ss, rhoLs * u3ss, ELss, B1ss, B2ss, B3ss};
      if (Ss > 0) {
        for (int n = 0; n < 8; n++) { F[n] = FL1[n] + SLs * (ULss[n] - ULs[n]); Ustar[n] = ULss[n]; }
      } else {
        for (int n = 0; n < 8; n++) { F[n] = FR1[n] + SRs * (URss[n] - URs[n]); Ustar[n] = URss[n]; }
      }
    }
  } else {
    if (Ss > 0) {
      for (int n = 0; n < 8; n++) { F[n] = FL1[n] + SL * (ULs[n] - UL1[n]); Ustar[n] = ULs[n]; }
    } else {
      for (int n = 0; n < 8; n++) { F[n] = FR1[n] + SR * (URs[n] - UR1[n]); Ustar[n] = URs[n]; }
    }
  }
}

/* flux_type dispatch of Lh.f90:220-230 (UL = left/lower state) */
void orc_riemann_1d(int flux_type, int direction, double SR, double SL,
                    const double UL[8], const double UR[8],
                    const double FL[8], const double FR[8],
                    double Fhat[8], double Ustar[8]) {
  double tmp[8];
  if (!Ustar) Ustar = tmp;
  switch (flux_type) {
    case ORC_LF: lf_flux(SR, SL, UL, UR, FL, FR, Fhat); break;
    case ORC_HLL: hll_flux(SR, SL, UL, UR, FL, FR, Fhat); break;
    case ORC_HLLC: hllc_flux(direction, SR, SL, UL, UR, FL, FR, Fhat, Ustar); break;
    default: hlld_flux(direction, SR, SL, UL, UR, FL, FR, Fhat, Ustar); break;
  }
}

/* corner (vertex) EMF: LF_Flux_2D.f90:13-60, HLL_Flux_2D.f90:17-109,
   HLLC_Flux_2D.f90:17-114, HLLD_Flux_2D.f90:17-114.  Arguments are the
   states in the four quadrants around the vertex. */
double orc_vertex_emf(int flux_type, const double URU[8], const double ULU[8],
                      const double URD[8], const double ULD[8]) {
  if (flux_type == ORC_LF) {
    double ax = fmax2(fmax2(fmax2(fabs(URU[1] / URU[0]), fabs(ULU[1] / ULU[0])), fabs(URD[1] / URD[0])), fabs(ULD[1] / ULD[0]));
    double ay = fmax2(fmax2(fmax2(fabs(URU[2] / URU[0]), fabs(ULU[2] / ULU[0])), fabs(URD[2] / URD[0])), fabs(ULD[2] / ULD[0]));
    double EzRU = (URU[2] / URU[0]) * URU[5] - (URU[1] / URU[0]) * URU[6];
    double EzLU = (ULU[2] / ULU[0]) * ULU[5] - (ULU[1] / ULU[0]) * ULU[6];
    double EzRD = (URD[2] / URD[0]) * URD[5] - (URD[1] / URD[0]) * URD[6];
    double EzLD = (ULD[2] / ULD[0]) * ULD[5] - (ULD[1] / ULD[0]) * ULD[6];
    return 0.25 * (EzRU + EzLU + EzRD + EzLD) - 0.25 * ay * (0.5 * (URU[5] + ULU[5]) - 0.5 * (URD[5] + ULD[5])) +
           0.25 * ax * (0.5 * (URU[6] + URD[6]) - 0.5 * (ULU[6] + ULD[6]));
  }
  double aRU, bRU, aLU, bLU, aRD, bRD, aLD, bLD;
  double cRU, dRU, cLU, dLU, cRD, dRD, cLD, dLD;
  orc_eigenvalueMm(URU, 1, 0, &aRU, &bRU);
  orc_eigenvalueMm(URU, 0, 1, &cRU, &dRU);
  orc_eigenvalueMm(ULU, 1, 0, &aLU, &bLU);
  orc_eigenvalueMm(ULU, 0, 1, &cLU, &dLU);
  orc_eigenvalueMm(URD, 1, 0, &aRD, &bRD);
  orc_eigenvalueMm(URD, 0, 1, &cRD, &dRD);
  orc_eigenvalueMm(ULD, 1, 0, &aLD, &bLD);
  orc_eigenvalueMm(ULD, 0, 1, &cLD, &dLD);
  double SR = fmax2(fmax2(fmax2(aRU, aLU), aRD), aLD);
  double SL = fmin2(fmin2(fmin2(bRU, bLU), bRD), bLD);
  double SU = fmax2(fmax2(fmax2(cRU, cLU), cRD), cLD);
  double SD = fmin2(fmin2(fmin2(dRU, dLU), dRD), dLD);
  double FxRU[8], FyRU[8], FxLU[8], FyLU[8], FxRD[8], FyRD[8], FxLD[8], FyLD[8];
  orc_mhd_flux(URU, FxRU, FyRU);
  orc_mhd_flux(ULU, FxLU, FyLU);
  orc_mhd_flux(URD, FxRD, FyRD);
  orc_mhd_flux(ULD, FxLD, FyLD);
  double FU[8], FD[8], FRr[8], FLl[8], UUs[8], UDs[8], URs[8], ULs[8];
  if (flux_type == ORC_HLL) {
    /* HLL1.f90:5-14 */
    double *Us[4] = {UUs, UDs, URs, ULs}, *Fs[4] = {FU, FD, FRr, FLl};
    const double *Ua[4] = {URU, URD, URU, ULU}, *Ub[4] = {ULU, ULD, URD, ULD};
    const double *Fa[4] = {FxRU, FxRD, FyRU, FyLU}, *Fb[4] = {FxLU, FxLD, FyRD, FyLD};
    double Sa[4] = {SR, SR, SU, SU}, Sb[4] = {SL, SL, SD, SD};
    for (int c = 0; c < 4; c++) {
      double sr = Sa[c], sl = Sb[c];
      for (int n = 0; n < 8; n++) {
        if (sr < 0) { Fs[c][n] = Fa[c][n]; Us[c][n] = Ua[c][n]; }
        else if (sl > 0) { Fs[c][n] = Fb[c][n]; Us[c][n] = Ub[c][n]; }
        else {
          Fs[c][n] = (sr * Fb[c][n] - sl * Fa[c][n] + sl * sr * (Ua[c][n] - Ub[c][n])) / (sr - sl);
          Us[c][n] = (sr * Ua[c][n] - sl * Ub[c][n] - (Fa[c][n] - Fb[c][n])) / (sr - sl);
        }
      }
    }
    double BxRU = URU[5], BxLU = ULU[5], BxRD = URD[5], BxLD = ULD[5];
    double ByRU = URU[6], ByLU = ULU[6], ByRD = URD[6], ByLD = ULD[6];
    /* corner fields for the uninitialised EzRU..EzLD (calculate_Ez.f90:5) */
    double EzRU = (URU[2] / URU[0]) * URU[5] - (URU[1] / URU[0]) * URU[6];
    double EzLU = (ULU[2] / ULU[0]) * ULU[5] - (ULU[1] / ULU[0]) * ULU[6];
    double EzRD = (URD[2] / URD[0]) * URD[5] - (URD[1] / URD[0]) * URD[6];
    double EzLD = (ULD[2] / ULD[0]) * ULD[5] - (ULD[1] / ULD[0]) * ULD[6];
    double BxUs = UUs[5], BxDs = UDs[5], ByRs = URs[6], ByLs = ULs[6];
    double EzRs = FRr[5], EzLs = FLl[5], EzUs = -FU[6], EzDs = -FD[6];
    double BxSS = (2 * SR * SU * BxRU - 2 * SL * SU * BxLU + 2 * SL * SD * BxLD - 2 * SR * SD * BxRD -
                   SR * (EzRU - EzRD) + SL * (EzLU - EzLD) - (SR - SL) * (EzUs - EzDs)) / (2 * (SR - SL) * (SU - SD));
    double BySS = (2 * SR * SU * ByRU - 2 * SL * SU * ByLU + 2 * SL * SD * ByLD - 2 * SR * SD * ByRD +
                   SU * (EzRU - EzLU) - SD * (EzRD - EzLD) + (SU - SD) * (EzRs - EzLs)) / (2 * (SR - SL) * (SU - SD));
    double EzSS = 0.25 * (EzRs + EzLs + EzUs + EzDs) - 0.25 * SU * (BxUs - BxSS) - 0.25 * SD * (BxDs - BxSS) +
                  0.25 * SR * (ByRs - BySS) + 0.25 * SL * (ByLs - BySS);
    if (SL > 0) return EzLs;
    if (SR < 0) return EzRs;
    if (SD > 0) return EzDs;
    if (SU < 0) return EzUs;
    return EzSS;
  }
  /* HLLC / HLLD: four 1D solves with the 2D bounds (HLLD_Flux_2D.f90:46-88) */
  orc_riemann_1d(flux_type, 1, SR, SL, ULU, URU, FxLU, FxRU, FU, UUs);
  orc_riemann_1d(flux_type, 1, SR, SL, ULD, URD, FxLD, FxRD, FD, UDs);
  orc_riemann_1d(flux_type, 2, SU, SD, URD, URU, FyRD, FyRU, FRr, URs);
  orc_riemann_1d(flux_type, 2, SU, SD, ULD, ULU, FyLD, FyLU, FLl, ULs);
  double EzRs = FRr[5], EzLs = FLl[5], EzUs = -FU[6], EzDs = -FD[6];
  double EzSS = 0.25 * (EzRs + EzLs + EzUs + EzDs);
  if (SL > 0) return EzLs;
  if (SR < 0) return EzRs;
  if (SD > 0) return EzDs;
  if (SU < 0) return EzUs;
  return EzSS;
}

/* eigenmatrix.f90:1-407 (+ compute_t_vector :411-433, pick_orthogonal
   :451-469).  R and L are column-major 8x8 (R[r + 8*c]). */
static void compute_t_vector(double n1, double n2, double n3, double B1, double B2, double B3,
                             double *t1, double *t2, double *t3) {
  double nB1 = n2 * B3 - n3 * B2, nB2 = n3 * B1 - n1 * B3, nB3 = n1 * B2 - n2 * B1;
  double tp1, tp2, tp3;
  if (nB1 * nB1 + nB2 * nB2 + nB3 * nB3 < 1e-7) {
    double v1 = n1, v2 = n2, v3 = n3;
    if (fabs(v1) >= fabs(v2) && fabs(v1) >= fabs(v3)) { tp2 = 1.0; tp3 = 1.0; tp1 = (-v2 - v3) / v1; }
    else if (fabs(v2) >= fabs(v1) && fabs(v2) >= fabs(v3)) { tp1 = 1.0; tp3 = 1.0; tp2 = (-v1 - v3) / v2; }
    else { tp1 = 1.0; tp2 = 1.0; tp3 = (-v1 - v2) / v3; }
  } else {
    tp1 = nB2 * n3 - nB3 * n2;
    tp2 = nB3 * n1 - nB1 * n3;
    tp3 = nB1 * n2 - nB2 * n1;
  }
  double tn = sqrt(tp1 * tp1 + tp2 * tp2 + tp3 * tp3);
  *t1 = tp1 / tn; *t2 = tp2 / tn; *t3 = tp3 / tn;
}

void orc_eigenmatrix(double evr[64], double evl[64], const double U[8], double n1, double n2, double n3) {
#define RU(r, c) ru[((r)-1) + 8 * ((c)-1)]
#define LUm(r, c) lu[((r)-1) + 8 * ((c)-1)]
#define EVR(r, c) evr[((r)-1) + 8 * ((c)-1)]
#define EVL(r, c) evl[((r)-1) + 8 * ((c)-1)]
  double rho = U[0], u1 = U[1] / rho, u2 = U[2] / rho, u3 = U[3] / rho, E = U[4];
  double B1 = U[5], B2 = U[6], B3 = U[7];
  const double gamma = 5.0 / 3.0;
  double ru[64], lu[64];
  double p = (gamma - 1.0) * (E - 0.5 * rho * (u1 * u1 + u2 * u2 + u3 * u3) - 0.5 * (B1 * B1 + B2 * B2 + B3 * B3));
  double t1, t2, t3;
  compute_t_vector(n1, n2, n3, B1, B2, B3, &t1, &t2, &t3);
  double rhosq = sqrt(fabs(rho));
  double g1 = gamma - 1.0;
  double a2 = fabs(gamma * p / rho);
  double a = sqrt(a2);
  double sqg2 = sqrt(1.0 / (2.0 * gamma));
  double sq12 = sqrt(0.5);
  double sqpr = sqrt(fabs(p)) / rho;
  double sqpor = sqrt(fabs(p / rho));
  double sq1og = sqrt(1.0 / gamma);
  double b1s = B1 / rhosq, b2s = B2 / rhosq, b3s = B3 / rhosq;
  double bN = b1s * n1 + b2s * n2 + b3s * n3;
  double d = a2 + (b1s * b1s + b2s * b2s + b3s * b3s);
  double cf = sqrt(0.5 * fabs(d + sqrt(fabs(d * d - 4.0 * a2 * (bN * bN)))));
  double cs = sqrt(0.5 * fabs(d - sqrt(fabs(d * d - 4.0 * a2 * (bN * bN)))));
  double beta1 = copysign(1.0, (b1s * n1 + b2s * n2 + b3s * n3) * 1.0);
  double alphaf, alphas;
  if (fabs(cf * cf - cs * cs) <= 1e-12) {
    alphaf = sin(atan(1.0));
    alphas = cos(atan(1.0));
  } else {
    alphaf = sqrt(fabs(a2 - cs * cs)) / sqrt(fabs(cf * cf - cs * cs));
    alphas = sqrt(fabs(cf * cf - a2)) / sqrt(fabs(cf * cf - cs * cs));
  }
  double TxN1 = n3 * t2 - n2 * t3, TxN2 = n1 * t3 - n3 * t1, TxN3 = n2 * t1 - n1 * t2;
  for (int i = 0; i < 64; i++) ru[i] = 0.0;
  RU(1, 1) = sqrt(g1 / gamma) * rhosq;
  RU(6, 2) = sq1og * a * n1; RU(7, 2) = sq1og * a * n2; RU(8, 2) = sq1og * a * n3;
  RU(2, 3) = -sq12 * (sqpr * (TxN1)); RU(3, 3) = -sq12 * (sqpr * (TxN2)); RU(4, 3) = -sq12 * (sqpr * (TxN3));
  RU(6, 3) = sq12 * sqpor * (TxN1); RU(7, 3) = sq12 * sqpor * (TxN2); RU(8, 3) = sq12 * sqpor * (TxN3);
  RU(1, 4) = RU(1, 3); RU(2, 4) = -RU(2, 3); RU(3, 4) = -RU(3, 3); RU(4, 4) = -RU(4, 3);
  RU(5, 4) = RU(5, 3); RU(6, 4) = RU(6, 3); RU(7, 4) = RU(7, 3); RU(8, 4) = RU(8, 3);
  double bst = (b1s * t1 + b2s * t2 + b3s * t3);
  RU(1, 5) = sqg2 * alphaf * rhosq;
  RU(2, 5) = sqg2 * ((alphaf * a2 * n1 + alphas * a * ((bst)*n1 - (bN)*t1))) / (rhosq * cf);
  RU(3, 5) = sqg2 * ((alphaf * a2 * n2 + alphas * a * ((bst)*n2 - (bN)*t2))) / (rhosq * cf);
  RU(4, 5) = sqg2 * ((alphaf * a2 * n3 + alphas * a * ((bst)*n3 - (bN)*t3))) / (rhosq * cf);
  RU(5, 5) = sqg2 * alphaf * rhosq * a2;
  RU(6, 5) = sqg2 * alphas * a * t1; RU(7, 5) = sqg2 * alphas * a * t2; RU(8, 5) = sqg2 * alphas * a * t3;
  RU(1, 6) = RU(1, 5); RU(2, 6) = -RU(2, 5); RU(3, 6) = -RU(3, 5); RU(4, 6) = -RU(4, 5);
  RU(5, 6) = RU(5, 5); RU(6, 6) = RU(6, 5); RU(7, 6) = RU(7, 5); RU(8, 6) = RU(8, 5);
  RU(1, 7) = sqg2 * alphas * rhosq;
  RU(2, 7) = beta1 * sqg2 * (alphaf * (cf * cf) * t1 + a * n1 * alphas * (bN)) / (rhosq * cf);
  RU(3, 7) = beta1 * sqg2 * (alphaf * (cf * cf) * t2 + alphas * a * (bN)*n2) / (rhosq * cf);
  RU(4, 7) = beta1 * sqg2 * (alphaf * (cf * cf) * t3 + alphas * a * (bN)*n3) / (rhosq * cf);
  RU(5, 7) = (a * a) * sqg2 * alphas * rhosq;
  RU(6, 7) = -sqg2 * alphaf * a * t1; RU(7, 7) = -sqg2 * alphaf * a * t2; RU(8, 7) = -sqg2 * alphaf * a * t3;
  RU(1, 8) = RU(1, 7); RU(2, 8) = -RU(2, 7); RU(3, 8) = -RU(3, 7); RU(4, 8) = -RU(4, 7);
  RU(5, 8) = RU(5, 7); RU(6, 8) = RU(6, 7); RU(7, 8) = RU(7, 7); RU(8, 8) = RU(8, 7);
  for (int m = 1; m <= 8; m++) {
    EVR(1, m) = RU(1, m) / g1;
    EVR(2, m) = (RU(1, m) * u1 + RU(2, m) * rho) / g1;
    EVR(3, m) = (RU(1, m) * u2 + RU(3, m) * rho) / g1;
    EVR(4, m) = (RU(1, m) * u3 + RU(4, m) * rho) / g1;
    EVR(5, m) = (RU(5, m) / g1 + B1 * RU(6, m) + B2 * RU(7, m) + B3 * RU(8, m) +
                 0.5 * RU(1, m) * (u1 * u1 + u2 * u2 + u3 * u3) + RU(2, m) * u1 * rho + RU(3, m) * u2 * rho +
                 RU(4, m) * u3 * rho) / g1;
    EVR(6, m) = RU(6, m) / g1;
    EVR(7, m) = RU(7, m) / g1;
    EVR(8, m) = RU(8, m) / g1;
  }
  /* left eigenvectors (eigenmatrix.f90:185-394) */
  compute_t_vector(n1, n2, n3, B1, B2, B3, &t1, &t2, &t3);
  rhosq = sqrt(rho);
  a2 = fabs(gamma * p / rho);
  a = sqrt(a2);
  b1s = B1 / rhosq; b2s = B2 / rhosq; b3s = B3 / rhosq;
  double BNs = b1s * n1 + b2s * n2 + b3s * n3;
  double BN = (B1 * n1 + B2 * n2 + B3 * n3);
  d = a2 + (b1s * b1s + b2s * b2s + b3s * b3s);
  cf = sqrt(0.5 * fabs(d + sqrt(fabs(d * d - 4.0 * a2 * (BNs * BNs)))));
  cs = sqrt(0.5 * fabs(d - sqrt(fabs(d * d - 4.0 * a2 * (BNs * BNs)))));
  if (fabs(cf * cf - cs * cs) <= 1e-12) {
    alphaf = sin(atan(1.0));
    alphas = cos(atan(1.0));
  } else {
    alphaf = sqrt(fabs(a2 - cs * cs)) / sqrt(fabs(cf * cf - cs * cs));
    alphas = sqrt(fabs(cf * cf - a2)) / sqrt(fabs(cf * cf - cs * cs));
  }
  double twosq = sqrt(2.0);
  beta1 = copysign(1.0, (BNs) * 1.0);
  double sqgam = sqrt(g1 / gamma);
  TxN1 = n3 * t2 - n2 * t3; TxN2 = n1 * t3 - n3 * t1; TxN3 = n2 * t1 - n1 * t2;
  double BT = (B1 * t1 + B2 * t2 + B3 * t3);
  double gu1sq = sqrt(1.0 / gamma);
  for (int i = 0; i < 64; i++) lu[i] = 0.0;
  LUm(1, 1) = 1.0 / (sqgam * rhosq);
  LUm(1, 5) = -1.0 / (a2 * sqgam * rhosq);
  double nen = (n3 * n3 * (t1 * t1 + t2 * t2) - 2.0 * n1 * n3 * t1 * t3 - 2.0 * n2 * t2 * (n1 * t1 + n3 * t3) +
                n2 * n2 * (t1 * t1 + t3 * t3) + n1 * n1 * (t2 * t2 + t3 * t3));
  double nen2 = a * gu1sq * nen;
  LUm(2, 6) = (-n2 * t1 * t2 - n3 * t1 * t3 + n1 * (t2 * t2 + t3 * t3)) / nen2;
  LUm(2, 7) = (-t2 * (n1 * t1 + n3 * t3) + n2 * (t1 * t1 + t3 * t3)) / nen2;
  LUm(2, 8) = (n3 * (t1 * t1 + t2 * t2) - (n1 * t1 + n2 * t2) * t3) / nen2;
  double nen3 = sqrt(2.0) * sqrt(fabs(p)) * nen;
  double nen31 = nen3 / sqrt(fabs(rho));
  LUm(3, 2) = rho * (-TxN1) / nen3; LUm(3, 3) = rho * (-TxN2) / nen3; LUm(3, 4) = rho * (-TxN3) / nen3;
  LUm(3, 6) = TxN1 / nen31; LUm(3, 7) = TxN2 / nen31; LUm(3, 8) = TxN3 / nen31;
  LUm(4, 1) = LUm(3, 1); LUm(4, 2) = -LUm(3, 2); LUm(4, 3) = -LUm(3, 3); LUm(4, 4) = -LUm(3, 4);
  LUm(4, 5) = LUm(3, 5); LUm(4, 6) = LUm(3, 6); LUm(4, 7) = LUm(3, 7); LUm(4, 8) = LUm(3, 8);
  double cf2 = cf * cf;
  double Term51 = rho * cf * ((n2 * t1 * t2 + n3 * t1 * t3 - n1 * (t2 * t2 + t3 * t3)) * rhosq * cf2 * alphaf -
                              a * BN * alphas * (-(n2 * n2) * t1 + n1 * n2 * t2 + n3 * TxN2));
  double Term52 = rho * cf * ((-t2 * (n1 * t1 + n3 * t3) + n2 * (t1 * t1 + t3 * t3)) * rhosq * cf2 * alphaf -
                              a * BN * alphas * (-n1 * n2 * t1 + (n1 * n1) * t2 + n3 * TxN1));
  double Term53 = rho * cf * ((n3 * (t1 * t1 + t2 * t2) - (n1 * t1 + n2 * t2) * t3) * rhosq * cf2 * alphaf -
                              a * (-n1 * n3 * t1 + (n1 * n1) * t3 + n2 * (-n3 * t2 + n2 * t3)) * BN * alphas);
  double Term54 = alphaf / (twosq * (a * a) * gu1sq * rhosq * (alphaf * alphaf + alphas * alphas));
  double Term55 = ((n2 * n2) * t1 - n1 * n2 * t2 + n3 * (n3 * t1 - n1 * t3)) * alphas;
  double Term56 = alphas * (-n1 * n2 * t1 + (n1 * n1) * t2 + n3 * TxN1);
  double Term57 = (-n1 * n3 * t1 + (n1 * n1) * t3 + n2 * (-n3 * t2 + n2 * t3)) * alphas;
  double nen51 = twosq * a * nen * gu1sq * (a * (BN * BN) * (alphas * alphas) + rhosq * cf2 * alphaf * (a * rhosq * alphaf + BT * alphas));
  double nen52 = twosq * a * gu1sq * (alphaf * alphaf + alphas * alphas) * nen;
  LUm(5, 2) = -Term51 / nen51; LUm(5, 3) = Term52 / nen51; LUm(5, 4) = Term53 / nen51;
  LUm(5, 5) = Term54;
  LUm(5, 6) = Term55 / nen52; LUm(5, 7) = Term56 / nen52; LUm(5, 8) = Term57 / nen52;
  LUm(6, 1) = LUm(5, 1); LUm(6, 2) = -LUm(5, 2); LUm(6, 3) = -LUm(5, 3); LUm(6, 4) = -LUm(5, 4);
  LUm(6, 5) = LUm(5, 5); LUm(6, 6) = LUm(5, 6); LUm(6, 7) = LUm(5, 7); LUm(6, 8) = LUm(5, 8);
  double Term71 = rho * cf * (a * ((n2 * n2) * t1 - n1 * n2 * t2 + n3 * (n3 * t1 - n1 * t3)) * rhosq * alphaf +
                              ((B3 * n2 * t1 - B2 * n3 * t1 - B3 * n1 * t2 + B2 * n1 * t3) * (-n3 * t2 + n2 * t3) +
                               B1 * ((n2 * n2) * (t1 * t1) + (n3 * n3) * (t1 * t1) - 2.0 * n1 * n2 * t1 * t2 -
                                     2.0 * n1 * n3 * t1 * t3 + (n1 * n1) * (t2 * t2 + t3 * t3))) * alphas);
  double Term72 = rho * cf * (((n3 * t1 - n1 * t3) * (B3 * n2 * t1 - B3 * n1 * t2 + B1 * n3 * t2 - B1 * n2 * t3) +
                               B2 * (((n1 * n1) + (n3 * n3)) * (t2 * t2) - 2.0 * n2 * t2 * (n1 * t1 + n3 * t3) +
                                     (n2 * n2) * (t1 * t1 + t3 * t3))) * alphas +
                              a * rhosq * alphaf * (-n1 * n2 * t1 + (n1 * n1) * t2 + n3 * TxN1));
  double Term73 = rho * cf * (a * (-n1 * n3 * t1 + (n1 * n1) * t3 + n2 * (-n3 * t2 + n2 * t3)) * rhosq * alphaf +
                              alphas * (B3 * ((n3 * n3) * (t1 * t1 + t2 * t2) - 2.0 * n3 * (n1 * t1 + n2 * t2) * t3 +
                                              ((n1 * n1) + (n2 * n2)) * (t3 * t3)) +
                                        B2 * (n3 * t1 - n1 * t3) * TxN3 + B1 * (-n3 * t2 + n2 * t3) * TxN3));
  double Term74 = alphas / (twosq * (a * a) * gu1sq * rhosq * (alphaf * alphaf + alphas * alphas));
  double Term75 = -alphaf * ((n2 * n2) * t1 - n1 * n2 * t2 + n3 * (n3 * t1 - n1 * t3));
  double Term76 = -alphaf * (-n1 * n2 * t1 + (n1 * n1) * t2 + n3 * TxN1);
  double Term77 = -alphaf * (-n1 * n3 * t1 + (n1 * n1) * t3 + n2 * (-n3 * t2 + n2 * t3));
  double nen71 = twosq * beta1 * nen * gu1sq *
                 (a * (BN * BN) * (alphas * alphas) + rhosq * cf2 * alphaf * (a * rhosq * alphaf + BT * alphas));
  double nen72 = nen52;
  LUm(7, 2) = Term71 / nen71; LUm(7, 3) = Term72 / nen71; LUm(7, 4) = Term73 / nen71;
  LUm(7, 5) = Term74;
  LUm(7, 6) = Term75 / nen72; LUm(7, 7) = Term76 / nen72; LUm(7, 8) = Term77 / nen72;
  LUm(8, 1) = LUm(7, 1); LUm(8, 2) = -LUm(7, 2); LUm(8, 3) = -LUm(7, 3); LUm(8, 4) = -LUm(7, 4);
  LUm(8, 5) = LUm(7, 5); LUm(8, 6) = LUm(7, 6); LUm(8, 7) = LUm(7, 7); LUm(8, 8) = LUm(7, 8);
  for (int m = 1; m <= 8; m++) {
    EVL(m, 1) = LUm(m, 1) * g1 - LUm(m, 2) * u1 * g1 / rho - LUm(m, 3) * u2 * g1 / rho - LUm(m, 4) * u3 * g1 / rho +
                LUm(m, 5) * (g1 * g1) * (u1 * u1 + u2 * u2 + u3 * u3) * .5;
    EVL(m, 2) = -LUm(m, 5) * u1 * (g1 * g1) + LUm(m, 2) * g1 / rho;
    EVL(m, 3) = -LUm(m, 5) * u2 * (g1 * g1) + LUm(m, 3) * g1 / rho;
    EVL(m, 4) = -LUm(m, 5) * u3 * (g1 * g1) + LUm(m, 4) * g1 / rho;
    EVL(m, 5) = LUm(m, 5) * (g1 * g1);
    EVL(m, 6) = LUm(m, 6) * g1 - B1 * LUm(m, 5) * (g1 * g1);
    EVL(m, 7) = LUm(m, 7) * g1 - B2 * LUm(m, 5) * (g1 * g1);
    EVL(m, 8) = LUm(m, 8) * g1 - B3 * LUm(m, 5) * (g1 * g1);
  }
  /* identity fallback (eigenmatrix.f90:397-404) */
  double sr = 0.0, sl = 0.0;
  for (int i = 0; i < 64; i++) sr += fabs(evr[i]);
  for (int i = 0; i < 64; i++) sl += fabs(evl[i]);
  if (sr > 1e4 || sl > 1e4) {
    for (int i = 0; i < 64; i++) { evr[i] = 0.0; evl[i] = 0.0; }
    for (int i = 0; i < 8; i++) { evr[i + 8 * i] = 1.0; evl[i + 8 * i] = 1.0; }
  }
#undef RU
#undef LUm
#undef EVR
#undef EVL
}

/* pressure.f90:6 */
static inline double pressure(const double U[8]) {
  return (GAMMA - 1) * (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0] -
                        0.5 * (U[5] * U[5] + U[6] * U[6] + U[7] * U[7]));
}

/* ---------------------------------------------------------------------- */
/* state management                                                       */

static double *xalloc(size_t n) {
  double *p = (double *)calloc(n, sizeof(double));
  return p;
}

orc_state *orc_create(const orc_params *p) {
  orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
  s->p = *p;
  if (s->p.CFL <= 0) s->p.CFL = 0.15;
  int Nx = p->Nx, Ny = p->Ny;
  s->hx = (p->xb - p->xa) / Nx;        /* init_data.f90:21-24 */
  s->hy = (p->yb - p->ya) / Ny;
  s->hx1 = 0.5 * s->hx;
  s->hy1 = 0.5 * s->hy;
  get_GLP(&s->T);
  get_basis(&s->T, s->hx1, s->hy1);
  s->nuh = (size_t)(Nx + 2) * (Ny + 2) * NP * NE;
  s->nbx = (size_t)(Nx + 1) * Ny * 3;
  s->nby = (size_t)Nx * (Ny + 1) * 3;
  s->uh = xalloc(s->nuh); s->du = xalloc(s->nuh); s->uI = xalloc(s->nuh);
  s->uII = xalloc(s->nuh); s->uh0 = xalloc(s->nuh); s->uhmod = xalloc(s->nuh);
  s->Bx = xalloc(s->nbx); s->dEz1 = xalloc(s->nbx); s->BxI = xalloc(s->nbx);
  s->BxII = xalloc(s->nbx); s->Bx0 = xalloc(s->nbx);
  s->By = xalloc(s->nby); s->dEz2 = xalloc(s->nby); s->ByI = xalloc(s->nby);
  s->ByII = xalloc(s->nby); s->By0 = xalloc(s->nby);
  s->Fxhat = xalloc((size_t)(Nx + 1) * Ny * NG * NE);
  s->Fyhat = xalloc((size_t)Nx * (Ny + 1) * NG * NE);
  s->EzRL = xalloc((size_t)(Nx + 1) * Ny * NG);
  s->EzUD = xalloc((size_t)Nx * (Ny + 1) * NG);
  s->EzVertex = xalloc((size_t)(Nx + 1) * (Ny + 1));
  s->trouble = (int32_t *)calloc((size_t)Nx * Ny, sizeof(int32_t));
  s->pmin = 1e6;
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
  return s;
}

void orc_destroy(orc_state *s) {
  if (!s) return;
  double *arr[] = {s->uh, s->du, s->uI, s->uII, s->uh0, s->uhmod, s->Bx, s->dEz1, s->BxI, s->BxII,
                   s->Bx0, s->By, s->dEz2, s->ByI, s->ByII, s->By0, s->Fxhat, s->Fyhat, s->EzRL,
                   s->EzUD, s->EzVertex};
  for (size_t i = 0; i < sizeof arr / sizeof arr[0]; i++) free(arr[i]);
  free(s->trouble);
  free(s);
}

double *orc_uh(orc_state *s) { return s->uh; }
double *orc_Bx(orc_state *s) { return s->Bx; }
double *orc_By(orc_state *s) { return s->By; }
double *orc_du(orc_state *s) { return s->du; }
double *orc_dEz1(orc_state *s) { return s->dEz1; }
double *orc_dEz2(orc_state *s) { return s->dEz2; }
double *orc_EzVertex(orc_state *s) { return s->EzVertex; }
int32_t *orc_trouble(orc_state *s) { return s->trouble; }
void orc_scalars(orc_state *s, double o[8]) {
  o[0] = s->t; o[1] = s->dt; o[2] = s->umax; o[3] = s->totaldiv;
  o[4] = s->pmin; o[5] = s->alphax; o[6] = s->alphay; o[7] = (double)s->count;
}
void orc_set_time(orc_state *s, double t, double dt) { s->t = t; s->dt = dt; }

/* ---------------------------------------------------------------------- */
/* initial conditions (init_data.f90:21-121 with init1/3/4.txt)           */

typedef struct {
  int ic;
  int full_energy;   /* init5-9: kinetic and magnetic energy include w and Bz */
  int shock;         /* init5, 6, 6b, 8b: shock = 1 (init_data.f90:92-121) */
  double par[32];
  double kx[16], ky[16], phu[16], phv[16];
} icdef_t;

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double u01(uint64_t *x) { return (double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0); }

/* conserved variables U1..U8 at (x, y) */
static void ic_eval(const icdef_t *c, double x, double y, double U[8]) {
  const double pi = 4 * atan(1.0);
  double rho = 1, v1 = 0, v2 = 0, v3 = 0, B1 = 0, B2 = 0, B3 = 0, p = 1;
  switch (c->ic) {
    case ORC_IC_ORSZAG_TANG: /* init3.txt:3-18 */
    case ORC_IC_OT_PERTURBED: {
      p = GAMMA;
      rho = GAMMA * GAMMA;
      v1 = -sin(y);
      v2 = sin(x);
      B1 = -sin(y);
      B2 = sin(2.0 * x);
      if (c->ic == ORC_IC_OT_PERTURBED) {
        double amp = c->par[0] / 16.0, du = 0, dv = 0;
        for (int m = 0; m < 16; m++) {
          du += sin(c->kx[m] * x + c->ky[m] * y + c->phu[m]);
          dv += sin(c->kx[m] * x + c->ky[m] * y + c->phv[m]);
        }
        v1 += amp * du;
        v2 += amp * dv;
      }
      break;
    }
    case ORC_IC_VORTEX: { /* init1.txt:3-19 */
      double r2 = x * x + y * y;
      p = 1 - r2 / (8.0 * (pi * pi)) * exp(1 - r2);
      rho = 1;
      v1 = 1 - 1.0 / (2.0 * pi) * y * exp(0.5 * (1 - r2));
      v2 = 1 + 1.0 / (2.0 * pi) * x * exp(0.5 * (1 - r2));
      B1 = -1.0 / (2.0 * pi) * y * exp(0.5 * (1 - r2));
      B2 = 1.0 / (2.0 * pi) * x * exp(0.5 * (1 - r2));
      break;
    }
    case ORC_IC_ROTOR: { /* init4.txt + Rotor_functions.f90:1-71; par: r0, r1, spin, p, B1, rho_in, xc, yc */
      double r0 = c->par[0], r1 = c->par[1], spin = c->par[2], xc = c->par[6], yc = c->par[7];
      double r = sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc));
      double f = (r1 - r) / (r1 - r0);
      double rin = c->par[5];
      if (r < r0) { rho = rin; v1 = -spin * (y - yc) / r0; v2 = spin * (x - xc) / r0; }
      else if (r < r1) { rho = 1.0 + (rin - 1.0) * f; v1 = -spin * f * (y - yc) / r; v2 = spin * f * (x - xc) / r; }
      else { rho = 1.0; v1 = 0.0; v2 = 0.0; }
      p = c->par[3];
      B1 = c->par[4];
      break;
    }
    case ORC_IC_BLAST: { /* init5.txt:3-18, p0.f90:1-14 */
      double r = pow(x * x + y * y, 0.5);
      p = r <= 0.1 ? 1000.0 : 0.1;
      B1 = 100.0 / pow(4 * pi, 0.5);
      break;
    }
    case ORC_IC_CLOUD_SHOCK: { /* init6.txt:3-18, Cloud_Shock_functions.f90:1-117 */
      double r = pow((x - 1.4) * (x - 1.4) + (y - 0.5) * (y - 0.5), 0.5);
      if (x <= 1.2) { rho = 3.88968; v3 = -0.05234; p = 14.2641; B3 = 3.9353; }
      else { rho = r >= 0.18 ? 1.0 : 5.0; v1 = -3.3156; p = 0.04; B3 = 1.0; }
      B1 = 1.0;
      break;
    }
    case ORC_IC_CLOUD_SHOCK_2: { /* init6b.txt:3-18, Cloud_Shock_functions2.f90:1-116 */
      double r = pow((x - 0.25) * (x - 0.25) + (y - 0.5) * (y - 0.5), 0.5);
      if (x <= 0.05) { rho = 3.88968; v1 = 11.2536; p = 167.345; B2 = 0.21826182; B3 = -0.21826182; }
      else { rho = r >= 0.15 ? 1.0 : 10.0; p = 1.0; B2 = 0.56418958; B3 = 0.56418958; }
      break;
    }
    case ORC_IC_MAGNETIC_LOOP: { /* init7.txt:3-18, Loop_functions.f90:1-37 */
      double r = pow(x * x + y * y, 0.5);
      v1 = 2.0;
      v2 = 1.0;
      if (r < 0.3) { B1 = -1e-3 * y / r; B2 = 1e-3 * x / r; }
      break;
    }
    case ORC_IC_BRIO_WU: /* init8.txt:3-18, BrioWu_functions.f90:1-37 */
      rho = x < 0 ? 1.0 : 0.1225; p = x < 0 ? 1.0 : 0.1; B1 = 0.75; B2 = x < 0 ? 1.0 : -1.0;
      break;
    case ORC_IC_BRIO_WU_Y: /* init8b.txt:3-18, BrioWu_functions.f90:40-76 */
      rho = y < 0 ? 1.0 : 0.1225; p = y < 0 ? 1.0 : 0.1; B1 = y < 0 ? 1.0 : -1.0; B2 = 0.75;
      break;
    case ORC_IC_RYU_JONES: /* init9.txt:3-18, RyuJones_functions.f90:1-49 */
      rho = x < 0 ? 1.08 : 1.0; p = x < 0 ? 0.95 : 1.0; v1 = x < 0 ? 1.2 : 0.0;
      B1 = 2.0 / pow(4 * pi, 0.5); B2 = (x < 0 ? 3.6 : 4.0) / pow(4 * pi, 0.5); B3 = 2.0 / pow(4 * pi, 0.5);
      break;
    case ORC_IC_RYU_JONES_Y: /* init9b.txt:3-18, RyuJones_functions.f90:52-100 */
      rho = y < 0 ? 1.08 : 1.0; p = y < 0 ? 0.95 : 1.0; v2 = y < 0 ? 1.2 : 0.0;
      B1 = (y < 0 ? 3.6 : 4.0) / pow(4 * pi, 0.5); B2 = 2.0 / pow(4 * pi, 0.5); B3 = 2.0 / pow(4 * pi, 0.5);
      break;
    case ORC_IC_RIEMANN4: { /* piecewise constant; par: x0, y0, then 4 x (rho,u,v,p,Bx,By) for NE, NW, SE, SW */
      int q = (x >= c->par[0] ? 0 : 1) + (y >= c->par[1] ? 0 : 2);
      const double *st = &c->par[2 + 6 * q];
      rho = st[0]; v1 = st[1]; v2 = st[2]; p = st[3]; B1 = st[4]; B2 = st[5];
      break;
    }
    default: break;
  }
  U[0] = rho;
  U[1] = rho * v1;
  U[2] = rho * v2;
  U[3] = rho * v3;
  if (c->full_energy)
    U[4] = p / GAMMA1 + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3) + 0.5 * (B1 * B1 + B2 * B2 + B3 * B3);
  else
    U[4] = p / GAMMA1 + 0.5 * rho * (v1 * v1 + v2 * v2) + 0.5 * (B1 * B1 + B2 * B2);
  U[5] = B1;
  U[6] = B2;
  U[7] = B3;
}

void orc_init_data(orc_state *s, int ic, const double *extra, int nextra) {
  icdef_t c;
  memset(&c, 0, sizeof c);
  c.ic = ic;
  c.full_energy = (ic >= 5 && ic <= 9) || (ic >= 61 && ic <= 91);
  c.shock = ic == ORC_IC_BLAST || ic == ORC_IC_CLOUD_SHOCK || ic == ORC_IC_CLOUD_SHOCK_2 || ic == ORC_IC_BRIO_WU_Y;
  if (ic == ORC_IC_OT_PERTURBED) {
    /* 16 modes, (kx, ky) = (1 + m mod 4, 1 + m div 4); phases from splitmix64 */
    c.par[0] = (nextra > 0) ? extra[0] : 1e-2;
    uint64_t seed = (nextra > 1) ? (uint64_t)extra[1] : 20261203ull;
    uint64_t st = seed;
    const double pi = 4 * atan(1.0);
    for (int m = 0; m < 16; m++) {
      c.kx[m] = 1 + (m % 4);
      c.ky[m] = 1 + (m / 4);
      c.phu[m] = 2 * pi * u01(&st);
      c.phv[m] = 2 * pi * u01(&st);
    }
  } else if (ic == ORC_IC_ROTOR) {
    double def[8] = {0.1, 0.115, 1.0, 0.5, 2.4 / sqrt(4 * (4 * atan(1.0))), 10.0, 0.5, 0.5};
    for (int i = 0; i < 8; i++) c.par[i] = (i < nextra) ? extra[i] : def[i];
  } else if (ic == ORC_IC_RIEMANN4) {
    for (int i = 0; i < 26 && i < nextra; i++) c.par[i] = extra[i];
  }
  int Nx = NXr, Ny = NYr;
  const tables_t *T = &s->T;
  double xa = s->p.xa, ya = s->p.ya, hx = s->hx, hy = s->hy, hx1 = s->hx1, hy1 = s->hy1;
  memset(s->uh, 0, s->nuh * sizeof(double));
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 1; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      double Xc = xa + (i - 0.5) * hx, Yc = ya + (j - 0.5) * hy;
      double Uq[NG][NG][8];
      for (int i1 = 0; i1 < NG; i1++)
        for (int j1 = 0; j1 < NG; j1++) ic_eval(&c, Xc + hx1 * T->lambda[i1], Yc + hy1 * T->lambda[j1], Uq[i1][j1]);
      for (int d = 1; d <= NP; d++) {
        if (d > 6) continue; /* init_data.f90:43 with n undefined: modes 7-10 stay 0 */
        for (int n = 1; n <= NE; n++) {
          double acc = 0.0;
          for (int i1 = 0; i1 < NG; i1++)
            for (int j1 = 0; j1 < NG; j1++)
              acc = acc + T->weight[i1] * T->weight[j1] * T->phiG[d - 1][i1][j1] * Uq[i1][j1][n - 1];
          s->uh[UHI(i, j, d, n)] = (0.25 * acc) / T->mm[d - 1];
        }
      }
    }
  memset(s->Bx, 0, s->nbx * sizeof(double));
  memset(s->By, 0, s->nby * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= Ny; j++)
    for (int i = 0; i <= Nx; i++) {
      double Yc = ya + (j - 0.5) * hy;
      double Uv[NG][8];
      for (int j1 = 0; j1 < NG; j1++) ic_eval(&c, xa + i * hx, Yc + hy1 * T->lambda[j1], Uv[j1]);
      for (int d = 1; d <= 3; d++) {
        double acc = 0.0;
        for (int j1 = 0; j1 < NG; j1++) acc = acc + 0.5 * T->weight[j1] * T->EzG[d - 1][j1] * Uv[j1][5];
        s->Bx[BXI(i, j, d)] = acc / T->mmE[d - 1];
      }
    }
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      double Xc = xa + (i - 0.5) * hx;
      double Uv[NG][8];
      for (int i1 = 0; i1 < NG; i1++) ic_eval(&c, Xc + hx1 * T->lambda[i1], ya + j * hy, Uv[i1]);
      for (int d = 1; d <= 3; d++) {
        double acc = 0.0;
        for (int i1 = 0; i1 < NG; i1++) acc = acc + 0.5 * T->weight[i1] * T->EzG[d - 1][i1] * Uv[i1][6];
        s->By[BYI(i, j, d)] = acc / T->mmE[d - 1];
      }
    }
  if (!c.shock) return;
  /* init_data.f90:92-121: P0 states (cell and face-centre values) */
  for (int j = 1; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      double U[8];
      ic_eval(&c, xa + (i - 0.5) * hx, ya + (j - 0.5) * hy, U);
      for (int n = 1; n <= NE; n++) {
        for (int d = 2; d <= NP; d++) s->uh[UHI(i, j, d, n)] = 0.0;
        s->uh[UHI(i, j, 1, n)] = U[n - 1];
      }
    }
  for (int j = 1; j <= Ny; j++)
    for (int i = 0; i <= Nx; i++) {
      double U[8];
      ic_eval(&c, xa + i * hx, ya + (j - 0.5) * hy, U);
      s->Bx[BXI(i, j, 1)] = U[5];
      s->Bx[BXI(i, j, 2)] = s->Bx[BXI(i, j, 3)] = 0.0;
    }
  for (int j = 0; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      double U[8];
      ic_eval(&c, xa + (i - 0.5) * hx, ya + j * hy, U);
      s->By[BYI(i, j, 1)] = U[6];
      s->By[BYI(i, j, 2)] = s->By[BYI(i, j, 3)] = 0.0;
    }
}

/* Loop_functions.f90:37-51 */
static double loop_Az(double x, double y) {
  double r = pow(x * x + y * y, 0.5);
  return r < 0.3 ? 1e-3 * (0.3 - r) : 0.0;
}

void orc_calculate_Az(orc_state *s) {
  int Nx = NXr, Ny = NYr;
  double xa = s->p.xa, ya = s->p.ya, hx = s->hx, hy = s->hy;
  for (int i = 0; i <= Nx; i++)
    for (int j = 1; j <= Ny; j++) {
      double X = xa + i * hx;
      s->Bx[BXI(i, j, 1)] = -(1.0 / hy) * (loop_Az(X, ya + j * hy) - loop_Az(X, ya + (j - 1) * hy));
    }
  for (int i = 1; i <= Nx; i++)
    for (int j = 0; j <= Ny; j++) {
      double Y = ya + j * hy;
      s->By[BYI(i, j, 1)] = (1.0 / hx) * (loop_Az(xa + i * hx, Y) - loop_Az(xa + (i - 1) * hx, Y));
    }
}

void orc_calculate_L2error(orc_state *s, double t20, double L2[8]) {
  int Nx = NXr, Ny = NYr;
  const tables_t *T = &s->T;
  icdef_t c;
  memset(&c, 0, sizeof c);
  c.ic = ORC_IC_VORTEX;
  double xa = s->p.xa, ya = s->p.ya, hx = s->hx, hy = s->hy, hx1 = s->hx1, hy1 = s->hy1;
  for (int n = 0; n < 8; n++) L2[n] = 0.0;
  for (int i = 1; i <= Nx; i++)
    for (int j = 1; j <= Ny; j++) {
      double Xc = xa + (i - 0.5) * hx, Yc = ya + (j - 0.5) * hy;
      for (int i1 = 0; i1 < NG; i1++)
        for (int j1 = 0; j1 < NG; j1++) {
          double U[8];
          ic_eval(&c, Xc + hx1 * T->lambda[i1] - t20, Yc + hy1 * T->lambda[j1] - t20, U);
          for (int n = 1; n <= NE; n++) {
            double u = 0.0;
            for (int d = 1; d <= NP; d++) u = u + s->uh[UHI(i, j, d, n)] * T->phiG[d - 1][i1][j1];
            double e = u - U[n - 1];
            L2[n - 1] = L2[n - 1] + T->weight[i1] * T->weight[j1] * (e * e);
          }
        }
    }
  for (int n = 0; n < 8; n++) L2[n] = pow(L2[n] * hx1 * hy1, 0.5);
}

/* ---------------------------------------------------------------------- */
/* set_bc.f90:5-27: x halos first, then y halos (so corners are right)    */
void orc_set_bc(orc_state *s) {
  int Nx = NXr, Ny = NYr;
  const orc_params *p = &s->p;
  for (int n = 1; n <= NE; n++)
    for (int d = 1; d <= NP; d++)
      for (int j = 0; j <= Ny + 1; j++) {
        if (p->bcR == 1) s->uh[UHI(Nx + 1, j, d, n)] = s->uh[UHI(1, j, d, n)];
        else if (p->bcR == 2) s->uh[UHI(Nx + 1, j, d, n)] = s->uh[UHI(Nx, j, d, n)];
        if (p->bcL == 1) s->uh[UHI(0, j, d, n)] = s->uh[UHI(Nx, j, d, n)];
        else if (p->bcL == 2) s->uh[UHI(0, j, d, n)] = s->uh[UHI(1, j, d, n)];
      }
  for (int n = 1; n <= NE; n++)
    for (int d = 1; d <= NP; d++)
      for (int i = 0; i <= Nx + 1; i++) {
        if (p->bcU == 1) s->uh[UHI(i, Ny + 1, d, n)] = s->uh[UHI(i, 1, d, n)];
        else if (p->bcU == 2) s->uh[UHI(i, Ny + 1, d, n)] = s->uh[UHI(i, Ny, d, n)];
        if (p->bcD == 1) s->uh[UHI(i, 0, d, n)] = s->uh[UHI(i, Ny, d, n)];
        else if (p->bcD == 2) s->uh[UHI(i, 0, d, n)] = s->uh[UHI(i, 1, d, n)];
      }
}

/* ---------------------------------------------------------------------- */
/* Lh.f90:1-305: DG residual.  Also stores the face EMFs of LhEz.f90:6-28,
   which use the same traces and trace fluxes. */
void orc_Lh(orc_state *s) {
  int Nx = NXr, Ny = NYr;
  const tables_t *T = &s->T;
  const int ft = s->p.flux_type;
  const double hx = s->hx, hy = s->hy;
  memset(s->du, 0, s->nuh * sizeof(double));

  /* x faces: Lh.f90:72-136 (traces + fluxes), :206-234 (Riemann),
     LhEz.f90:6-16 (face EMF) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 1; j <= Ny; j++)
    for (int i = 0; i <= Nx; i++) {
      for (int q = 0; q < NG; q++) {
        double UR[8], UL[8], FR[8], FL[8], F[8], Us[8];
        for (int n = 1; n <= NE; n++) {
          double a = 0.0, b = 0.0;
          for (int d = 1; d <= NP; d++) {
            a = a + s->uh[UHI(i, j, d, n)] * T->phiGR[d - 1][q];
            b = b + s->uh[UHI(i + 1, j, d, n)] * T->phiGL[d - 1][q];
          }
          UR[n - 1] = a;  /* right trace of cell i   (UR(i,j,q,:))   */
          UL[n - 1] = b;  /* left trace of cell i+1  (UL(i+1,j,q,:)) */
        }
        lh_flux(UR, FR, NULL, 0);
        lh_flux(UL, FL, NULL, 0);
        double SRmax, SRmin, SLmax, SLmin;
        orc_eigenvalueMm(UR, 1, 0, &SRmax, &SRmin);
        orc_eigenvalueMm(UL, 1, 0, &SLmax, &SLmin);
        double SR = fmax2(SRmax, SLmax), SL = fmin2(SRmin, SLmin);
        orc_riemann_1d(ft, 1, SR, SL, /*UL1=*/UR, /*UR1=*/UL, /*FL1=*/FR, /*FR1=*/FL, F, Us);
        for (int n = 0; n < NE; n++) s->Fxhat[FXI(i, j) + q * NE + n] = F[n];
        double sa = UR[1] / UR[0], sb = UL[1] / UL[0];
        double S = fmax2(fabs(sa), fabs(sb));
        s->EzRL[EXI(i, j) + q] = -0.5 * (FR[6] + FL[6] - S * (UL[6] - UR[6]));
      }
    }
  /* y faces: Lh.f90:139-203, :237-264; LhEz.f90:18-28 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      for (int q = 0; q < NG; q++) {
        double UU[8], UD[8], FU[8], FD[8], F[8], Us[8];
        for (int n = 1; n <= NE; n++) {
          double a = 0.0, b = 0.0;
          for (int d = 1; d <= NP; d++) {
            a = a + s->uh[UHI(i, j, d, n)] * T->phiGU[d - 1][q];
            b = b + s->uh[UHI(i, j + 1, d, n)] * T->phiGD[d - 1][q];
          }
          UU[n - 1] = a;  /* upper trace of cell j   */
          UD[n - 1] = b;  /* lower trace of cell j+1 */
        }
        lh_flux(UU, NULL, FU, 0);
        lh_flux(UD, NULL, FD, 1);
        double SRmax, SRmin, SLmax, SLmin;
        orc_eigenvalueMm(UU, 0, 1, &SRmax, &SRmin);
        orc_eigenvalueMm(UD, 0, 1, &SLmax, &SLmin);
        double SR = fmax2(SRmax, SLmax), SL = fmin2(SRmin, SLmin);
        orc_riemann_1d(ft, 2, SR, SL, UU, UD, FU, FD, F, Us);
        for (int n = 0; n < NE; n++) s->Fyhat[FYI(i, j) + q * NE + n] = F[n];
        double sa = UU[2] / UU[0], sb = UD[2] / UD[0];
        double S = fmax2(fabs(sa), fabs(sb));
        s->EzUD[EYI(i, j) + q] = 0.5 * (FU[5] + FD[5] - S * (UD[5] - UU[5]));
      }
    }
  /* volume (Lh.f90:11-68) and surface (Lh.f90:268-304) integrals */
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 1; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      double uG[NE][NG][NG], Fx[NE][NG][NG], Fy[NE][NG][NG];
      for (int n = 1; n <= NE; n++)
        for (int i1 = 0; i1 < NG; i1++)
          for (int j1 = 0; j1 < NG; j1++) {
            double acc = 0.0;
            for (int d = 1; d <= NP; d++) acc = acc + s->uh[UHI(i, j, d, n)] * T->phiG[d - 1][i1][j1];
            uG[n - 1][i1][j1] = acc;
          }
      for (int i1 = 0; i1 < NG; i1++)
        for (int j1 = 0; j1 < NG; j1++) {
          double U[8], fx[8], fy[8];
          for (int n = 0; n < NE; n++) U[n] = uG[n][i1][j1];
          lh_flux(U, fx, fy, 0);
          for (int n = 0; n < NE; n++) { Fx[n][i1][j1] = fx[n]; Fy[n][i1][j1] = fy[n]; }
        }
      for (int d = 1; d <= NP; d++)
        for (int n = 1; n <= NE; n++) {
          if (!active(d, n)) continue;
          double acc = 0.0;
          if (d >= 2)
            for (int i1 = 0; i1 < NG; i1++)
              for (int j1 = 0; j1 < NG; j1++)
                acc = acc + 0.25 * T->weight[i1] * T->weight[j1] *
                                (Fx[n - 1][i1][j1] * T->phixG[d - 1][i1][j1] + Fy[n - 1][i1][j1] * T->phiyG[d - 1][i1][j1]);
          for (int q = 0; q < NG; q++)
            acc = acc - (0.5 / hx) * T->weight[q] *
                            (s->Fxhat[FXI(i, j) + q * NE + n - 1] * T->phiGR[d - 1][q] -
                             s->Fxhat[FXI(i - 1, j) + q * NE + n - 1] * T->phiGL[d - 1][q]);
          for (int q = 0; q < NG; q++)
            acc = acc - (0.5 / hy) * T->weight[q] *
                            (s->Fyhat[FYI(i, j) + q * NE + n - 1] * T->phiGU[d - 1][q] -
                             s->Fyhat[FYI(i, j - 1) + q * NE + n - 1] * T->phiGD[d - 1][q]);
          s->du[UHI(i, j, d, n)] = acc / T->mm[d - 1];
        }
    }
}

/* LhEz.f90:33-96: vertex EMF and the constrained-transport right-hand side.
   Requires orc_Lh first (face EMFs). */
void orc_LhEz(orc_state *s) {
  int Nx = NXr, Ny = NYr;
  const tables_t *T = &s->T;
  const int ft = s->p.flux_type;
  const double hx = s->hx, hy = s->hy;
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j <= Ny; j++)
    for (int i = 0; i <= Nx; i++) {
      double URU1[8], ULU1[8], URD1[8], ULD1[8];
      for (int n = 1; n <= NE; n++) {
        double a = 0, b = 0, c = 0, e = 0;
        for (int d = 1; d <= NP; d++) {
          a = a + s->uh[UHI(i + 1, j + 1, d, n)] * T->phiLD[d - 1];
          b = b + s->uh[UHI(i, j + 1, d, n)] * T->phiRD[d - 1];
          c = c + s->uh[UHI(i + 1, j, d, n)] * T->phiLU[d - 1];
          e = e + s->uh[UHI(i, j, d, n)] * T->phiRU[d - 1];
        }
        URU1[n - 1] = a; ULU1[n - 1] = b; URD1[n - 1] = c; ULD1[n - 1] = e;
      }
      s->EzVertex[EVI(i, j)] = orc_vertex_emf(ft, URU1, ULU1, URD1, ULD1);
    }
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 1; j <= Ny; j++)
    for (int i = 0; i <= Nx; i++)
      for (int d = 1; d <= 3; d++) {
        double acc = 0.0;
        for (int q = 0; q < NG; q++) acc = acc + 0.5 * T->weight[q] * s->EzRL[EXI(i, j) + q] * T->EzyG[d - 1][q];
        s->dEz1[BXI(i, j, d)] = (acc - s->EzVertex[EVI(i, j)] * T->EzU[d - 1] / hy +
                                 s->EzVertex[EVI(i, j - 1)] * T->EzD[d - 1] / hy) / T->mmE[d - 1];
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++)
      for (int d = 1; d <= 3; d++) {
        double acc = 0.0;
        for (int q = 0; q < NG; q++) acc = acc - 0.5 * T->weight[q] * s->EzUD[EYI(i, j) + q] * T->EzxG[d - 1][q];
        s->dEz2[BYI(i, j, d)] = (acc + s->EzVertex[EVI(i, j)] * T->EzR[d - 1] / hx -
                                 s->EzVertex[EVI(i - 1, j)] * T->EzL[d - 1] / hx) / T->mmE[d - 1];
      }
}

/* ---------------------------------------------------------------------- */
/* TVB_Limiter_P2.f90:16-126 + minmod.f90:5-27 */
static void minmod8(const double DU[8], const double DR[8], const double DL[8], double Mh2, double out[8]) {
  for (int i = 0; i < 8; i++) {
    if (fabs(DU[i]) <= Mh2) {
      out[i] = DU[i];
    } else {
      double a = fsign1(DU[i]), b = fsign1(DR[i]), c = fsign1(DL[i]);
      double sv = (a + b + c) / 3.0;
      if (fabs(sv) == 1) out[i] = sv * fmin2(fmin2(fabs(DU[i]), fabs(DR[i])), fabs(DL[i]));
      else out[i] = 0.0;
    }
  }
}
static inline void matvec8(const double A[64], const double x[8], double y[8]) {
  for (int r = 0; r < 8; r++) {
    double acc = 0.0;
    for (int c = 0; c < 8; c++) acc = acc + A[r + 8 * c] * x[c];
    y[r] = acc;
  }
}

/* limit one cell; writes uhmod.  Reads uh (the current, unlimited stage
   value, whose halos hold whatever the caller left there). */
static void tvb_cell(orc_state *s, int i, int j) {
  double U[8], R[64], L[64];
  double dR[8], dL[8], r1[8], l1[8], tR[8], tL[8], tU[8], mod[8];
  double R1mod[8], L1mod[8], U1mod[8], D1mod[8];
  int change[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double M = s->p.M;
  for (int n = 1; n <= NE; n++) U[n - 1] = s->uh[UHI(i, j, 1, n)];
  /* x direction */
  orc_eigenmatrix(R, L, U, 1.0, 0.0, 0.0);
  for (int n = 1; n <= NE; n++) {
    dR[n - 1] = s->uh[UHI(i + 1, j, 1, n)] - s->uh[UHI(i, j, 1, n)];
    dL[n - 1] = s->uh[UHI(i, j, 1, n)] - s->uh[UHI(i - 1, j, 1, n)];
    r1[n - 1] = s->uh[UHI(i, j, 2, n)] + (2.0 / 3.0) * s->uh[UHI(i, j, 4, n)];
    l1[n - 1] = s->uh[UHI(i, j, 2, n)] - (2.0 / 3.0) * s->uh[UHI(i, j, 4, n)];
  }
  matvec8(L, dR, tR);
  matvec8(L, dL, tL);
  matvec8(L, r1, tU);
  minmod8(tU, tR, tL, M * (s->hx * s->hx), mod);
  matvec8(R, mod, R1mod);
  for (int d = 0; d < 8; d++) if (fabs(R1mod[d] - r1[d]) > 1e-6) change[d] = 1;
  matvec8(L, l1, tU);
  minmod8(tU, tR, tL, M * (s->hx * s->hx), mod);
  matvec8(R, mod, L1mod);
  for (int d = 0; d < 8; d++) if (fabs(L1mod[d] - l1[d]) > 1e-6) change[d] = 1;
  /* y direction */
  orc_eigenmatrix(R, L, U, 0.0, 1.0, 0.0);
  for (int n = 1; n <= NE; n++) {
    dR[n - 1] = s->uh[UHI(i, j + 1, 1, n)] - s->uh[UHI(i, j, 1, n)];
    dL[n - 1] = s->uh[UHI(i, j, 1, n)] - s->uh[UHI(i, j - 1, 1, n)];
    r1[n - 1] = s->uh[UHI(i, j, 3, n)] + (2.0 / 3.0) * s->uh[UHI(i, j, 6, n)];
    l1[n - 1] = s->uh[UHI(i, j, 3, n)] - (2.0 / 3.0) * s->uh[UHI(i, j, 6, n)];
  }
  matvec8(L, dR, tR);
  matvec8(L, dL, tL);
  matvec8(L, r1, tU);
  minmod8(tU, tR, tL, M * (s->hy * s->hy), mod);
  matvec8(R, mod, U1mod);
  for (int d = 0; d < 8; d++) if (fabs(U1mod[d] - r1[d]) > 1e-6) change[d] = 1;
  matvec8(L, l1, tU);
  minmod8(tU, tR, tL, M * (s->hy * s->hy), mod);
  matvec8(R, mod, D1mod);
  for (int d = 0; d < 8; d++) if (fabs(D1mod[d] - l1[d]) > 1e-6) change[d] = 1;
  for (int n = 1; n <= NE; n++) {
    if (!change[n - 1]) continue;
    for (int d = 7; d <= NP; d++) s->uhmod[UHI(i, j, d, n)] = 0.0;
    s->uhmod[UHI(i, j, 5, n)] = 0.0;
    s->uhmod[UHI(i, j, 2, n)] = 0.5 * (R1mod[n - 1] + L1mod[n - 1]);
    s->uhmod[UHI(i, j, 3, n)] = 0.5 * (U1mod[n - 1] + D1mod[n - 1]);
    s->uhmod[UHI(i, j, 4, n)] = (3.0 / 4.0) * (R1mod[n - 1] - L1mod[n - 1]);
    s->uhmod[UHI(i, j, 6, n)] = (3.0 / 4.0) * (U1mod[n - 1] - D1mod[n - 1]);
  }
}

void orc_TVB_Limiter_P2(orc_state *s) {
  int Nx = NXr, Ny = NYr;
  if (s->p.indicator == ORC_IND_KXRCF) orc_KXRCF_Detector(s);  /* TVB_Limiter_P2.f90:12 */
  memcpy(s->uhmod, s->uh, s->nuh * sizeof(double));
  if (s->p.indicator != ORC_IND_NONE) {
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
    for (int j = 1; j <= Ny; j++)
      for (int i = 1; i <= Nx; i++) {
        if (s->p.indicator == ORC_IND_KXRCF && s->trouble[(i - 1) + (size_t)Nx * (j - 1)] != 1) continue;
        tvb_cell(s, i, j);
      }
  }
  memcpy(s->uh, s->uhmod, s->nuh * sizeof(double));
  if (s->p.pp) orc_pp_Limiter(s);   /* TVB_Limiter_P2.f90:123 when enabled */
  orc_set_bc(s);                    /* TVB_Limiter_P2.f90:126 */
  /* :128-187 compute a B-face limiter whose result is discarded (:186-187) */
}

/* ---------------------------------------------------------------------- */
/* pp_Limiter.f90:1-125 */
static void calculate_tq(const double ubar[8], const double uq[8], double *tq) {
  double ta = 0, tb = 1, t = 0, ut[8];
  while (tb - ta > 1e-14) {
    t = 0.5 * (ta + tb);
    for (int n = 0; n < 8; n++) ut[n] = t * uq[n] + (1 - t) * ubar[n];
    if (pressure(ut) < 1e-12) tb = t;
    else ta = t;
  }
  *tq = ta;
}
static double norm8(const double x[8]) {
  double d = 0;
  for (int i = 0; i < 8; i++) d = d + x[i] * x[i];
  return sqrt(d);
}

void orc_pp_Limiter(orc_state *s) {
  int Nx = NXr, Ny = NYr;
  const tables_t *T = &s->T;
  const double eps = 1e-13;
  /* density pass */
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 1; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      double eta = 1.0;
      double rbar = s->uh[UHI(i, j, 1, 1)];
      for (int i1 = 0; i1 < NG; i1++)
        for (int j1 = 0; j1 < NG; j1++)
          for (int f = 0; f < 2; f++) {
            double rq = 0.0;
            for (int d = 1; d <= NP; d++) rq = rq + s->uh[UHI(i, j, d, 1)] * T->phiGLL[f][d - 1][i1][j1];
            double eta1 = fabs((rbar - eps) / (rbar - rq));
            if (eta1 < 1) eta = eta1;
          }
      if (eta < 1)
        for (int d = 2; d <= NP; d++) s->uh[UHI(i, j, d, 1)] = 0.5 * eta * s->uh[UHI(i, j, d, 1)];
    }
  /* pressure pass */
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
  for (int j = 1; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      double eta = 1.0, eta1 = 1.0;
      double ubar[8], uq[8], sq[8], dd[8];
      for (int n = 1; n <= NE; n++) ubar[n - 1] = s->uh[UHI(i, j, 1, n)];
      for (int i1 = 0; i1 < NG; i1++)
        for (int j1 = 0; j1 < NG; j1++)
          for (int f = 0; f < 2; f++) {
            for (int n = 1; n <= NE; n++) {
              double a = 0.0;
              for (int d = 1; d <= NP; d++) a = a + s->uh[UHI(i, j, d, n)] * T->phiGLL[f][d - 1][i1][j1];
              uq[n - 1] = a;
            }
            double pq = pressure(uq);
            if (pq < 0) {
              double tq;
              calculate_tq(ubar, uq, &tq);
              for (int n = 0; n < 8; n++) sq[n] = tq * uq[n] + (1 - tq) * ubar[n];
              for (int n = 0; n < 8; n++) dd[n] = sq[n] - ubar[n];
              double n1 = norm8(dd);
              for (int n = 0; n < 8; n++) dd[n] = uq[n] - ubar[n];
              double n2 = norm8(dd);
              eta1 = n1 / n2;
            }
            if (eta1 < eta) eta = eta1;
          }
      if (eta < 1) eta = 0.5 * eta;
      for (int n = 1; n <= NE; n++)
        for (int d = 2; d <= NP; d++) s->uh[UHI(i, j, d, n)] = eta * s->uh[UHI(i, j, d, n)];
    }
}

/* ---------------------------------------------------------------------- */
/* KXRCF_Detector.f90:7-107 (fixed variant: uG local, flags reset, Ck parameter) */
void orc_KXRCF_Detector(orc_state *s) {
  int Nx = NXr, Ny = NYr;
  const tables_t *T = &s->T;
  const double hx = s->hx, hy = s->hy, hx1 = s->hx1, hy1 = s->hy1, Ck = s->p.Ck;
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 1; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      double I[8] = {0, 0, 0, 0, 0, 0, 0, 0}, Lij = 1e-12, uinf[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      /* traces of this cell and of its neighbours on the shared faces */
      if (s->uh[UHI(i + 1, j, 1, 2)] < 0) {   /* right side inflow */
        for (int q = 0; q < NG; q++) {
          for (int n = 1; n <= NE; n++) {
            double a = 0, b = 0;
            for (int d = 1; d <= NP; d++) {
              a = a + s->uh[UHI(i, j, d, n)] * T->phiGR[d - 1][q];
              b = b + s->uh[UHI(i + 1, j, d, n)] * T->phiGL[d - 1][q];
            }
            I[n - 1] = I[n - 1] + hy1 * T->weight[q] * (a - b);
          }
          Lij = Lij + hy;
        }
      }
      if (s->uh[UHI(i - 1, j, 1, 2)] > 0) {   /* left side inflow */
        for (int q = 0; q < NG; q++) {
          for (int n = 1; n <= NE; n++) {
            double a = 0, b = 0;
            for (int d = 1; d <= NP; d++) {
              a = a + s->uh[UHI(i, j, d, n)] * T->phiGL[d - 1][q];
              b = b + s->uh[UHI(i - 1, j, d, n)] * T->phiGR[d - 1][q];
            }
            I[n - 1] = I[n - 1] + hy1 * T->weight[q] * (a - b);
          }
          Lij = Lij + hy;
        }
      }
      if (s->uh[UHI(i, j + 1, 1, 3)] < 0) {   /* upper side inflow */
        for (int q = 0; q < NG; q++) {
          for (int n = 1; n <= NE; n++) {
            double a = 0, b = 0;
            for (int d = 1; d <= NP; d++) {
              a = a + s->uh[UHI(i, j, d, n)] * T->phiGU[d - 1][q];
              b = b + s->uh[UHI(i, j + 1, d, n)] * T->phiGD[d - 1][q];
            }
            I[n - 1] = I[n - 1] + hx1 * T->weight[q] * (a - b);
          }
          Lij = Lij + hx;
        }
      }
      if (s->uh[UHI(i, j - 1, 1, 3)] > 0) {   /* lower side inflow */
        for (int q = 0; q < NG; q++) {
          for (int n = 1; n <= NE; n++) {
            double a = 0, b = 0;
            for (int d = 1; d <= NP; d++) {
              a = a + s->uh[UHI(i, j, d, n)] * T->phiGD[d - 1][q];
              b = b + s->uh[UHI(i, j - 1, d, n)] * T->phiGU[d - 1][q];
            }
            I[n - 1] = I[n - 1] + hx1 * T->weight[q] * (a - b);
          }
          Lij = Lij + hx;
        }
      }
      for (int i1 = 0; i1 < NG; i1++)
        for (int j1 = 0; j1 < NG; j1++)
          for (int n = 1; n <= NE; n++) {
            double a = 0.0;
            for (int d = 1; d <= NP; d++) a = a + s->uh[UHI(i, j, d, n)] * T->phiG[d - 1][i1][j1];
            if (fabs(a) + 1e-12 > uinf[n - 1]) uinf[n - 1] = fabs(a) + 1e-12;
          }
      int flag = 0;
      double hk = pow(hx * hy, 0.25 * (2 + 1));
      for (int n = 0; n < NE; n++) {
        double In = fabs(I[n]) / (hk * Lij * uinf[n]);
        if (In >= Ck) flag = 1;
      }
      s->trouble[(i - 1) + (size_t)Nx * (j - 1)] = flag;
    }
}

/* ---------------------------------------------------------------------- */
/* div_free_Balsara.f90:13-79 */
void orc_div_free_Balsara(orc_state *s) {
  int Nx = NXr, Ny = NYr;
  const double hx = s->hx, hy = s->hy;
  for (int n = 6; n <= 7; n++)
    for (int d = 1; d <= NP; d++)
      for (int j = 0; j <= Ny + 1; j++)
        for (int i = 0; i <= Nx + 1; i++) s->uh[UHI(i, j, d, n)] = 0.0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 1; j <= Ny; j++)
    for (int i = 1; i <= Nx; i++) {
      double a0R = s->Bx[BXI(i, j, 1)], a1R = s->Bx[BXI(i, j, 2)], a2R = s->Bx[BXI(i, j, 3)];
      double a0L = s->Bx[BXI(i - 1, j, 1)], a1L = s->Bx[BXI(i - 1, j, 2)], a2L = s->Bx[BXI(i - 1, j, 3)];
      double b0U = s->By[BYI(i, j, 1)], b1U = s->By[BYI(i, j, 2)], b2U = s->By[BYI(i, j, 3)];
      double b0D = s->By[BYI(i, j - 1, 1)], b1D = s->By[BYI(i, j - 1, 2)], b2D = s->By[BYI(i, j - 1, 3)];
      double rxy = hx / hy, ryx = hy / hx;
      double h = 1.0 / 2.0;
      s->uh[UHI(i, j, 1, 6)] = h * (a0R + a0L) + (1.0 / 6.0) * rxy * (b1U - b1D);
      s->uh[UHI(i, j, 2, 6)] = h * (a0R - a0L) + (1.0 / 15.0) * rxy * (b2U - b2D);
      s->uh[UHI(i, j, 3, 6)] = h * (a1R + a1L);
      s->uh[UHI(i, j, 4, 6)] = -(1.0 / 4.0) * rxy * (b1U - b1D);
      s->uh[UHI(i, j, 5, 6)] = h * (a1R - a1L);
      s->uh[UHI(i, j, 6, 6)] = h * (a2R + a2L);
      s->uh[UHI(i, j, 7, 6)] = -(1.0 / 6.0) * rxy * (b2U - b2D);
      s->uh[UHI(i, j, 9, 6)] = h * (a2R - a2L);
      s->uh[UHI(i, j, 1, 7)] = h * (b0U + b0D) + (1.0 / 6.0) * ryx * (a1R - a1L);
      s->uh[UHI(i, j, 3, 7)] = h * (b0U - b0D) + (1.0 / 15.0) * ryx * (a2R - a2L);
      s->uh[UHI(i, j, 2, 7)] = h * (b1U + b1D);
      s->uh[UHI(i, j, 6, 7)] = -(1.0 / 4.0) * ryx * (a1R - a1L);
      s->uh[UHI(i, j, 5, 7)] = h * (b1U - b1D);
      s->uh[UHI(i, j, 4, 7)] = h * (b2U + b2D);
      s->uh[UHI(i, j, 10, 7)] = -(1.0 / 6.0) * ryx * (a2R - a2L);
      s->uh[UHI(i, j, 8, 7)] = h * (b2U - b2D);
    }
}

/* ---------------------------------------------------------------------- */
/* calculate_dt.f90:5-27 (CFL is a parameter; the file hard-codes 0.15)    */
void orc_calculate_dt(orc_state *s) {
  int Nx = NXr, Ny = NYr;
  double ax = 0, ay = 0;
  for (int i = 1; i <= Nx; i++)
    for (int j = 1; j <= Ny; j++) {
      double U[8], a1, a2;
      for (int n = 1; n <= NE; n++) U[n - 1] = s->uh[UHI(i, j, 1, n)];
      orc_eigenvalueMm(U, 1, 0, &a1, &a2);
      if (fabs(a1) > ax || fabs(a2) > ax) ax = fmax2(fabs(a1), fabs(a2));
      orc_eigenvalueMm(U, 0, 1, &a1, &a2);
      if (fabs(a1) > ay || fabs(a2) > ay) ay = fmax2(fabs(a1), fabs(a2));
    }
  s->alphax = ax;
  s->alphay = ay;
  s->dt = s->p.CFL / (ax / s->hx + ay / s->hy);
}

/* calculate_umax.f90:5-13 */
void orc_calculate_umax(orc_state *s) {
  double m = 0;
  for (int i = 1; i <= NXr; i++)
    for (int j = 1; j <= NYr; j++)
      if (fabs(s->uh[UHI(i, j, 1, 1)]) > m) m = fabs(s->uh[UHI(i, j, 1, 1)]);
  s->umax = m;
}

/* calculate_pmin.f90:5-26 */
void orc_calculate_pmin(orc_state *s) {
  const tables_t *T = &s->T;
  double pm = 1000000;
  for (int i = 1; i <= NXr; i++)
    for (int j = 1; j <= NYr; j++)
      for (int i1 = 0; i1 < NG; i1++)
        for (int j1 = 0; j1 < NG; j1++) {
          double U[8];
          for (int n = 1; n <= NE; n++) {
            double a = 0.0;
            for (int d = 1; d <= NP; d++) a = a + s->uh[UHI(i, j, d, n)] * T->phiG[d - 1][i1][j1];
            U[n - 1] = a;
          }
          double p1 = pressure(U);
          if (p1 < pm) pm = p1;
        }
  s->pmin = pm;
}

/* calculate_totaldiv.f90:5-23 */
void orc_calculate_totaldiv(orc_state *s) {
  const tables_t *T = &s->T;
  double tot = 0;
  for (int i = 1; i <= NXr; i++)
    for (int j = 1; j <= NYr; j++) {
      double g[NG][NG];
      for (int i1 = 0; i1 < NG; i1++)
        for (int j1 = 0; j1 < NG; j1++) {
          double a = 0.0;
          for (int d = 1; d <= NP; d++)
            a = a + s->uh[UHI(i, j, d, 6)] * T->phixG[d - 1][i1][j1] + s->uh[UHI(i, j, d, 7)] * T->phiyG[d - 1][i1][j1];
          g[i1][j1] = a;
        }
      for (int i1 = 0; i1 < NG; i1++)
        for (int j1 = 0; j1 < NG; j1++) tot = tot + T->weight[i1] * T->weight[j1] * (g[i1][j1] * g[i1][j1]);
    }
  s->totaldiv = sqrt(tot * s->hx1 * s->hy1);
}

/* ---------------------------------------------------------------------- */
/* RK3.f90 */
void orc_rk3_prologue(orc_state *s) {
  s->t = 0;
  s->count = 0;
  orc_set_bc(s);
  orc_TVB_Limiter_P2(s);
  orc_div_free_Balsara(s);
  orc_calculate_umax(s);
  orc_calculate_totaldiv(s);
  orc_calculate_pmin(s);
}

static void axpy_whole(double *out, const double *x, double a, const double *y, size_t n) {
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < n; k++) out[k] = x[k] + a * y[k];
}
static void lin2(double *out, double a, const double *x, double b, const double *y, size_t n) {
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < n; k++) out[k] = a * x[k] + b * y[k];
}

/* RK3.f90:74-139 (the three stages) */
void orc_rk3_stages(orc_state *s, double dt) {
  size_t nu = s->nuh, nx = s->nbx, ny = s->nby;
  /* stage 1 */
  orc_set_bc(s);
  orc_Lh(s);
  orc_LhEz(s);
  axpy_whole(s->uI, s->uh, dt, s->du, nu);
  axpy_whole(s->BxI, s->Bx, dt, s->dEz1, nx);
  axpy_whole(s->ByI, s->By, dt, s->dEz2, ny);
  memcpy(s->uh0, s->uh, nu * sizeof(double));
  memcpy(s->Bx0, s->Bx, nx * sizeof(double));
  memcpy(s->By0, s->By, ny * sizeof(double));
  memcpy(s->uh, s->uI, nu * sizeof(double));
  memcpy(s->Bx, s->BxI, nx * sizeof(double));
  memcpy(s->By, s->ByI, ny * sizeof(double));
  orc_TVB_Limiter_P2(s);
  orc_div_free_Balsara(s);
  orc_set_bc(s);
  /* stage 2 */
  orc_Lh(s);
  orc_LhEz(s);
  axpy_whole(s->uh, s->uI, dt, s->du, nu);
  axpy_whole(s->Bx, s->BxI, dt, s->dEz1, nx);
  axpy_whole(s->By, s->ByI, dt, s->dEz2, ny);
  lin2(s->uII, 3.0 / 4.0, s->uh0, 1.0 / 4.0, s->uh, nu);
  lin2(s->BxII, 3.0 / 4.0, s->Bx0, 1.0 / 4.0, s->Bx, nx);
  lin2(s->ByII, 3.0 / 4.0, s->By0, 1.0 / 4.0, s->By, ny);
  memcpy(s->uh, s->uII, nu * sizeof(double));
  memcpy(s->Bx, s->BxII, nx * sizeof(double));
  memcpy(s->By, s->ByII, ny * sizeof(double));
  orc_TVB_Limiter_P2(s);
  orc_div_free_Balsara(s);
  /* stage 3 */
  orc_set_bc(s);
  orc_Lh(s);
  orc_LhEz(s);
  axpy_whole(s->Bx, s->BxII, dt, s->dEz1, nx);
  axpy_whole(s->By, s->ByII, dt, s->dEz2, ny);
  axpy_whole(s->uh, s->uII, dt, s->du, nu);
  lin2(s->uh, 1.0 / 3.0, s->uh0, 2.0 / 3.0, s->uh, nu);
  lin2(s->Bx, 1.0 / 3.0, s->Bx0, 2.0 / 3.0, s->Bx, nx);
  lin2(s->By, 1.0 / 3.0, s->By0, 2.0 / 3.0, s->By, ny);
  orc_TVB_Limiter_P2(s);
  orc_div_free_Balsara(s);
}

/* RK3.f90:58-174 (one pass of the time loop, without frame output) */
int orc_rk3_step(orc_state *s, double tend) {
  if (!(s->t < tend)) return 0;
  orc_calculate_dt(s);
  if (s->t + s->dt >= tend) {
    s->dt = tend - s->t;
    s->t = tend;
  } else {
    s->t = s->t + s->dt;
  }
  if (s->dt < 1e-7) s->t = tend;
  orc_rk3_stages(s, s->dt);
  s->count++;
  orc_calculate_umax(s);
  orc_calculate_totaldiv(s);
  orc_calculate_pmin(s);
  return 1;
}
