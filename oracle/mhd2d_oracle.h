/*
 * mhd2d_oracle.h -- CPU restatement of the 2D ideal-MHD P2 globally
 * divergence-free DG step of GDF_Balsara_P2_0316 (Fortran, COMMON-block
 * code).  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and
 * the CPU-baseline leg of bench.py use it as the checker.  Nothing in the
 * product library links or calls it.
 *
 * Arrays use the reference's own column-major layouts (GDF_Balsara_P2_0316/
 * com.txt:22,48) so that dumps of the reference and of this restatement can
 * be compared element by element:
 *   uh(0:Nx+1, 0:Ny+1, 10, 8)   Bx(0:Nx, 1:Ny, 3)   By(1:Nx, 0:Ny, 3)
 *   du  like uh;  dEz1 like Bx;  dEz2 like By;  EzVertex(0:Nx, 0:Ny)
 */
#ifndef MHD2D_ORACLE_H
#define MHD2D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LF = 1, ORC_HLL = 2, ORC_HLLC = 3, ORC_HLLD = 4 };
enum { ORC_IND_NONE = 0, ORC_IND_ALL = 1, ORC_IND_KXRCF = 2 };
enum { ORC_IC_ORSZAG_TANG = 3, ORC_IC_VORTEX = 1, ORC_IC_ROTOR = 4,
       ORC_IC_BLAST = 5, ORC_IC_CLOUD_SHOCK = 6, ORC_IC_CLOUD_SHOCK_2 = 61, ORC_IC_MAGNETIC_LOOP = 7,
       ORC_IC_BRIO_WU = 8, ORC_IC_BRIO_WU_Y = 81, ORC_IC_RYU_JONES = 9, ORC_IC_RYU_JONES_Y = 91,
       ORC_IC_OT_PERTURBED = 103, ORC_IC_RIEMANN4 = 104 };

typedef struct orc_params {
  int Nx, Ny;
  int flux_type;               /* 1 LF, 2 HLL, 3 HLLC, 4 HLLD (com.txt:1)   */
  int bcR, bcL, bcU, bcD;      /* 1 periodic, 2 outflow (set_bc.f90)        */
  int pp;                      /* 1: pp_Limiter after TVB (TVB_Limiter_P2.f90:123) */
  int indicator;               /* which cells TVB limits                    */
  int nthreads;                /* OpenMP threads (0 = runtime default)      */
  double xa, xb, ya, yb;
  double M;                    /* TVB constant (initN.txt: M = ...)         */
  double CFL;                  /* calculate_dt.f90:8 hard-codes 0.15        */
  double Ck;                   /* KXRCF threshold                           */
} orc_params;

typedef struct orc_state orc_state;

orc_state *orc_create(const orc_params *p);
void orc_destroy(orc_state *s);

/* raw access to the reference-layout arrays (owned by the state) */
double *orc_uh(orc_state *s);
double *orc_Bx(orc_state *s);
double *orc_By(orc_state *s);
double *orc_du(orc_state *s);
double *orc_dEz1(orc_state *s);
double *orc_dEz2(orc_state *s);
double *orc_EzVertex(orc_state *s);
int32_t *orc_trouble(orc_state *s);
/* scalars: t, dt, umax, totaldiv, pmin, alphax, alphay, count */
void orc_scalars(orc_state *s, double out[8]);
void orc_set_time(orc_state *s, double t, double dt);

/* IC projection (init_data.f90:21-121); extra[] = IC-family parameters */
void orc_init_data(orc_state *s, int ic, const double *extra, int nextra);
/* calculate_Az.f90:1-26: mode 1 of Bx, By from the magnetic loop's vector potential */
void orc_calculate_Az(orc_state *s);
/* calculate_L2error.f90:1-37: L2 error of each variable against the vortex
   (init1.txt) translated by (t20, t20) */
void orc_calculate_L2error(orc_state *s, double t20, double L2[8]);

/* the reference's subroutines, one call each */
void orc_set_bc(orc_state *s);
void orc_Lh(orc_state *s);
void orc_LhEz(orc_state *s);
void orc_TVB_Limiter_P2(orc_state *s);
void orc_pp_Limiter(orc_state *s);
void orc_KXRCF_Detector(orc_state *s);
void orc_div_free_Balsara(orc_state *s);
void orc_calculate_dt(orc_state *s);
void orc_calculate_umax(orc_state *s);
void orc_calculate_pmin(orc_state *s);
void orc_calculate_totaldiv(orc_state *s);

/* RK3.f90:8-18 (limit + diagnostics before the loop) */
void orc_rk3_prologue(orc_state *s);
/* one pass of the RK3.f90:58-174 loop body: dt, clip to tend, 3 stages,
   diagnostics.  Returns 1 if a step was taken, 0 if t >= tend. */
int orc_rk3_step(orc_state *s, double tend);
/* the three stages with a given dt (no dt computation, no clip) */
void orc_rk3_stages(orc_state *s, double dt);

/* pointwise helpers exposed for unit tests */
void orc_eigenvalueMm(const double U[8], int n1, int n2, double *amax, double *amin);
void orc_riemann_1d(int flux_type, int direction, double SR, double SL,
                    const double UL[8], const double UR[8],
                    const double FL[8], const double FR[8],
                    double Fhat[8], double Ustar[8]);
double orc_vertex_emf(int flux_type, const double URU[8], const double ULU[8],
                      const double URD[8], const double ULD[8]);
void orc_eigenmatrix(double R[64], double L[64], const double U[8],
                     double n1, double n2, double n3);
void orc_mhd_flux(const double U[8], double Fx[8], double Fy[8]);

#ifdef __cplusplus
}
#endif
#endif
