/*
 * orc_internal.h -- data structures shared by the oracle (uft_oracle.c) and
 * the threaded reference-structure CPU baseline (cpu_ref.c).
 *
 * TEST INFRASTRUCTURE ONLY (see uft_oracle.c).
 */
#pragma once
#include <stdint.h>

#define ORC_MAXM 8

typedef struct {
  int64_t n;
  int64_t *rp;
  int32_t *ci;
  double *v;
} csr_t;

typedef struct {
  /* OptPara defaults (OptPara.cpp:66-78) overridable like the options DB (:92-104) */
  double Emin, Emax, penal, volfrac, rmin, nu, movlim, Xmin, Xmax, x0, beta0, betamax;
  int proj, m, maxItr;
  double rtol_el, rtol_f;   /* -ksp_rtol for elasticity / filter (both 1e-5 by default) */
  int maxit_el, maxit_f;
  int ke_cache;             /* 0: recompute KE twice per iteration like the reference */
} orc_params;

typedef struct {
  int npe, nd;
  int64_t nn, ne, ndof;
  int32_t *elem;
  double *xyz;
  int nfixed;
  int32_t *fixed;
  int nlc, nload;
  int32_t *ldofs;
  double *lvals;
  orc_params P;
  /* derived */
  uint8_t *isfixed;
  double *vol, TV;
  double *KEc;   /* cached KE (ne * nd * nd) when ke_cache */
  double *Tel;   /* ne * npe */
  csr_t K, KF;
  double *KFdiag;
  double *U, *F, *TX, *RX;
  /* design state (OptPara.h:28-52) */
  double *X, *XFiled, *XPhys, *Xold, *dfdx, *dgdx, *xmin, *xmax;
  double fx, gx[ORC_MAXM];
  double beta, btemp;
  /* MMA state (MMA.h:76-109) */
  int mk;
  double asyminit, asymdec, asyminc;
  int robust, conmod;   /* RobustAsymptotesType, constraintModification (MMA.h) */
  int load_separate;    /* multi-load-case form: one solve per case, summed compliance */
  double *Ucase;        /* displacements of every case (nlc * ndof) in that form */
  double a[ORC_MAXM], c[ORC_MAXM], d[ORC_MAXM], y[ORC_MAXM], z;
  double lam[ORC_MAXM], mu[ORC_MAXM], s[2 * ORC_MAXM], b[ORC_MAXM], grad[ORC_MAXM], Hess[ORC_MAXM * ORC_MAXM];
  double *L, *Uas, *alpha, *betav, *p0, *q0, *pij, *qij, *xo1, *xo2;
  /* diagnostics */
  int it_el, it_f, dip_steps;
  double relres_el;
} orc_ctx;

double orc_hex8_ke(const double *X, double nu, double *KE);
double orc_tet4_ke(const double *X, double nu, double *KE);
void orc_hex8_kehel(const double *X, double rmin, double *KF, double *T);
void orc_tet4_kehel(const double *X, double rmin, double *KF, double *T);
void orc_mma_movelimit(orc_ctx *C, const double *x, double *xmin, double *xmax);
void orc_mma_update(orc_ctx *C, double *x, const double *dfdx, const double *gx, const double *dgdx,
                    const double *xmin, const double *xmax);
double orc_design_change(orc_ctx *C, const double *x, double *xold);
void orc_heaviside(orc_ctx *C, const double *xf, double *xp, int itr);
