/*
 * uft_oracle.c -- CPU FP64 restatement of the reference SIMP design iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library never links it.
 *
 * Every routine cites the reference file:line it restates (paths relative to
 * the reference repository root).  The linear algebra that the reference
 * delegates to PETSc (MPIAIJ assembly, MatZeroRowsColumns, KSP CG + Jacobi,
 * FGMRES + GAMG) is restated by its mathematical definition: the assembled
 * matrix is built explicitly in CSR, Dirichlet rows/columns are eliminated
 * with a unit diagonal, and the systems are solved with Jacobi-PCG to the
 * requested relative tolerance.  GAMG is a solver choice, not a
 * result-defining step, so the oracle only has to solve the same system
 * tightly.
 *
 * Pinning (see tests/test_oracle_*.py): element matrices against closed forms,
 * invariants and 50-digit mpmath brute force; solves against dense/scipy
 * direct solves; MMA against a brute-force dual bisection; integer mesh/DOF/BC
 * indexing against hand-built meshes and the reference BC fixture facts.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

/* ------------------------------------------------------------------------- */
/* Synthetic cantilever generator (workload definition, not in the reference) */
/* ------------------------------------------------------------------------- */

/* SplitMix64 (Steele, Lea, Flood 2014): portable reproducible stream. */
static uint64_t sm64_next(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double sm64_u01(uint64_t *s) { return (double)(sm64_next(s) >> 11) * 0x1.0p-53; }

int64_t orc_box_nodes(int nx, int ny, int nz) { return (int64_t)(nx + 1) * (ny + 1) * (nz + 1); }
int64_t orc_box_elems(int etype, int nx, int ny, int nz) {
  int64_t h = (int64_t)nx * ny * nz;
  return etype == 4 ? 6 * h : h;
}

/* Node id = i + (nx+1)(j + (ny+1)k).  Interior nodes are jittered by
 * (2u-1)*jitter*h_a per axis, three draws per interior node in node order.
 * Hex local order is the reference's "diagonal numbering"
 * (SolveFem.cpp:352-360): (i,j,k),(i+1,j,k),(i,j+1,k),(i+1,j+1,k), then k+1.
 * Tet4: each hex split into 6 positively oriented tets around the 0-7 diagonal. */
int orc_gen_box(int etype, int nx, int ny, int nz, double lx, double ly, double lz,
                double jitter, uint64_t seed, double *xyz, int32_t *elem) {
  if (nx < 1 || ny < 1 || nz < 1 || (etype != 8 && etype != 4)) return -1;
  uint64_t s = seed;
  double hx = lx / nx, hy = ly / ny, hz = lz / nz;
  int64_t n = 0;
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++, n++) {
        double x = lx * (double)i / (double)nx;
        double y = ly * (double)j / (double)ny;
        double z = lz * (double)k / (double)nz;
        if (jitter != 0.0 && i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz) {
          x += (2.0 * sm64_u01(&s) - 1.0) * jitter * hx;
          y += (2.0 * sm64_u01(&s) - 1.0) * jitter * hy;
          z += (2.0 * sm64_u01(&s) - 1.0) * jitter * hz;
        }
        xyz[3 * n] = x; xyz[3 * n + 1] = y; xyz[3 * n + 2] = z;
      }
  static const int tets[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7},
                                 {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};
  int64_t e = 0;
  int64_t sx = nx + 1, sxy = (int64_t)(nx + 1) * (ny + 1);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int64_t b = i + sx * j + sxy * k;
        int32_t hv[8] = {(int32_t)b, (int32_t)(b + 1), (int32_t)(b + sx), (int32_t)(b + sx + 1),
                         (int32_t)(b + sxy), (int32_t)(b + sxy + 1), (int32_t)(b + sxy + sx),
                         (int32_t)(b + sxy + sx + 1)};
        if (etype == 8) {
          for (int a = 0; a < 8; a++) elem[8 * e + a] = hv[a];
          e++;
        } else {
          for (int t = 0; t < 6; t++, e++)
            for (int a = 0; a < 4; a++) elem[4 * e + a] = hv[tets[t][a]];
        }
      }
  return 0;
}

/* Cantilever BCs.  Fixed: all three DOFs of the i=0 nodes in ascending node
 * order as triples {3n,3n+1,3n+2} (GetBcs/GetBCs.m:46-51).  Load on the
 * y-DOF 3n+1 (GetBcs/GetBCs.m:67):
 *   kind 0: nodes i=nx, j=0, all k, value -1/(nz+1) each (bottom edge of x=Lx);
 *   kind 1: 3x3 node patch i=nx, j in {ny/2-1..ny/2+1}, k in {nz/2-1..nz/2+1}, -1/9;
 *   kind 2: eight load cases, each a 3x3 patch on x=Lx (-1/9 per node,
 *           unit total) centred at (j, k) = (1 + floor(u (ny-1)),
 *           1 + floor(u (nz-1))) with u from the SplitMix64 stream of seed 8
 *           (j then k, case by case); 72 entries, 9 per case. */
int orc_box_bcs(int nx, int ny, int nz, int load_kind, int32_t *fixed, int *nfixed,
                int32_t *ldofs, double *lvals, int *nload) {
  int64_t sx = nx + 1, sxy = (int64_t)(nx + 1) * (ny + 1);
  int nf = 0;
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++) {
      int64_t n = sx * j + sxy * k;
      fixed[nf++] = (int32_t)(3 * n);
      fixed[nf++] = (int32_t)(3 * n + 1);
      fixed[nf++] = (int32_t)(3 * n + 2);
    }
  *nfixed = nf;
  int nl = 0;
  if (load_kind == 0) {
    for (int k = 0; k <= nz; k++) {
      int64_t n = nx + sxy * k;
      ldofs[nl] = (int32_t)(3 * n + 1);
      lvals[nl++] = -1.0 / (double)(nz + 1);
    }
  } else if (load_kind == 1) {
    int jc = ny / 2, kc = nz / 2;
    for (int k = kc - 1; k <= kc + 1; k++)
      for (int j = jc - 1; j <= jc + 1; j++) {
        int64_t n = nx + sx * j + sxy * k;
        ldofs[nl] = (int32_t)(3 * n + 1);
        lvals[nl++] = -1.0 / 9.0;
      }
  } else {
    uint64_t st = 8;
    for (int lc = 0; lc < 8; lc++) {
      int jc = 1 + (int)floor(sm64_u01(&st) * (ny - 1));
      int kc = 1 + (int)floor(sm64_u01(&st) * (nz - 1));
      for (int k = kc - 1; k <= kc + 1; k++)
        for (int j = jc - 1; j <= jc + 1; j++) {
          int64_t n = nx + sx * j + sxy * k;
          ldofs[nl] = (int32_t)(3 * n + 1);
          lvals[nl++] = -1.0 / 9.0;
        }
    }
  }
  *nload = nl;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Element routines                                                           */
/* ------------------------------------------------------------------------- */

/* Truncated 3-point Gauss rule used by the reference for both elasticity and
 * Helmholtz (SolveFem.cpp:364-365; PDEFilter.cpp:195-196). */
static const double GP3[3] = {-0.775, 0.0, 0.775};
static const double GW3[3] = {0.556, 0.889, 0.556};

static double det3(const double J[3][3]) {
  return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
         J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}
static void inv3(const double J[3][3], double det, double Ji[3][3]) {
  double id = 1.0 / det;
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
}

/* Trilinear shape functions, diagonal numbering: node a sits at
 * (sx,sy,sz) = (a&1 ? +1:-1, a&2 ? +1:-1, a&4 ? +1:-1)
 * (shape-function comment SolveFem.cpp:352-360; derivative rows :390-395). */
static void hex8_shape(double xi, double eta, double zeta, double N[8], double dN[3][8]) {
  for (int a = 0; a < 8; a++) {
    double sx = (a & 1) ? 1.0 : -1.0, sy = (a & 2) ? 1.0 : -1.0, sz = (a & 4) ? 1.0 : -1.0;
    double fx = 1.0 + sx * xi, fy = 1.0 + sy * eta, fz = 1.0 + sz * zeta;
    N[a] = 0.125 * fx * fy * fz;
    dN[0][a] = 0.125 * sx * fy * fz;
    dN[1][a] = 0.125 * sy * fx * fz;
    dN[2][a] = 0.125 * sz * fx * fy;
  }
}

/* Isotropic E=1 elasticity matrix with engineering shear strains in the order
 * xy, yz, zx (SolveFem.cpp:370-376). */
static void dmat(double nu, double D[6][6]) {
  double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double mu = 1.0 / (2.0 * (1.0 + nu));
  memset(D, 0, 36 * sizeof(double));
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) D[i][j] = lam;
    D[i][i] = lam + 2.0 * mu;
    D[3 + i][3 + i] = mu;
  }
}

/* B (6 x 3*npe) over component-blocked columns [u1..un, v1..vn, w1..wn]
 * (SolveFem.cpp:401; tet4/Tet4Isoparametric.cpp:84). */
static void bmat(int npe, const double *dx, const double *dy, const double *dz, double *B) {
  int nd = 3 * npe;
  memset(B, 0, 6 * nd * sizeof(double));
  for (int a = 0; a < npe; a++) {
    B[0 * nd + a] = dx[a];
    B[1 * nd + npe + a] = dy[a];
    B[2 * nd + 2 * npe + a] = dz[a];
    B[3 * nd + a] = dy[a];
    B[3 * nd + npe + a] = dx[a];
    B[4 * nd + npe + a] = dz[a];
    B[4 * nd + 2 * npe + a] = dy[a];
    B[5 * nd + a] = dz[a];
    B[5 * nd + 2 * npe + a] = dx[a];
  }
}

/* KE += s * B^T D B */
static void add_btdb(int nd, const double *B, const double D[6][6], double s, double *KE) {
  double DB[6 * 24];
  for (int i = 0; i < 6; i++)
    for (int k = 0; k < nd; k++) {
      double acc = 0.0;
      for (int j = 0; j < 6; j++) acc += D[i][j] * B[j * nd + k];
      DB[i * nd + k] = acc;
    }
  for (int p = 0; p < nd; p++)
    for (int q = 0; q < nd; q++) {
      double acc = 0.0;
      for (int i = 0; i < 6; i++) acc += B[i * nd + p] * DB[i * nd + q];
      KE[p * nd + q] += s * acc;
    }
}

/* SolveFem::Hex8Isoparametric (SolveFem.cpp:339-408): 27-point truncated rule,
 * KE = sum W_i W_j W_k B^T D B detJ with signed detJ.  X is 8x3 row-major.
 * Returns the minimum detJ over the Gauss points. */
double orc_hex8_ke(const double *X, double nu, double *KE) {
  double D[6][6];
  dmat(nu, D);
  memset(KE, 0, 576 * sizeof(double));
  double mindet = 1e300;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        double N[8], dN[3][8], J[3][3], Ji[3][3];
        hex8_shape(GP3[i], GP3[j], GP3[k], N, dN);
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) {
            double acc = 0.0;
            for (int a = 0; a < 8; a++) acc += dN[r][a] * X[3 * a + c];
            J[r][c] = acc;
          }
        double dj = det3(J);
        if (dj < mindet) mindet = dj;
        inv3(J, dj, Ji);
        double dx[8], dy[8], dz[8];
        for (int a = 0; a < 8; a++) {
          dx[a] = Ji[0][0] * dN[0][a] + Ji[0][1] * dN[1][a] + Ji[0][2] * dN[2][a];
          dy[a] = Ji[1][0] * dN[0][a] + Ji[1][1] * dN[1][a] + Ji[1][2] * dN[2][a];
          dz[a] = Ji[2][0] * dN[0][a] + Ji[2][1] * dN[1][a] + Ji[2][2] * dN[2][a];
        }
        double B[6 * 24];
        bmat(8, dx, dy, dz, B);
        add_btdb(24, B, D, GW3[i] * GW3[j] * GW3[k] * dj, KE);
      }
  return mindet;
}

/* tet4/Tet4Isoparametric.cpp:1-41: constant-strain tet,
 * KE = (1/6) B^T D B sqrt(det(J J^T)).  Returns signed det J. */
double orc_tet4_ke(const double *X, double nu, double *KE) {
  double D[6][6];
  dmat(nu, D);
  memset(KE, 0, 144 * sizeof(double));
  double J[3][3], Ji[3][3], JJ[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) J[r][c] = X[3 * (r + 1) + c] - X[c];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) JJ[r][c] = J[r][0] * J[c][0] + J[r][1] * J[c][1] + J[r][2] * J[c][2];
  double dj = det3(J);
  double ad = sqrt(det3(JJ));
  inv3(J, dj, Ji);
  static const double dN[3][4] = {{-1, 1, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 0, 1}};
  double dx[4], dy[4], dz[4];
  for (int a = 0; a < 4; a++) {
    dx[a] = Ji[0][0] * dN[0][a] + Ji[0][1] * dN[1][a] + Ji[0][2] * dN[2][a];
    dy[a] = Ji[1][0] * dN[0][a] + Ji[1][1] * dN[1][a] + Ji[1][2] * dN[2][a];
    dz[a] = Ji[2][0] * dN[0][a] + Ji[2][1] * dN[1][a] + Ji[2][2] * dN[2][a];
  }
  double B[6 * 12];
  bmat(4, dx, dy, dz, B);
  add_btdb(12, B, D, ad / 6.0, KE);
  return dj;
}

static double tri6(const double *a, const double *b, const double *c, const double *o) {
  /* (o-a) . ((o-b) x (o-c)) / 6 */
  double u[3] = {o[0] - a[0], o[1] - a[1], o[2] - a[2]};
  double v[3] = {o[0] - b[0], o[1] - b[1], o[2] - b[2]};
  double w[3] = {o[0] - c[0], o[1] - c[1], o[2] - c[2]};
  double cx = v[1] * w[2] - v[2] * w[1], cy = v[2] * w[0] - v[0] * w[2], cz = v[0] * w[1] - v[1] * w[0];
  return (u[0] * cx + u[1] * cy + u[2] * cz) / 6.0;
}

/* SolveFem::ComputeVolOfHex (SolveFem.cpp:423-467): sum of |signed volume| of
 * the five tets 6-124, 7-143, 1-657, 4-678, 6-147 (1-based apex-base). */
double orc_hex8_vol(const double *X) {
  const double *c[8];
  for (int a = 0; a < 8; a++) c[a] = X + 3 * a;
  double p1 = tri6(c[0], c[1], c[3], c[5]);
  double p2 = tri6(c[0], c[3], c[2], c[6]);
  double p3 = tri6(c[5], c[4], c[6], c[0]);
  double p4 = tri6(c[5], c[6], c[7], c[3]);
  double p5 = tri6(c[0], c[3], c[6], c[5]);
  return fabs(p1) + fabs(p2) + fabs(p3) + fabs(p4) + fabs(p5);
}

/* Tet4 volume |det J|/6 (the reference has no tet volume routine). */
double orc_tet4_vol(const double *X) {
  double J[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) J[r][c] = X[3 * (r + 1) + c] - X[c];
  return fabs(det3(J)) / 6.0;
}

/* PDEFilter::KeHel (PDEFilter.cpp:185-240): r = rmin/(2 sqrt 3);
 * KF = sum W (r^2 Dn^T Dn + N^T N) detJ, T = sum W N detJ, truncated rule. */
void orc_hex8_kehel(const double *X, double rmin, double *KF, double *T) {
  double r = 1.0 / (2.0 * sqrt(3.0)) * rmin;
  double r2 = r * r;
  memset(KF, 0, 64 * sizeof(double));
  memset(T, 0, 8 * sizeof(double));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        double N[8], dN[3][8], J[3][3], Ji[3][3];
        hex8_shape(GP3[i], GP3[j], GP3[k], N, dN);
        for (int rr = 0; rr < 3; rr++)
          for (int c = 0; c < 3; c++) {
            double acc = 0.0;
            for (int a = 0; a < 8; a++) acc += dN[rr][a] * X[3 * a + c];
            J[rr][c] = acc;
          }
        double dj = det3(J);
        inv3(J, dj, Ji);
        double G[3][8];
        for (int a = 0; a < 8; a++)
          for (int rr = 0; rr < 3; rr++)
            G[rr][a] = Ji[rr][0] * dN[0][a] + Ji[rr][1] * dN[1][a] + Ji[rr][2] * dN[2][a];
        double w = GW3[i] * GW3[j] * GW3[k] * dj;
        for (int a = 0; a < 8; a++) {
          for (int b = 0; b < 8; b++)
            KF[a * 8 + b] += w * (r2 * (G[0][a] * G[0][b] + G[1][a] * G[1][b] + G[2][a] * G[2][b]) + N[a] * N[b]);
          T[a] += w * N[a];
        }
      }
}

/* tet4/KeHel.cpp:1-47: 4-point Hammer rule, weight 1/24, signed det J. */
void orc_tet4_kehel(const double *X, double rmin, double *KF, double *T) {
  double r = 1.0 / (2.0 * sqrt(3.0)) * rmin;
  double r2 = r * r;
  double a = (5.0 + 3.0 * sqrt(5.0)) / 20.0, b = (5.0 - sqrt(5.0)) / 20.0;
  double G[4][3] = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b}};
  double J[3][3], Ji[3][3];
  for (int rr = 0; rr < 3; rr++)
    for (int c = 0; c < 3; c++) J[rr][c] = X[3 * (rr + 1) + c] - X[c];
  double dj = det3(J);
  inv3(J, dj, Ji);
  static const double dN[3][4] = {{-1, 1, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 0, 1}};
  double Gd[3][4];
  for (int q = 0; q < 4; q++)
    for (int rr = 0; rr < 3; rr++)
      Gd[rr][q] = Ji[rr][0] * dN[0][q] + Ji[rr][1] * dN[1][q] + Ji[rr][2] * dN[2][q];
  memset(KF, 0, 16 * sizeof(double));
  memset(T, 0, 4 * sizeof(double));
  for (int g = 0; g < 4; g++) {
    double s = G[g][0], t = G[g][1], v = G[g][2];
    double N[4] = {1.0 - s - t - v, s, t, v};
    double w = 1.0 / 24.0 * dj;
    for (int p = 0; p < 4; p++) {
      for (int q = 0; q < 4; q++)
        KF[p * 4 + q] += w * (r2 * (Gd[0][p] * Gd[0][q] + Gd[1][p] * Gd[1][q] + Gd[2][p] * Gd[2][q]) + N[p] * N[q]);
      T[p] += w * N[p];
    }
  }
}

/* SolveFem::GetDofCoord (SolveFem.cpp:38-57): component-blocked local order,
 * column k -> 3*node[k mod npe] + k/npe. */
void orc_dofmap(const int32_t *elem, int64_t ne, int npe, int32_t *dof) {
  int nd = 3 * npe;
  for (int64_t e = 0; e < ne; e++)
    for (int k = 0; k < nd; k++) dof[e * nd + k] = 3 * elem[e * npe + (k % npe)] + k / npe;
}

/* ------------------------------------------------------------------------- */
/* Sparse helpers                                                             */
/* ------------------------------------------------------------------------- */


static void csr_free(csr_t *A) {
  free(A->rp); free(A->ci); free(A->v);
  memset(A, 0, sizeof(*A));
}

static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* Node adjacency (nodes sharing an element, including self), sorted. */
static void node_graph(int64_t nn, int64_t ne, int npe, const int32_t *elem, int64_t **rp_out, int32_t **ci_out) {
  int64_t *cnt = calloc(nn + 1, sizeof(int64_t));
  for (int64_t e = 0; e < ne; e++)
    for (int a = 0; a < npe; a++) cnt[elem[e * npe + a] + 1] += npe;
  for (int64_t i = 0; i < nn; i++) cnt[i + 1] += cnt[i];
  int32_t *tmp = malloc(cnt[nn] * sizeof(int32_t));
  int64_t *pos = malloc(nn * sizeof(int64_t));
  for (int64_t i = 0; i < nn; i++) pos[i] = cnt[i];
  for (int64_t e = 0; e < ne; e++)
    for (int a = 0; a < npe; a++) {
      int32_t n = elem[e * npe + a];
      for (int b = 0; b < npe; b++) tmp[pos[n]++] = elem[e * npe + b];
    }
  int64_t *rp = malloc((nn + 1) * sizeof(int64_t));
  int32_t *ci = malloc(cnt[nn] * sizeof(int32_t));
  rp[0] = 0;
  int64_t w = 0;
  for (int64_t i = 0; i < nn; i++) {
    int64_t s = cnt[i], t = cnt[i + 1];
    qsort(tmp + s, t - s, sizeof(int32_t), cmp_i32);
    for (int64_t k = s; k < t; k++)
      if (k == s || tmp[k] != tmp[k - 1]) ci[w++] = tmp[k];
    rp[i + 1] = w;
  }
  free(cnt); free(tmp); free(pos);
  *rp_out = rp;
  *ci_out = ci;
}

static int64_t csr_find(const csr_t *A, int64_t row, int32_t col) {
  int64_t lo = A->rp[row], hi = A->rp[row + 1] - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) >> 1;
    if (A->ci[mid] == col) return mid;
    if (A->ci[mid] < col) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

static void csr_mv(const csr_t *A, const double *x, double *y) {
  for (int64_t i = 0; i < A->n; i++) {
    double acc = 0.0;
    for (int64_t k = A->rp[i]; k < A->rp[i + 1]; k++) acc += A->v[k] * x[A->ci[k]];
    y[i] = acc;
  }
}

static double dot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* ------------------------------------------------------------------------- */
/* Oracle context                                                             */
/* ------------------------------------------------------------------------- */


void orc_default_params(orc_params *P) {
  P->maxItr = 200; P->rmin = 0.01; P->penal = 3.0; P->Emin = 1.0e-9; P->Emax = 1.0e6;
  P->movlim = 0.2; P->Xmin = 0.01; P->Xmax = 1.0; P->volfrac = 0.12; P->x0 = 0.5;
  P->beta0 = 1.0; P->betamax = 8.0; P->proj = 0; P->m = 1; P->nu = 0.3;
  P->rtol_el = 1.0e-5; P->rtol_f = 1.0e-5; P->maxit_el = 100000; P->maxit_f = 10000;
  P->ke_cache = 1;
}
size_t orc_params_size(void) { return sizeof(orc_params); }

static void elem_coords(const orc_ctx *C, int64_t e, double *X) {
  for (int a = 0; a < C->npe; a++) {
    int32_t n = C->elem[e * C->npe + a];
    X[3 * a] = C->xyz[3 * n]; X[3 * a + 1] = C->xyz[3 * n + 1]; X[3 * a + 2] = C->xyz[3 * n + 2];
  }
}
static double elem_ke(const orc_ctx *C, int64_t e, double *KE) {
  if (C->KEc) { memcpy(KE, C->KEc + e * C->nd * C->nd, C->nd * C->nd * sizeof(double)); return 1.0; }
  double X[24];
  elem_coords(C, e, X);
  return C->npe == 8 ? orc_hex8_ke(X, C->P.nu, KE) : orc_tet4_ke(X, C->P.nu, KE);
}

void orc_destroy(orc_ctx *C);

orc_ctx *orc_create(int npe, int64_t nn, int64_t ne, const int32_t *elem, const double *xyz,
                    int nfixed, const int32_t *fixed, int nlc, int nload, const int32_t *ldofs,
                    const double *lvals, const orc_params *P) {
  if (npe != 8 && npe != 4) return NULL;
  if (P->m < 1 || P->m > ORC_MAXM) return NULL;
  orc_ctx *C = calloc(1, sizeof(orc_ctx));
  C->npe = npe; C->nd = 3 * npe; C->nn = nn; C->ne = ne; C->ndof = 3 * nn;
  C->P = *P;
  C->elem = malloc(ne * npe * sizeof(int32_t)); memcpy(C->elem, elem, ne * npe * sizeof(int32_t));
  C->xyz = malloc(nn * 3 * sizeof(double)); memcpy(C->xyz, xyz, nn * 3 * sizeof(double));
  C->nfixed = nfixed;
  C->fixed = malloc((nfixed + 1) * sizeof(int32_t)); memcpy(C->fixed, fixed, nfixed * sizeof(int32_t));
  C->nlc = nlc; C->nload = nload;
  C->ldofs = malloc((nlc * nload + 1) * sizeof(int32_t)); memcpy(C->ldofs, ldofs, nlc * nload * sizeof(int32_t));
  C->lvals = malloc((nlc * nload + 1) * sizeof(double)); memcpy(C->lvals, lvals, nlc * nload * sizeof(double));
  C->isfixed = calloc(C->ndof, 1);
  for (int i = 0; i < nfixed; i++) C->isfixed[fixed[i]] = 1;
  for (int64_t e = 0; e < ne; e++)
    for (int a = 0; a < npe; a++)
      if (elem[e * npe + a] < 0 || elem[e * npe + a] >= nn) { orc_destroy(C); return NULL; }

  /* SolveFem ctor: volumes and total volume (SolveFem.cpp:59-89) */
  C->vol = malloc(ne * sizeof(double));
  C->TV = 0.0;
  for (int64_t e = 0; e < ne; e++) {
    double X[24];
    elem_coords(C, e, X);
    C->vol[e] = npe == 8 ? orc_hex8_vol(X) : orc_tet4_vol(X);
    C->TV += C->vol[e];
  }
  if (P->ke_cache) {
    C->KEc = NULL;
    double *KEc = malloc(ne * C->nd * C->nd * sizeof(double));
    for (int64_t e = 0; e < ne; e++) {
      double X[24];
      elem_coords(C, e, X);
      if (npe == 8) orc_hex8_ke(X, P->nu, KEc + e * 576); else orc_tet4_ke(X, P->nu, KEc + e * 144);
    }
    C->KEc = KEc;
  }

  /* sparsity: node graph expanded to 3x3 blocks */
  int64_t *nrp; int32_t *nci;
  node_graph(nn, ne, npe, elem, &nrp, &nci);
  int64_t nnzb = nrp[nn];
  C->K.n = 3 * nn;
  C->K.rp = malloc((3 * nn + 1) * sizeof(int64_t));
  C->K.ci = malloc(9 * nnzb * sizeof(int32_t));
  C->K.v = calloc(9 * nnzb, sizeof(double));
  C->K.rp[0] = 0;
  for (int64_t i = 0; i < nn; i++)
    for (int c = 0; c < 3; c++) {
      int64_t r = 3 * i + c, w = C->K.rp[r];
      for (int64_t k = nrp[i]; k < nrp[i + 1]; k++)
        for (int cc = 0; cc < 3; cc++) C->K.ci[w++] = 3 * nci[k] + cc;
      C->K.rp[r + 1] = w;
    }
  /* PDEFilter::PrepareFilter (PDEFilter.cpp:36-140): KF assembled once. */
  C->KF.n = nn; C->KF.rp = nrp; C->KF.ci = nci; C->KF.v = calloc(nnzb, sizeof(double));
  C->Tel = malloc(ne * npe * sizeof(double));
  for (int64_t e = 0; e < ne; e++) {
    double X[24], KFe[64], Te[8];
    elem_coords(C, e, X);
    if (npe == 8) orc_hex8_kehel(X, P->rmin, KFe, Te); else orc_tet4_kehel(X, P->rmin, KFe, Te);
    for (int a = 0; a < npe; a++) {
      C->Tel[e * npe + a] = Te[a];
      for (int b2 = 0; b2 < npe; b2++)
        C->KF.v[csr_find(&C->KF, elem[e * npe + a], elem[e * npe + b2])] += KFe[a * npe + b2];
    }
  }
  C->KFdiag = malloc(nn * sizeof(double));
  for (int64_t i = 0; i < nn; i++) C->KFdiag[i] = C->KF.v[csr_find(&C->KF, i, (int32_t)i)];

  C->U = calloc(C->ndof, sizeof(double));
  C->F = calloc(C->ndof, sizeof(double));
  C->TX = calloc(nn, sizeof(double));
  C->RX = calloc(nn, sizeof(double));
  int m = P->m;
  C->X = malloc(ne * sizeof(double)); C->XFiled = malloc(ne * sizeof(double));
  C->XPhys = malloc(ne * sizeof(double)); C->Xold = malloc(ne * sizeof(double));
  C->dfdx = calloc(ne, sizeof(double)); C->dgdx = calloc(ne * m, sizeof(double));
  C->xmin = calloc(ne, sizeof(double)); C->xmax = calloc(ne, sizeof(double));
  /* OptPara.cpp:150-153 */
  for (int64_t e = 0; e < ne; e++) C->X[e] = C->XFiled[e] = C->XPhys[e] = C->Xold[e] = P->volfrac;
  /* MMA(n, m, x, a, c, d) with a=0, c=1000, d=0 (OptPara.cpp:166-174; MMA.cpp:178-224) */
  C->mk = 0; C->asyminit = 0.5; C->asymdec = 0.7; C->asyminc = 1.2; C->robust = 0; C->conmod = 0;
  for (int j = 0; j < m; j++) { C->a[j] = 0.0; C->c[j] = 1000.0; C->d[j] = 0.0; }
  C->L = calloc(ne, sizeof(double)); C->Uas = calloc(ne, sizeof(double));
  C->alpha = calloc(ne, sizeof(double)); C->betav = calloc(ne, sizeof(double));
  C->p0 = calloc(ne, sizeof(double)); C->q0 = calloc(ne, sizeof(double));
  C->pij = calloc(ne * m, sizeof(double)); C->qij = calloc(ne * m, sizeof(double));
  C->xo1 = calloc(ne, sizeof(double)); C->xo2 = calloc(ne, sizeof(double));
  return C;
}

void orc_destroy(orc_ctx *C) {
  if (!C) return;
  free(C->elem); free(C->xyz); free(C->fixed); free(C->ldofs); free(C->lvals); free(C->isfixed);
  free(C->vol); free(C->KEc); free(C->Tel); csr_free(&C->K); csr_free(&C->KF); free(C->KFdiag);
  free(C->U); free(C->F); free(C->TX); free(C->RX);
  free(C->X); free(C->XFiled); free(C->XPhys); free(C->Xold); free(C->dfdx); free(C->dgdx);
  free(C->xmin); free(C->xmax);
  free(C->L); free(C->Uas); free(C->alpha); free(C->betav); free(C->p0); free(C->q0);
  free(C->pij); free(C->qij); free(C->xo1); free(C->xo2); free(C->Ucase);
  free(C);
}

/* accessors used by the ctypes wrapper */
double *orc_vec(orc_ctx *C, int which) {
  switch (which) {
    case 0: return C->X; case 1: return C->XFiled; case 2: return C->XPhys; case 3: return C->Xold;
    case 4: return C->dfdx; case 5: return C->dgdx; case 6: return C->xmin; case 7: return C->xmax;
    case 8: return C->vol; case 9: return C->U; case 10: return C->RX; case 11: return C->F;
    case 12: return C->L; case 13: return C->Uas; case 14: return C->xo1; case 15: return C->xo2;
    case 16: return C->Tel; case 17: return C->KFdiag;
    case 18: return C->alpha; case 19: return C->betav; case 20: return C->p0; case 21: return C->q0;
    case 22: return C->pij; case 23: return C->qij;
  }
  return NULL;
}
double orc_total_volume(orc_ctx *C) { return C->TV; }
void orc_diag(orc_ctx *C, int *it_el, int *it_f, double *relres_el, int *dip_steps) {
  *it_el = C->it_el; *it_f = C->it_f; *relres_el = C->relres_el; *dip_steps = C->dip_steps;
}
void orc_set_params(orc_ctx *C, const orc_params *P) { C->P = *P; }
int orc_mma_k(orc_ctx *C) { return C->mk; }
double orc_mma_lam(orc_ctx *C, int j) { return C->lam[j]; }


/* ------------------------------------------------------------------------- */
/* SolveFem                                                                   */
/* ------------------------------------------------------------------------- */

static double simp_E(const orc_params *P, double x) { return P->Emin + pow(x, P->penal) * (P->Emax - P->Emin); }

/* SolveFem::AssembleFem (SolveFem.cpp:267-337): K = sum E(x_e) KE_e, loads
 * INSERTed per case, MatZeroRowsColumns(fixed, diag=1), then F[fixed]=0. */
void orc_assemble(orc_ctx *C, const double *xphys) {
  int nd = C->nd, npe = C->npe;
  memset(C->K.v, 0, C->K.rp[C->K.n] * sizeof(double));
  memset(C->F, 0, C->ndof * sizeof(double));
  memset(C->U, 0, C->ndof * sizeof(double));
  for (int lc = 0; lc < C->nlc; lc++)
    for (int i = 0; i < C->nload; i++) C->F[C->ldofs[lc * C->nload + i]] = C->lvals[lc * C->nload + i];
  double KE[576];
  for (int64_t e = 0; e < C->ne; e++) {
    elem_ke(C, e, KE);
    double dens = simp_E(&C->P, xphys[e]);
    int32_t dof[24];
    for (int k = 0; k < nd; k++) dof[k] = 3 * C->elem[e * npe + (k % npe)] + k / npe;
    for (int p = 0; p < nd; p++)
      for (int q = 0; q < nd; q++) C->K.v[csr_find(&C->K, dof[p], dof[q])] += KE[p * nd + q] * dens;
  }
  for (int64_t r = 0; r < C->K.n; r++)
    for (int64_t k = C->K.rp[r]; k < C->K.rp[r + 1]; k++) {
      int32_t c = C->K.ci[k];
      if (C->isfixed[r] || C->isfixed[c]) C->K.v[k] = (r == c) ? 1.0 : 0.0;
    }
  for (int i = 0; i < C->nfixed; i++) C->F[C->fixed[i]] = 0.0;
}

/* Apply the assembled Dirichlet-eliminated K (for parity tests). */
void orc_apply_K(orc_ctx *C, const double *u, double *y) { csr_mv(&C->K, u, y); }

/* Jacobi-PCG on the eliminated system, x0 = 0 (U zeroed, SolveFem.cpp:273),
 * stop when ||F - K U||_2 <= rtol ||F||_2 (right-preconditioned FGMRES
 * semantics of SolveFem.cpp:126-139).  The test is on the true residual:
 * when the recursively updated residual passes it, F - K U is formed and,
 * if that one has drifted above the tolerance, CG restarts from it. */
static int pcg_el(orc_ctx *C) {
  int64_t n = C->ndof;
  double *r = malloc(n * sizeof(double)), *z = malloc(n * sizeof(double));
  double *p = malloc(n * sizeof(double)), *q = malloc(n * sizeof(double)), *dinv = malloc(n * sizeof(double));
  for (int64_t i = 0; i < n; i++) {
    dinv[i] = 1.0 / C->K.v[csr_find(&C->K, i, (int32_t)i)];
    C->U[i] = 0.0; r[i] = C->F[i]; z[i] = dinv[i] * r[i]; p[i] = z[i];
  }
  double bn = sqrt(dot(n, C->F, C->F));
  double rz = dot(n, r, z), rn = bn, tn = bn;
  const double tol = C->P.rtol_el * bn;
  int it = 0;
  if (bn == 0.0) { C->it_el = 0; C->relres_el = 0.0; free(r); free(z); free(p); free(q); free(dinv); return 0; }
  const int progress = getenv("ORC_PROGRESS") != NULL;
  while (it < C->P.maxit_el) {
    if (rn <= tol) {
      csr_mv(&C->K, C->U, q);
      for (int64_t i = 0; i < n; i++) r[i] = C->F[i] - q[i];
      tn = sqrt(dot(n, r, r));
      if (tn <= tol) break;
      /* restart from the true residual */
      rn = tn;
      for (int64_t i = 0; i < n; i++) { z[i] = dinv[i] * r[i]; p[i] = z[i]; }
      rz = dot(n, r, z);
    }
    if (progress && it % 500 == 0) { fprintf(stderr, "[oracle pcg] it %d relres %.3e\n", it, rn / bn); fflush(stderr); }
    csr_mv(&C->K, p, q);
    double al = rz / dot(n, p, q);
    for (int64_t i = 0; i < n; i++) { C->U[i] += al * p[i]; r[i] -= al * q[i]; }
    rn = sqrt(dot(n, r, r));
    it++;
    for (int64_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
    double rz2 = dot(n, r, z);
    double be = rz2 / rz;
    rz = rz2;
    for (int64_t i = 0; i < n; i++) p[i] = z[i] + be * p[i];
  }
  /* true residual diagnostic (SolveFem.cpp:142-152) */
  csr_mv(&C->K, C->U, q);
  for (int64_t i = 0; i < n; i++) q[i] = C->F[i] - q[i];
  tn = sqrt(dot(n, q, q));
  C->relres_el = tn / bn;
  C->it_el = it;
  free(r); free(z); free(p); free(q); free(dinv);
  return tn <= tol ? 0 : 1;
}

/* SolveFem::ComputeObjectiveAndSensitivities (SolveFem.cpp:159-265):
 * uKu_e = u_e^T KE_e u_e, fx = sum E(x_e) uKu_e, df_e = -p x^(p-1)(Emax-Emin) uKu_e,
 * gx = sum x_e V_e / TV - volfrac, dg_e = V_e / TV. */
/* Multi-load-case form (an extension; the reference superposes its load
 * cases into one right-hand side, SolveFem.cpp:277-283): every case is
 * solved on its own and the compliance and its sensitivities are summed. */
static int fem_obj_sens_cases(orc_ctx *C, const double *xphys, double *fx, double *gx, double *dfdx,
                              double *dgdx) {
  const orc_params *P = &C->P;
  int nd = C->nd, npe = C->npe, rc = 0;
  double KE[576];
  if (!C->Ucase) C->Ucase = calloc((size_t)C->nlc * C->ndof, sizeof(double));
  orc_assemble(C, xphys);
  for (int64_t e = 0; e < C->ne; e++) dfdx[e] = 0.0;
  double f = 0.0;
  for (int lc = 0; lc < C->nlc; lc++) {
    memset(C->F, 0, C->ndof * sizeof(double));
    for (int i = 0; i < C->nload; i++) C->F[C->ldofs[lc * C->nload + i]] = C->lvals[lc * C->nload + i];
    for (int i = 0; i < C->nfixed; i++) C->F[C->fixed[i]] = 0.0;
    int r = pcg_el(C);
    if (r) rc = r;
    memcpy(C->Ucase + (size_t)lc * C->ndof, C->U, C->ndof * sizeof(double));
    for (int64_t e = 0; e < C->ne; e++) {
      elem_ke(C, e, KE);
      double ue[24];
      for (int k = 0; k < nd; k++) ue[k] = C->U[3 * C->elem[e * npe + (k % npe)] + k / npe];
      double uKu = 0.0;
      for (int p = 0; p < nd; p++)
        for (int q = 0; q < nd; q++) uKu += ue[p] * KE[p * nd + q] * ue[q];
      f += simp_E(P, xphys[e]) * uKu;
      dfdx[e] += -1.0 * P->penal * pow(xphys[e], P->penal - 1) * (P->Emax - P->Emin) * uKu;
    }
  }
  double g = 0.0;
  for (int64_t e = 0; e < C->ne; e++) {
    g += xphys[e] * C->vol[e];
    dgdx[e] = C->vol[e] / C->TV;
  }
  fx[0] = f;
  gx[0] = g / C->TV - P->volfrac;
  return rc;
}
void orc_set_load_separate(orc_ctx *C, int on) { C->load_separate = on; }
const double *orc_case_u(orc_ctx *C, int lc) { return C->Ucase ? C->Ucase + (size_t)lc * C->ndof : NULL; }

int orc_fem_obj_sens(orc_ctx *C, const double *xphys, double *fx, double *gx, double *dfdx, double *dgdx) {
  if (C->load_separate && C->nlc > 1) return fem_obj_sens_cases(C, xphys, fx, gx, dfdx, dgdx);
  orc_assemble(C, xphys);
  int rc = pcg_el(C);
  int nd = C->nd, npe = C->npe;
  double KE[576];
  const orc_params *P = &C->P;
  double f = 0.0, g = 0.0;
  for (int64_t e = 0; e < C->ne; e++) {
    elem_ke(C, e, KE);
    double ue[24];
    for (int k = 0; k < nd; k++) ue[k] = C->U[3 * C->elem[e * npe + (k % npe)] + k / npe];
    double uKu = 0.0;
    for (int p = 0; p < nd; p++)
      for (int q = 0; q < nd; q++) uKu += ue[p] * KE[p * nd + q] * ue[q];
    f += (P->Emin + pow(xphys[e], P->penal) * (P->Emax - P->Emin)) * uKu;
    dfdx[e] = -1.0 * P->penal * pow(xphys[e], P->penal - 1) * (P->Emax - P->Emin) * uKu;
    g += xphys[e] * C->vol[e];
    dgdx[e] = C->vol[e] / C->TV;
  }
  fx[0] = f;
  gx[0] = g / C->TV - P->volfrac;
  /* the reference fills only constraint 0 (SolveFem.cpp:234-249) */
  return rc;
}

/* ------------------------------------------------------------------------- */
/* PDEFilter                                                                  */
/* ------------------------------------------------------------------------- */

/* PETSc KSPCG + PCJACOBI semantics (PDEFilter.cpp:117-137): nonzero initial
 * guess (RX of the previous solve), preconditioned residual norm, relative to
 * ||D^-1 b|| (KSPConvergedDefault with a nonzero guess), atol 1e-50,
 * dtol 1e5, maxit 1e4. */
static int cg_filter(orc_ctx *C, const double *b, double *x) {
  int64_t n = C->nn;
  double *r = malloc(n * sizeof(double)), *z = malloc(n * sizeof(double));
  double *p = malloc(n * sizeof(double)), *q = malloc(n * sizeof(double));
  csr_mv(&C->KF, x, q);
  double snorm = 0.0;
  for (int64_t i = 0; i < n; i++) {
    double zb = b[i] / C->KFdiag[i];
    snorm += zb * zb;
    r[i] = b[i] - q[i];
    z[i] = r[i] / C->KFdiag[i];
  }
  snorm = sqrt(snorm);
  double dp = sqrt(dot(n, z, z));
  if (snorm == 0.0) snorm = dp;
  double ttol = C->P.rtol_f * snorm;
  if (ttol < 1e-50) ttol = 1e-50;
  int it = 0, rc = 0;
  double beta = dot(n, z, r), betaold = 0.0;
  if (!(dp <= ttol)) {
    for (it = 0; it < C->P.maxit_f;) {
      if (it == 0) memcpy(p, z, n * sizeof(double));
      else { double bb = beta / betaold; for (int64_t i = 0; i < n; i++) p[i] = z[i] + bb * p[i]; }
      csr_mv(&C->KF, p, q);
      double al = beta / dot(n, p, q);
      for (int64_t i = 0; i < n; i++) { x[i] += al * p[i]; r[i] -= al * q[i]; z[i] = r[i] / C->KFdiag[i]; }
      dp = sqrt(dot(n, z, z));
      it++;
      if (dp <= ttol) break;
      if (dp >= 1e5 * snorm) { rc = 2; break; }
      betaold = beta;
      beta = dot(n, z, r);
    }
    if (it >= C->P.maxit_f && !(dp <= ttol)) rc = 1;
  }
  C->it_f = it;
  free(r); free(z); free(p); free(q);
  return rc;
}

/* out_e = (1/npe) sum_{a} RX[node_a(e)]   (TFX^T RX, PDEFilter.cpp:97,149) */
static void tfx_t(const orc_ctx *C, const double *RX, double *out) {
  int npe = C->npe;
  double w = 1.0 / (double)npe;
  for (int64_t e = 0; e < C->ne; e++) {
    double s = 0.0;
    for (int a = 0; a < npe; a++) s += w * RX[C->elem[e * npe + a]];
    out[e] = s;
  }
}
/* TX = TF x : TX_n = sum_{e,a: node_a(e)=n} T_e[a] x_e  (PDEFilter.cpp:106,147) */
static void tf_mult(const orc_ctx *C, const double *x, double *TX) {
  memset(TX, 0, C->nn * sizeof(double));
  for (int64_t e = 0; e < C->ne; e++)
    for (int a = 0; a < C->npe; a++) TX[C->elem[e * C->npe + a]] += C->Tel[e * C->npe + a] * x[e];
}

/* PDEFilter::Filter (PDEFilter.cpp:142-165) */
int orc_filter(orc_ctx *C, const double *x, double *xf) {
  tf_mult(C, x, C->TX);
  int rc = cg_filter(C, C->TX, C->RX);
  tfx_t(C, C->RX, xf);
  for (int64_t e = 0; e < C->ne; e++) {
    if (xf[e] < C->P.Xmin) xf[e] = C->P.Xmin;
    else if (xf[e] > 1.0) xf[e] = 1.0;
  }
  return rc;
}

/* PDEFilter::ComputeFilteredSens (PDEFilter.cpp:167-183): forward operator in
 * place on dfdx and each dgdx[i], no clamp. */
int orc_filter_sens(orc_ctx *C, double *dfdx, double *dgdx) {
  int rc = 0;
  tf_mult(C, dfdx, C->TX);
  rc |= cg_filter(C, C->TX, C->RX);
  tfx_t(C, C->RX, dfdx);
  for (int j = 0; j < C->P.m; j++) {
    tf_mult(C, dgdx + j * C->ne, C->TX);
    rc |= cg_filter(C, C->TX, C->RX);
    tfx_t(C, C->RX, dgdx + j * C->ne);
  }
  return rc;
}

/* PDEFilter::HeavisideFilter (PDEFilter.cpp:242-293) */
void orc_heaviside(orc_ctx *C, const double *xf, double *xp, int itr) {
  const orc_params *P = &C->P;
  if (!P->proj) { memcpy(xp, xf, C->ne * sizeof(double)); return; }
  if (itr < 30) C->beta = P->beta0;
  else if (itr % 10 == 0) C->beta = P->betamax < C->btemp + 1 ? P->betamax : C->btemp + 1;
  double bt = C->beta;
  for (int64_t i = 0; i < C->ne; i++) {
    xp[i] = (tanh(bt * P->x0) + tanh(bt * (xf[i] - P->x0))) / (tanh(bt * P->x0) + tanh(bt * (1 - P->x0)));
    if (xp[i] < P->Xmin) xp[i] = P->Xmin;
  }
  C->btemp = bt;
}

/* Projection sharpness beta of the last HeavisideFilter call (PDEFilter.h:51). */
double orc_beta(orc_ctx *C) { return C->beta; }

/* PDEFilter::ComputeProjectedSens (PDEFilter.cpp:295-341) */
void orc_projected_sens(orc_ctx *C, const double *xf, double *dfdx, double *dgdx) {
  const orc_params *P = &C->P;
  double bt = C->beta;
  for (int64_t i = 0; i < C->ne; i++) {
    double th = tanh(bt * (xf[i] - P->x0));
    double dx = bt * (1.0 - th * th) / (tanh(bt * P->x0) + tanh(bt * (1 - P->x0)));
    dfdx[i] *= dx;
    for (int j = 0; j < P->m; j++) dgdx[j * C->ne + i] *= dx;
  }
}

/* ------------------------------------------------------------------------- */
/* MMA (Svanberg 1987, dual interior point of Aage et al. 2014)               */
/* ------------------------------------------------------------------------- */

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* MMA::SetOuterMovelimit (MMA.cpp:370-389) */
void orc_mma_movelimit(orc_ctx *C, const double *x, double *xmin, double *xmax) {
  const orc_params *P = &C->P;
  for (int64_t i = 0; i < C->ne; i++) {
    xmax[i] = dmin(P->Xmax, x[i] + P->movlim);
    xmin[i] = dmax(P->Xmin, x[i] - P->movlim);
  }
}

/* MMA::DesignChange (MMA.cpp:391-410): inf-norm of x - xold; xold := x */
double orc_design_change(orc_ctx *C, const double *x, double *xold) {
  double ch = 0.0;
  for (int64_t i = 0; i < C->ne; i++) {
    ch = dmax(ch, fabs(x[i] - xold[i]));
    xold[i] = x[i];
  }
  return ch;
}

/* x(lambda): closed-form minimiser of the separable Lagrangian, clamped to
 * [alpha, beta]; also y(lambda), z(lambda) (MMA.cpp:678-728). */
static void mma_xyz(orc_ctx *C, double *x) {
  int m = C->P.m;
  int64_t n = C->ne;
  double lamai = 0.0;
  for (int j = 0; j < m; j++) {
    if (C->lam[j] < 0.0) C->lam[j] = 0.0;
    C->y[j] = dmax(0.0, C->lam[j] - C->c[j]);
    lamai += C->lam[j] * C->a[j];
  }
  C->z = dmax(0.0, 10.0 * (lamai - 1.0));
  for (int64_t i = 0; i < n; i++) {
    double pl = C->p0[i], ql = C->q0[i];
    for (int j = 0; j < m; j++) { pl += C->pij[j * n + i] * C->lam[j]; ql += C->qij[j * n + i] * C->lam[j]; }
    double xi = (sqrt(pl) * C->L[i] + sqrt(ql) * C->Uas[i]) / (sqrt(pl) + sqrt(ql));
    if (xi < C->alpha[i]) xi = C->alpha[i];
    if (xi > C->betav[i]) xi = C->betav[i];
    x[i] = xi;
  }
}

/* sum_i pij/(U-x) + qij/(x-L) for each constraint j (MMA.cpp:741-746, 901-907) */
static void mma_gsum(orc_ctx *C, const double *x, double *out) {
  int m = C->P.m;
  int64_t n = C->ne;
  for (int j = 0; j < m; j++) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += C->pij[j * n + i] / (C->Uas[i] - x[i]) + C->qij[j * n + i] / (x[i] - C->L[i]);
    out[j] = s;
  }
}

/* Dual Hessian (MMA.cpp:767-868) */
static void mma_hess(orc_ctx *C, const double *x) {
  int m = C->P.m;
  int64_t n = C->ne;
  double H[ORC_MAXM * ORC_MAXM];
  memset(H, 0, sizeof(H));
  for (int64_t i = 0; i < n; i++) {
    double pl = C->p0[i], ql = C->q0[i], PQ[ORC_MAXM];
    double ux = C->Uas[i] - x[i], xl = x[i] - C->L[i];
    for (int j = 0; j < m; j++) {
      pl += C->pij[j * n + i] * C->lam[j];
      ql += C->qij[j * n + i] * C->lam[j];
      PQ[j] = C->pij[j * n + i] / pow(ux, 2.0) - C->qij[j * n + i] / pow(xl, 2.0);
    }
    double df2 = -1.0 / (2.0 * pl / pow(ux, 3.0) + 2.0 * ql / pow(xl, 3.0));
    double xp = (sqrt(pl) * C->L[i] + sqrt(ql) * C->Uas[i]) / (sqrt(pl) + sqrt(ql));
    if (xp < C->alpha[i]) df2 = 0.0;
    if (xp > C->betav[i]) df2 = 0.0;
    for (int j = 0; j < m; j++)
      for (int k = 0; k < m; k++) H[j * m + k] += PQ[j] * df2 * PQ[k];
  }
  double lamai = 0.0;
  for (int j = 0; j < m; j++) {
    if (C->lam[j] < 0.0) C->lam[j] = 0.0;
    lamai += C->lam[j] * C->a[j];
    if (C->lam[j] > C->c[j]) H[j * m + j] += -1.0;
    H[j * m + j] += -C->mu[j] / C->lam[j];
  }
  if (lamai > 0.0)
    for (int j = 0; j < m; j++)
      for (int k = 0; k < m; k++) H[j * m + k] += -10.0 * C->a[j] * C->a[k];
  double tr = 0.0;
  for (int j = 0; j < m; j++) tr += H[j * m + j];
  double corr = 1e-4 * tr / m;
  if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
  for (int j = 0; j < m; j++) H[j * m + j] += corr;
  memcpy(C->Hess, H, sizeof(H));
}

/* In-place LU without pivoting and the matching solve (MMA.cpp:936-969). */
static void lu_solve(double *K, double *x, int nn) {
  for (int s = 0; s < nn - 1; s++)
    for (int i = s + 1; i < nn; i++) {
      K[i * nn + s] = K[i * nn + s] / K[s * nn + s];
      for (int j = s + 1; j < nn; j++) K[i * nn + j] -= K[i * nn + s] * K[s * nn + j];
    }
  for (int i = 1; i < nn; i++) {
    double acc = 0.0;
    for (int j = 0; j < i; j++) acc -= K[i * nn + j] * x[j];
    x[i] += acc;
  }
  x[nn - 1] = x[nn - 1] / K[(nn - 1) * nn + (nn - 1)];
  for (int i = nn - 2; i >= 0; i--) {
    double acc = x[i];
    for (int j = i + 1; j < nn; j++) acc -= K[i * nn + j] * x[j];
    x[i] = acc / K[i * nn + i];
  }
}

/* MMA::Update = GenSub + xold shift + SolveDIP (MMA.cpp:483-502, 506-676) */
void orc_mma_update(orc_ctx *C, double *x, const double *dfdx, const double *gx, const double *dgdx,
                    const double *xmin, const double *xmax) {
  int m = C->P.m;
  int64_t n = C->ne;
  /* GenSub: asymptotes (MMA.cpp:511-578) */
  C->mk++;
  if (C->mk < 3) {
    for (int64_t i = 0; i < n; i++) {
      C->L[i] = x[i] - C->asyminit * (xmax[i] - xmin[i]);
      C->Uas[i] = x[i] + C->asyminit * (xmax[i] - xmin[i]);
    }
  } else {
    for (int64_t i = 0; i < n; i++) {
      double hv = (x[i] - C->xo1[i]) * (C->xo1[i] - C->xo2[i]);
      double gm = hv < 0.0 ? C->asymdec : (hv > 0.0 ? C->asyminc : 1.0);
      double Li = x[i] - gm * (C->xo1[i] - C->L[i]);
      double Ui = x[i] + gm * (C->Uas[i] - C->xo1[i]);
      double xmi = dmax(1.0e-5, xmax[i] - xmin[i]);
      if (C->robust == 0) {
        Li = dmax(Li, x[i] - 10.0 * xmi);
        Li = dmin(Li, x[i] - 0.01 * xmi);
        Ui = dmax(Ui, x[i] + 0.01 * xmi);
        Ui = dmin(Ui, x[i] + 10.0 * xmi);
      } else { /* RobustAsymptotesType 1 (MMA.cpp:557-575) */
        Li = dmax(Li, x[i] - 100.0 * xmi);
        Li = dmin(Li, x[i] - 1.0e-4 * xmi);
        Ui = dmax(Ui, x[i] + 1.0e-4 * xmi);
        Ui = dmin(Ui, x[i] + 100.0 * xmi);
        double lo = xmin[i] - 1.0e-5, hi = xmax[i] + 1.0e-5;
        if (x[i] < lo) { Li = x[i] - (hi - x[i]) / 0.9; Ui = x[i] + (hi - x[i]) / 0.9; }
        if (x[i] > hi) { Li = x[i] - (x[i] - lo) / 0.9; Ui = x[i] + (x[i] - lo) / 0.9; }
      }
      C->L[i] = Li; C->Uas[i] = Ui;
    }
  }
  /* bounds and approximation coefficients (MMA.cpp:579-602) */
  const double feps = 1.0e-6;
  for (int64_t i = 0; i < n; i++) {
    double Li = C->L[i], Ui = C->Uas[i], xi = x[i];
    C->alpha[i] = dmax(xmin[i], 0.9 * Li + 0.1 * xi);
    C->betav[i] = dmin(xmax[i], 0.9 * Ui + 0.1 * xi);
    double dfp = dmax(0.0, dfdx[i]), dfm = dmax(0.0, -1.0 * dfdx[i]);
    C->p0[i] = pow(Ui - xi, 2.0) * (dfp + 0.001 * fabs(dfdx[i]) + 0.5 * feps / (Ui - Li));
    C->q0[i] = pow(xi - Li, 2.0) * (dfm + 0.001 * fabs(dfdx[i]) + 0.5 * feps / (Ui - Li));
    for (int j = 0; j < m; j++) {
      double g = dgdx[j * n + i];
      double reg = C->conmod ? 0.001 * fabs(g) + 0.5 * feps / (Ui - Li) : 0.0; /* MMA.cpp:588-597 */
      C->pij[j * n + i] = pow(Ui - xi, 2.0) * (dmax(0.0, g) + reg);
      C->qij[j * n + i] = pow(xi - Li, 2.0) * (dmax(0.0, -1.0 * g) + reg);
    }
  }
  /* b = sum(p/(U-x) + q/(x-L)) - gx (MMA.cpp:603-620) */
  double bs[ORC_MAXM];
  mma_gsum(C, x, bs);
  for (int j = 0; j < m; j++) C->b[j] = bs[j] - gx[j];
  /* xo2 := xo1; xo1 := x (MMA.cpp:496-497) */
  memcpy(C->xo2, C->xo1, n * sizeof(double));
  memcpy(C->xo1, x, n * sizeof(double));

  /* SolveDIP (MMA.cpp:639-676) */
  for (int j = 0; j < m; j++) { C->lam[j] = C->c[j] / 2.0; C->mu[j] = 1.0; }
  double tol = 1.0e-9 * sqrt((double)(m + n));
  double epsi = 1.0, err = 1.0;
  int steps = 0;
  while (epsi > tol) {
    int loop = 0;
    while (err > 0.9 * epsi && loop < 100) {
      loop++;
      steps++;
      mma_xyz(C, x);
      double gs[ORC_MAXM];
      mma_gsum(C, x, gs);
      for (int j = 0; j < m; j++) C->grad[j] = gs[j] - C->b[j] - C->a[j] * C->z - C->y[j];
      for (int j = 0; j < m; j++) C->grad[j] = -1.0 * C->grad[j] - epsi / C->lam[j];
      mma_hess(C, x);
      lu_solve(C->Hess, C->grad, m);
      for (int j = 0; j < m; j++) C->s[j] = C->grad[j];
      for (int j = 0; j < m; j++) C->s[m + j] = -C->mu[j] + epsi / C->lam[j] - C->s[j] * C->mu[j] / C->lam[j];
      /* DualLineSearch (MMA.cpp:870-888) */
      double theta = 1.005;
      for (int j = 0; j < m; j++) {
        if (theta < -1.01 * C->s[j] / C->lam[j]) theta = -1.01 * C->s[j] / C->lam[j];
        if (theta < -1.01 * C->s[j + m] / C->mu[j]) theta = -1.01 * C->s[j + m] / C->mu[j];
      }
      theta = 1.0 / theta;
      for (int j = 0; j < m; j++) { C->lam[j] += theta * C->s[j]; C->mu[j] += theta * C->s[j + m]; }
      mma_xyz(C, x);
      /* DualResidual (MMA.cpp:890-934) */
      mma_gsum(C, x, gs);
      double nrI = 0.0;
      for (int j = 0; j < m; j++) {
        double r1 = gs[j] - C->b[j] - C->a[j] * C->z - C->y[j] + C->mu[j];
        double r2 = C->mu[j] * C->lam[j] - epsi;
        if (nrI < fabs(r1)) nrI = fabs(r1);
        if (nrI < fabs(r2)) nrI = fabs(r2);
      }
      err = nrI;
    }
    epsi = epsi * 0.1;
  }
  C->dip_steps = steps;
}

/* ------------------------------------------------------------------------- */
/* The design loop of main.cpp:64-119                                         */
/* ------------------------------------------------------------------------- */

/* One design iteration (main.cpp:72-107).  Returns the rc of the solves. */
int orc_design_iteration(orc_ctx *C, int itr, double *fx, double *gx, double *ch) {
  int rc = orc_fem_obj_sens(C, C->XPhys, &C->fx, C->gx, C->dfdx, C->dgdx);
  if (C->P.proj) orc_projected_sens(C, C->XFiled, C->dfdx, C->dgdx);
  rc |= orc_filter_sens(C, C->dfdx, C->dgdx);
  orc_mma_movelimit(C, C->X, C->xmin, C->xmax);
  orc_mma_update(C, C->X, C->dfdx, C->gx, C->dgdx, C->xmin, C->xmax);
  *ch = orc_design_change(C, C->X, C->Xold);
  rc |= orc_filter(C, C->X, C->XFiled);
  orc_heaviside(C, C->XFiled, C->XPhys, itr);
  *fx = C->fx;
  for (int j = 0; j < C->P.m; j++) gx[j] = C->gx[j];
  return rc;
}

/* Initial filter + projection (main.cpp:64-66) */
int orc_init_design(orc_ctx *C) {
  int rc = orc_filter(C, C->X, C->XFiled);
  orc_heaviside(C, C->XFiled, C->XPhys, 0);
  return rc;
}

/* Full loop: while (itr < maxItr && ch > 0.01) (main.cpp:70).  Histories are
 * written for each executed iteration; returns the iteration count. */
int orc_run(orc_ctx *C, int maxitr, double *fx_hist, double *gx_hist, double *ch_hist) {
  orc_init_design(C);
  double ch = 1.0;
  int itr = 0;
  while (itr < maxitr && ch > 0.01) {
    itr++;
    double fx, gx[ORC_MAXM];
    orc_design_iteration(C, itr, &fx, gx, &ch);
    if (fx_hist) fx_hist[itr - 1] = fx;
    if (gx_hist) gx_hist[itr - 1] = gx[0];
    if (ch_hist) ch_hist[itr - 1] = ch;
  }
  return itr;
}

/* ------------------------------------------------------------------------- */
/* Bounded CPU-baseline sample (used only by bench.py's cpu_baseline leg)     */
/* ------------------------------------------------------------------------- */
#include <time.h>
static double wclock(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* Times the pieces of one reference-structure design iteration on the full
 * mesh held by C without running a full solve:
 *   out[0] seconds per element of Hex8Isoparametric/Tet4Isoparametric (KE)
 *   out[1] seconds per element of the 24x24 (12x12) CSR scatter-add
 *   out[2] seconds per element of the u^T KE u sensitivity loop body
 *   out[3] seconds per Jacobi-PCG iteration on the full assembled K pattern
 *   out[4] seconds per filter CG iteration on the full KF
 *   out[5] seconds of one MMA update (GenSub + dual interior point) on ne vars
 * ne_sample elements are used for the per-element timings; the CSR values
 * come from the cached KE of the first ne_sample elements replicated
 * (timings do not depend on the values). */
int orc_sample_costs(orc_ctx *C, int64_t ne_sample, int n_it, double *out) {
  if (ne_sample > C->ne) ne_sample = C->ne;
  int nd = C->nd, npe = C->npe;
  double KE[576];
  double t0 = wclock(), sink = 0.0;
  for (int64_t e = 0; e < ne_sample; e++) {
    double X[24];
    elem_coords(C, e, X);
    if (npe == 8) orc_hex8_ke(X, C->P.nu, KE); else orc_tet4_ke(X, C->P.nu, KE);
    sink += KE[e % (nd * nd)];
  }
  out[0] = (wclock() - t0) / ne_sample;
  memset(C->K.v, 0, C->K.rp[C->K.n] * sizeof(double));
  t0 = wclock();
  for (int64_t e = 0; e < ne_sample; e++) {
    int32_t dof[24];
    for (int k = 0; k < nd; k++) dof[k] = 3 * C->elem[e * npe + (k % npe)] + k / npe;
    for (int p = 0; p < nd; p++)
      for (int q = 0; q < nd; q++) C->K.v[csr_find(&C->K, dof[p], dof[q])] += KE[p * nd + q] * 0.5;
  }
  out[1] = (wclock() - t0) / ne_sample;
  t0 = wclock();
  for (int64_t e = 0; e < ne_sample; e++) {
    double ue[24];
    for (int k = 0; k < nd; k++) ue[k] = C->U[3 * C->elem[e * npe + (k % npe)] + k / npe] + 1e-3 * k;
    double uKu = 0.0;
    for (int p = 0; p < nd; p++)
      for (int q = 0; q < nd; q++) uKu += ue[p] * KE[p * nd + q] * ue[q];
    sink += uKu;
  }
  out[2] = (wclock() - t0) / ne_sample;
  /* full-size K values from the last element matrix replicated, SPD-shifted */
  memset(C->K.v, 0, C->K.rp[C->K.n] * sizeof(double));
  for (int64_t e = 0; e < C->ne; e++) {
    int32_t dof[24];
    for (int k = 0; k < nd; k++) dof[k] = 3 * C->elem[e * npe + (k % npe)] + k / npe;
    for (int p = 0; p < nd; p++)
      for (int q = 0; q < nd; q++) C->K.v[csr_find(&C->K, dof[p], dof[q])] += KE[p * nd + q] + (p == q ? 1e-3 : 0.0);
  }
  for (int64_t i = 0; i < C->ndof; i++) C->F[i] = 1.0;
  double rtol = C->P.rtol_el;
  int maxit = C->P.maxit_el;
  C->P.rtol_el = 1e-300;
  C->P.maxit_el = n_it;
  t0 = wclock();
  pcg_el(C);
  out[3] = (wclock() - t0) / n_it;
  C->P.rtol_el = rtol;
  C->P.maxit_el = maxit;
  double rtf = C->P.rtol_f;
  int mf = C->P.maxit_f;
  C->P.rtol_f = 1e-300;
  C->P.maxit_f = n_it;
  for (int64_t i = 0; i < C->nn; i++) { C->TX[i] = 1.0 + 1e-3 * (i % 7); C->RX[i] = 0.0; }
  t0 = wclock();
  cg_filter(C, C->TX, C->RX);
  out[4] = (wclock() - t0) / (C->it_f > 0 ? C->it_f : 1);
  C->P.rtol_f = rtf;
  C->P.maxit_f = mf;
  double gx[ORC_MAXM] = {0.0};
  for (int64_t e = 0; e < C->ne; e++) { C->dfdx[e] = -1.0 - (e % 13); C->dgdx[e] = C->vol[e] / C->TV; }
  orc_mma_movelimit(C, C->X, C->xmin, C->xmax);
  t0 = wclock();
  orc_mma_update(C, C->X, C->dfdx, gx, C->dgdx, C->xmin, C->xmax);
  out[5] = wclock() - t0;
  out[6] = sink;
  return 0;
}

/* MMA settings: SetAsymptotes (MMA.cpp:347-355), SetRobustAsymptotesType
 * (MMA.cpp:357-368), ConstraintModification (MMA.h:33-36), user a, c, d
 * (MMA.h:13, OptPara.cpp:166-174). */
void orc_mma_settings(orc_ctx *C, double init, double dec, double inc, int robust, int conmod) {
  C->asyminit = init; C->asymdec = dec; C->asyminc = inc; C->robust = robust; C->conmod = conmod;
}
void orc_mma_set_acd(orc_ctx *C, const double *a, const double *c, const double *d) {
  for (int j = 0; j < C->P.m; j++) { C->a[j] = a[j]; C->c[j] = c[j]; C->d[j] = d[j]; }
}
/* MMA::Restart (MMA.cpp:302-342): copy the state out; the restart
 * constructors (MMA.cpp:5-176) copy it in with iteration counter k. */
void orc_mma_state_get(orc_ctx *C, int *k, double *xo1, double *xo2, double *U, double *L) {
  size_t b = C->ne * sizeof(double);
  *k = C->mk; memcpy(xo1, C->xo1, b); memcpy(xo2, C->xo2, b); memcpy(U, C->Uas, b); memcpy(L, C->L, b);
}
void orc_mma_state_set(orc_ctx *C, int k, const double *xo1, const double *xo2, const double *U, const double *L) {
  size_t b = C->ne * sizeof(double);
  C->mk = k; memcpy(C->xo1, xo1, b); memcpy(C->xo2, xo2, b); memcpy(C->Uas, U, b); memcpy(C->L, L, b);
}
/* MMA::KKTresidual (MMA.cpp:412-480) with the multipliers lam, y, z of the
 * last update. */
void orc_mma_kkt(orc_ctx *C, const double *x, const double *dfdx, const double *gx, const double *dgdx,
                 const double *xmin, const double *xmax, double *norm2, double *norminf) {
  int m = C->P.m;
  int64_t n = C->ne;
  double n2 = 0.0, ni = 0.0;
  for (int64_t i = 0; i < n; i++) {
    double ri = dfdx[i];
    for (int j = 0; j < m; j++) ri += C->lam[j] * dgdx[j * n + i];
    double mumin = (x[i] < xmin[i] + 1.0e-5 && ri > 0.0) ? ri : 0.0;
    double mumax = (x[i] > xmax[i] - 1.0e-5 && ri < 0.0) ? -ri : 0.0;
    ri += -mumin + mumax;
    double r1 = mumin * (x[i] - xmin[i]), r2 = mumax * (xmax[i] - x[i]);
    n2 += ri * ri + r1 * r1 + r2 * r2;
    ni = dmax(ni, dmax(fabs(ri), dmax(fabs(r1), fabs(r2))));
  }
  double rl = 0.0;
  for (int j = 0; j < m; j++) rl += C->lam[j] * (C->a[j] * C->z + C->y[j] - gx[j]);
  n2 += rl * rl;
  *norm2 = sqrt(n2);
  *norminf = dmax(ni, fabs(rl));
}

/* SolveFem::ComputeStress (stress/ComputeStress.cpp:53-104): centroid strain
 * B(0,0,0) u_e of each hex8, sigma = Emax D B u_e scaled by sqrt(x_e), von
 * Mises of the six components (shear order xy, yz, zx).  tet4: the element's
 * constant strain (an extension: the reference has the hex8 routine only). */
static int inv3_cof(const double *J, double *Ji) {
  double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) + J[2] * (J[3] * J[7] - J[4] * J[6]);
  if (det == 0.0) return -1;
  Ji[0] = (J[4] * J[8] - J[5] * J[7]) / det; Ji[1] = (J[2] * J[7] - J[1] * J[8]) / det; Ji[2] = (J[1] * J[5] - J[2] * J[4]) / det;
  Ji[3] = (J[5] * J[6] - J[3] * J[8]) / det; Ji[4] = (J[0] * J[8] - J[2] * J[6]) / det; Ji[5] = (J[2] * J[3] - J[0] * J[5]) / det;
  Ji[6] = (J[3] * J[7] - J[4] * J[6]) / det; Ji[7] = (J[1] * J[6] - J[0] * J[7]) / det; Ji[8] = (J[0] * J[4] - J[1] * J[3]) / det;
  return 0;
}
void orc_stress(orc_ctx *C, const double *xphys, const double *u, double *vms) {
  const int npe = C->npe;
  const double nu = C->P.nu, Emax = C->P.Emax;
  const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 1.0 / (2.0 * (1.0 + nu));
  /* reference derivatives at the centroid: hex8 diagonal order, tet4 linear */
  double dNr[3][8];
  if (npe == 8) {
    for (int a = 0; a < 8; a++) {
      double sx = (a & 1) ? 1.0 : -1.0, sy = (a & 2) ? 1.0 : -1.0, sz = (a & 4) ? 1.0 : -1.0;
      dNr[0][a] = 0.125 * sx; dNr[1][a] = 0.125 * sy; dNr[2][a] = 0.125 * sz;
    }
  } else {
    double t[3][4] = {{-1, 1, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 0, 1}};
    for (int r = 0; r < 3; r++) for (int a = 0; a < 4; a++) dNr[r][a] = t[r][a];
  }
  for (int64_t e = 0; e < C->ne; e++) {
    double J[9] = {0}, Ji[9];
    for (int r = 0; r < 3; r++)
      for (int a = 0; a < npe; a++) {
        const double *X = &C->xyz[3 * (int64_t)C->elem[e * npe + a]];
        for (int c = 0; c < 3; c++) J[r * 3 + c] += dNr[r][a] * X[c];
      }
    inv3_cof(J, Ji);
    double g[3][8];
    for (int r = 0; r < 3; r++)
      for (int a = 0; a < npe; a++) g[r][a] = Ji[r * 3 + 0] * dNr[0][a] + Ji[r * 3 + 1] * dNr[1][a] + Ji[r * 3 + 2] * dNr[2][a];
    double eps[6] = {0};
    for (int a = 0; a < npe; a++) {
      const double *ua = &u[3 * (int64_t)C->elem[e * npe + a]];
      eps[0] += g[0][a] * ua[0];
      eps[1] += g[1][a] * ua[1];
      eps[2] += g[2][a] * ua[2];
      eps[3] += g[1][a] * ua[0] + g[0][a] * ua[1];
      eps[4] += g[2][a] * ua[1] + g[1][a] * ua[2];
      eps[5] += g[2][a] * ua[0] + g[0][a] * ua[2];
    }
    double tr = eps[0] + eps[1] + eps[2], sc = Emax * sqrt(xphys[e]), sg[6];
    for (int k = 0; k < 3; k++) sg[k] = sc * (lam * tr + 2.0 * mu * eps[k]);
    for (int k = 3; k < 6; k++) sg[k] = sc * mu * eps[k];
    vms[e] = (1.0 / sqrt(2.0)) * sqrt(pow(sg[0] - sg[1], 2) + pow(sg[0] - sg[2], 2) + pow(sg[1] - sg[2], 2) +
                                      6.0 * (sg[3] * sg[3] + sg[4] * sg[4] + sg[5] * sg[5]));
  }
}
