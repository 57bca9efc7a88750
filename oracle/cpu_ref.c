/*
 * cpu_ref.c -- threaded CPU restatement of the reference's per-iteration
 * structure, for bench.py's cpu_baseline leg.
 *
 * TEST INFRASTRUCTURE ONLY (the same status as uft_oracle.c): loaded by
 * tests/ and by bench.py's cpu_baseline leg; the product library never
 * links it.
 *
 * One design iteration of main.cpp:72-107 as the reference performs it,
 * with OpenMP threads in place of MPI ranks (README.md:22 runs
 * mpiexec -np 8):
 *   SolveFem::AssembleFem (SolveFem.cpp:267-337): KE recomputed for every
 *     element (SolveFem.cpp:301), K re-assembled, MatZeroRowsColumns, loads;
 *   SolveFem::ComputeFem (SolveFem.cpp:91-157): a smoothed-aggregation
 *     multigrid rebuilt from the new matrix, as PCSetUp of PCGAMG does at
 *     every call (4 levels, PCGAMGSetNlevels, SolveFem.cpp:130-131; six
 *     rigid-body modes from the coordinates, PCSetCoordinates :132; one
 *     Jacobi smoothing step of every tentative prolongator, GAMG's default;
 *     Chebyshev(2)/Jacobi smoothers on [0.1, 1.1] lambda_max; V-cycle; dense
 *     Cholesky on the coarsest level) preconditioning FGMRES(30) from U = 0
 *     (SolveFem.cpp:126-139);
 *   ComputeObjectiveAndSensitivities (SolveFem.cpp:159-265): KE recomputed
 *     again (SolveFem.cpp:214);
 *   PDEFilter: Jacobi-CG solves on the assembled KF (PDEFilter.cpp:117-183);
 *   MMA: the oracle's update (serial, MMA.cpp:483-676).
 * With ke_cache = 1 the element matrices are computed once at setup (the
 * "KE-cached" variant).  Aggregation, prolongators and coarse operators are
 * all rebuilt inside the iteration.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "orc_internal.h"

static double crf_clock(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* ------------------------------------------------------------------------- */
/* threaded sparse kernels                                                    */
/* ------------------------------------------------------------------------- */

static void pmv(const csr_t *A, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < A->n; i++) {
    double acc = 0.0;
    for (int64_t k = A->rp[i]; k < A->rp[i + 1]; k++) acc += A->v[k] * x[A->ci[k]];
    y[i] = acc;
  }
}

static double pdot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

static void csr_release(csr_t *A) {
  free(A->rp);
  free(A->ci);
  free(A->v);
  memset(A, 0, sizeof(*A));
}

/* C = A B (Gustavson, rows in parallel; per-thread dense accumulator over the
 * ncol columns of B); columns of each row in order of first appearance */
static void spgemm(const csr_t *A, const csr_t *B, int64_t ncol, csr_t *C) {
  const int64_t n = A->n;
  C->n = n;
  C->rp = calloc(n + 1, sizeof(int64_t));
#pragma omp parallel
  {
    int64_t *mark = malloc(ncol * sizeof(int64_t));
    for (int64_t j = 0; j < ncol; j++) mark[j] = -1;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
      int64_t c = 0;
      for (int64_t k = A->rp[i]; k < A->rp[i + 1]; k++) {
        const int32_t a = A->ci[k];
        for (int64_t l = B->rp[a]; l < B->rp[a + 1]; l++) {
          const int32_t j = B->ci[l];
          if (mark[j] != i) {
            mark[j] = i;
            c++;
          }
        }
      }
      C->rp[i + 1] = c;
    }
    free(mark);
  }
  for (int64_t i = 0; i < n; i++) C->rp[i + 1] += C->rp[i];
  C->ci = malloc((C->rp[n] + 1) * sizeof(int32_t));
  C->v = malloc((C->rp[n] + 1) * sizeof(double));
#pragma omp parallel
  {
    int64_t *pos = malloc(ncol * sizeof(int64_t));
    for (int64_t j = 0; j < ncol; j++) pos[j] = -1;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
      const int64_t r0 = C->rp[i];
      int64_t w = r0;
      for (int64_t k = A->rp[i]; k < A->rp[i + 1]; k++) {
        const int32_t a = A->ci[k];
        const double av = A->v[k];
        for (int64_t l = B->rp[a]; l < B->rp[a + 1]; l++) {
          const int32_t j = B->ci[l];
          if (pos[j] < r0) {
            pos[j] = w;
            C->ci[w] = j;
            C->v[w++] = av * B->v[l];
          } else {
            C->v[pos[j]] += av * B->v[l];
          }
        }
      }
    }
    free(pos);
  }
}

/* B = A^T (A has ncol columns) */
static void transpose(const csr_t *A, int64_t ncol, csr_t *B) {
  B->n = ncol;
  B->rp = calloc(ncol + 1, sizeof(int64_t));
  for (int64_t k = 0; k < A->rp[A->n]; k++) B->rp[A->ci[k] + 1]++;
  for (int64_t j = 0; j < ncol; j++) B->rp[j + 1] += B->rp[j];
  B->ci = malloc((B->rp[ncol] + 1) * sizeof(int32_t));
  B->v = malloc((B->rp[ncol] + 1) * sizeof(double));
  int64_t *pos = malloc((ncol + 1) * sizeof(int64_t));
  memcpy(pos, B->rp, ncol * sizeof(int64_t));
  for (int64_t i = 0; i < A->n; i++)
    for (int64_t k = A->rp[i]; k < A->rp[i + 1]; k++) {
      const int64_t w = pos[A->ci[k]]++;
      B->ci[w] = (int32_t)i;
      B->v[w] = A->v[k];
    }
  free(pos);
}

/* ------------------------------------------------------------------------- */
/* aggregation multigrid, rebuilt every solve                                 */
/* ------------------------------------------------------------------------- */

#define CRF_MAXLEV 4

typedef struct {
  csr_t A, P, R;  /* A of this level; P to the next level, R = P^T */
  int64_t n;
  double *dinv, lmax;
  double *x, *b, *r, *d, *t;
} crf_level;

typedef struct {
  int nlev;
  crf_level L[CRF_MAXLEV];
  double *cf;  /* Cholesky factor of the coarsest operator (row-major, lower) */
  int64_t nc;
} crf_amg;

/* greedy aggregation (Vanek, Mandel, Brezina 1996, phases 1-3) on a graph
 * whose vertices with skip[v] set are left out (agg -1) */
static int64_t crf_aggregate(int64_t n, const int64_t *rp, const int32_t *ci, const uint8_t *skip, int32_t *agg) {
  int32_t na = 0;
  for (int64_t i = 0; i < n; i++) agg[i] = -1;
  for (int64_t i = 0; i < n; i++) {
    if (agg[i] >= 0 || (skip && skip[i])) continue;
    int free_nb = 1;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
      const int32_t j = ci[k];
      if (j != i && agg[j] >= 0) {
        free_nb = 0;
        break;
      }
    }
    if (!free_nb) continue;
    agg[i] = na;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++)
      if (!(skip && skip[ci[k]])) agg[ci[k]] = na;
    na++;
  }
  int32_t *tent = malloc(n * sizeof(int32_t));
  memcpy(tent, agg, n * sizeof(int32_t));
  for (int64_t i = 0; i < n; i++) {
    if (agg[i] >= 0 || (skip && skip[i])) continue;
    int32_t best = -1, bestc = 0;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
      const int32_t a = agg[ci[k]];
      if (a < 0) continue;
      int32_t c = 0;
      for (int64_t k2 = rp[i]; k2 < rp[i + 1]; k2++)
        if (agg[ci[k2]] == a) c++;
      if (c > bestc || (c == bestc && a < best)) {
        best = a;
        bestc = c;
      }
    }
    tent[i] = best;
  }
  memcpy(agg, tent, n * sizeof(int32_t));
  free(tent);
  for (int64_t i = 0; i < n; i++) {
    if (agg[i] >= 0 || (skip && skip[i])) continue;
    agg[i] = na;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++)
      if (agg[ci[k]] < 0 && !(skip && skip[ci[k]])) agg[ci[k]] = na;
    na++;
  }
  return na;
}

/* two-pass modified Gram-Schmidt on the 6 columns of B (rows x 6); Q in B,
 * R (6x6) in Rf; columns that fall below 1e-10 of their norm become zero */
static void crf_qr6(double *B, int64_t rows, double *Rf) {
  memset(Rf, 0, 36 * sizeof(double));
  double n0[6];
  for (int j = 0; j < 6; j++) {
    double s = 0.0;
    for (int64_t r = 0; r < rows; r++) s += B[r * 6 + j] * B[r * 6 + j];
    n0[j] = sqrt(s);
  }
  for (int j = 0; j < 6; j++) {
    for (int pass = 0; pass < 2; pass++)
      for (int i = 0; i < j; i++) {
        double s = 0.0;
        for (int64_t r = 0; r < rows; r++) s += B[r * 6 + i] * B[r * 6 + j];
        Rf[i * 6 + j] += s;
        for (int64_t r = 0; r < rows; r++) B[r * 6 + j] -= s * B[r * 6 + i];
      }
    double s = 0.0;
    for (int64_t r = 0; r < rows; r++) s += B[r * 6 + j] * B[r * 6 + j];
    s = sqrt(s);
    if (s <= 1e-10 * n0[j] || s == 0.0) {
      for (int64_t r = 0; r < rows; r++) B[r * 6 + j] = 0.0;
      Rf[j * 6 + j] = 0.0;
    } else {
      for (int64_t r = 0; r < rows; r++) B[r * 6 + j] /= s;
      Rf[j * 6 + j] = s;
    }
  }
}

/* tentative prolongator from aggregates agg[unit] (units of bs unknowns) and
 * the near-null space Bn (nunit*bs x 6, row-major); rows of skipped unknowns
 * (skip[row]) are zero.  Returns the coarse near-null space (na*6 x 6). */
static double *crf_tentative(int64_t nunit, int bs, const int32_t *agg, int64_t na, const double *Bn,
                             const uint8_t *skip, csr_t *P) {
  int64_t *cnt = calloc(na + 1, sizeof(int64_t));
  for (int64_t u = 0; u < nunit; u++)
    if (agg[u] >= 0) cnt[agg[u] + 1]++;
  for (int64_t a = 0; a < na; a++) cnt[a + 1] += cnt[a];
  int32_t *mem = malloc((cnt[na] + 1) * sizeof(int32_t));
  int64_t *pos = malloc((na + 1) * sizeof(int64_t));
  memcpy(pos, cnt, na * sizeof(int64_t));
  for (int64_t u = 0; u < nunit; u++)
    if (agg[u] >= 0) mem[pos[agg[u]]++] = (int32_t)u;
  const int64_t n = nunit * bs;
  P->n = n;
  P->rp = calloc(n + 1, sizeof(int64_t));
  for (int64_t u = 0; u < nunit; u++)
    for (int c = 0; c < bs; c++) P->rp[u * bs + c + 1] = agg[u] >= 0 ? 6 : 0;
  for (int64_t i = 0; i < n; i++) P->rp[i + 1] += P->rp[i];
  P->ci = malloc((P->rp[n] + 1) * sizeof(int32_t));
  P->v = malloc((P->rp[n] + 1) * sizeof(double));
  double *Bc = calloc(na * 36, sizeof(double));
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t a = 0; a < na; a++) {
    const int64_t k = cnt[a + 1] - cnt[a], rows = k * bs;
    double *B = malloc(rows * 6 * sizeof(double));
    for (int64_t q = 0; q < k; q++) {
      const int64_t u = mem[cnt[a] + q];
      for (int c = 0; c < bs; c++) {
        const int64_t row = u * bs + c;
        for (int j = 0; j < 6; j++) B[(q * bs + c) * 6 + j] = (skip && skip[row]) ? 0.0 : Bn[row * 6 + j];
      }
    }
    crf_qr6(B, rows, Bc + a * 36);
    for (int64_t q = 0; q < k; q++) {
      const int64_t u = mem[cnt[a] + q];
      for (int c = 0; c < bs; c++) {
        const int64_t row = u * bs + c;
        const int64_t w = P->rp[row];
        for (int j = 0; j < 6; j++) {
          P->ci[w + j] = (int32_t)(a * 6 + j);
          P->v[w + j] = B[(q * bs + c) * 6 + j];
        }
      }
    }
    free(B);
  }
  free(cnt);
  free(mem);
  free(pos);
  return Bc;
}

static void crf_alloc_vecs(crf_level *L) {
  L->dinv = malloc(L->n * sizeof(double));
  L->x = malloc(L->n * sizeof(double));
  L->b = malloc(L->n * sizeof(double));
  L->r = malloc(L->n * sizeof(double));
  L->d = malloc(L->n * sizeof(double));
  L->t = malloc(L->n * sizeof(double));
}

static void crf_diag(crf_level *L) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < L->n; i++) {
    double d = 0.0;
    for (int64_t k = L->A.rp[i]; k < L->A.rp[i + 1]; k++)
      if (L->A.ci[k] == i) d = L->A.v[k];
    L->dinv[i] = d > 0.0 ? 1.0 / d : 1.0;
  }
}

/* lambda_max(D^-1 A) by 10 Lanczos (Jacobi-CG) steps from a fixed
 * pseudo-random vector, the largest Ritz value by bisection on the Sturm
 * sequence of the tridiagonal matrix */
static double crf_lmax(crf_level *L) {
  const int64_t n = L->n;
  const int ns = 10;
  double *r = L->r, *z = L->d, *p = L->t, *q = L->x;
  uint64_t s = 12345;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    uint64_t h = s + (uint64_t)i * 0x9E3779B97F4A7C15ULL;
    h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ULL;
    h = (h ^ (h >> 27)) * 0x94D049BB133111EBULL;
    h ^= h >> 31;
    r[i] = (double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
  double al[16], be[16];
  int m = 0;
  double rz = 0.0, alo = 0.0, bo = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rz)
  for (int64_t i = 0; i < n; i++) {
    z[i] = L->dinv[i] * r[i];
    p[i] = z[i];
    rz += r[i] * z[i];
  }
  for (int k = 0; k < ns && rz > 0.0; k++) {
    pmv(&L->A, p, q);
    const double pq = pdot(n, p, q);
    if (!(pq > 0.0)) break;
    const double a = rz / pq;
    double rzn = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rzn)
    for (int64_t i = 0; i < n; i++) {
      r[i] -= a * q[i];
      z[i] = L->dinv[i] * r[i];
      rzn += r[i] * z[i];
    }
    const double b = rzn / rz;
    /* Lanczos tridiagonal from the CG coefficients: T_kk = 1/a_k +
     * b_{k-1}/a_{k-1}, T_k,k+1 = sqrt(b_k)/a_k */
    al[m] = 1.0 / a + (k > 0 ? bo / alo : 0.0);
    be[m] = sqrt(b) / a;
    m++;
    alo = a;
    bo = b;
    rz = rzn;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) p[i] = z[i] + b * p[i];
  }
  if (m == 0) return 1.0;
  /* largest eigenvalue of T (diag al, off-diagonal be[0..m-2]) by bisection */
  double lo = 0.0, hi = 0.0;
  for (int i = 0; i < m; i++) {
    double rad = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i < m - 1 ? fabs(be[i]) : 0.0);
    if (al[i] + rad > hi) hi = al[i] + rad;
  }
  for (int it = 0; it < 100; it++) {
    const double mid = 0.5 * (lo + hi);
    int cnt = 0;
    double d = 1.0;
    for (int i = 0; i < m; i++) {
      d = al[i] - mid - (i > 0 ? be[i - 1] * be[i - 1] / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0.0) cnt++;
    }
    if (cnt == m) hi = mid; else lo = mid;
  }
  return hi;
}

static void crf_amg_free(crf_amg *H) {
  for (int l = 0; l < H->nlev; l++) {
    crf_level *L = &H->L[l];
    if (l > 0) csr_release(&L->A);
    csr_release(&L->P);
    csr_release(&L->R);
    free(L->dinv); free(L->x); free(L->b); free(L->r); free(L->d); free(L->t);
  }
  free(H->cf);
  memset(H, 0, sizeof(*H));
}

/* P = Pt - omega D^-1 (A Pt), omega = 4 / (3 lambda_max) */
static void crf_smooth_p(crf_level *L, const csr_t *Pt, int64_t ncol, csr_t *P) {
  csr_t AP;
  spgemm(&L->A, Pt, ncol, &AP);
  const double omega = 4.0 / (3.0 * L->lmax);
  /* P = Pt - omega D^-1 AP: row-wise merge (Pt's columns are a subset of
   * AP's whenever the diagonal of A is nonzero; merge generally anyway) */
  const int64_t n = L->n;
  P->n = n;
  P->rp = calloc(n + 1, sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) P->rp[i + 1] = AP.rp[i + 1] - AP.rp[i] + Pt->rp[i + 1] - Pt->rp[i];
  for (int64_t i = 0; i < n; i++) P->rp[i + 1] += P->rp[i];
  P->ci = malloc((P->rp[n] + 1) * sizeof(int32_t));
  P->v = malloc((P->rp[n] + 1) * sizeof(double));
  int64_t *len = malloc(n * sizeof(int64_t));
#pragma omp parallel
  {
    int64_t *pos = malloc(ncol * sizeof(int64_t));
    for (int64_t j = 0; j < ncol; j++) pos[j] = -1;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
      const int64_t r0 = P->rp[i];
      int64_t w = r0;
      const double s = -omega * L->dinv[i];
      for (int64_t k = AP.rp[i]; k < AP.rp[i + 1]; k++) {
        pos[AP.ci[k]] = w;
        P->ci[w] = AP.ci[k];
        P->v[w++] = s * AP.v[k];
      }
      for (int64_t k = Pt->rp[i]; k < Pt->rp[i + 1]; k++) {
        const int32_t j = Pt->ci[k];
        if (pos[j] >= r0) {
          P->v[pos[j]] += Pt->v[k];
        } else {
          P->ci[w] = j;
          P->v[w++] = Pt->v[k];
        }
      }
      len[i] = w - r0;
    }
    free(pos);
  }
  /* compact */
  int64_t w = 0;
  for (int64_t i = 0; i < n; i++) {
    const int64_t r0 = P->rp[i];
    if (w != r0) {
      memmove(P->ci + w, P->ci + r0, len[i] * sizeof(int32_t));
      memmove(P->v + w, P->v + r0, len[i] * sizeof(double));
    }
    P->rp[i] = w;
    w += len[i];
  }
  P->rp[n] = w;
  free(len);
  csr_release(&AP);
}

/* Galerkin A_c = R (A P) with R = P^T; zero rows (dropped near-null columns)
 * get a unit diagonal */
static void crf_galerkin(crf_level *L, int64_t nc, csr_t *Ac) {
  transpose(&L->P, nc, &L->R);
  csr_t AP;
  spgemm(&L->A, &L->P, nc, &AP);
  spgemm(&L->R, &AP, nc, Ac);
  csr_release(&AP);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nc; i++) {
    int64_t kd = -1;
    double s = 0.0;
    for (int64_t k = Ac->rp[i]; k < Ac->rp[i + 1]; k++) {
      s += fabs(Ac->v[k]);
      if (Ac->ci[k] == i) kd = k;
    }
    if (s == 0.0 && kd >= 0) Ac->v[kd] = 1.0;
  }
}

/* node graph of the fine level, coordinates for the rigid-body modes */
typedef struct {
  int64_t nn;
  int64_t *rp;
  int32_t *ci;
  const double *xyz;
} crf_mesh;

/* PCSetUp of PCGAMG on A (SolveFem.cpp:129-136): everything rebuilt */
static void crf_amg_setup(crf_amg *H, const csr_t *A0, const crf_mesh *M, const uint8_t *isfixed) {
  memset(H, 0, sizeof(*H));
  crf_level *L0 = &H->L[0];
  L0->A = *A0;  /* borrowed */
  L0->n = A0->n;
  crf_alloc_vecs(L0);
  crf_diag(L0);
  /* level 0 -> 1: aggregates of the node graph (nodes whose three DOFs are
   * all fixed are left out), rigid-body modes about each aggregate's mean */
  const int64_t nn = M->nn;
  uint8_t *skipn = malloc(nn);
  for (int64_t i = 0; i < nn; i++) skipn[i] = isfixed[3 * i] && isfixed[3 * i + 1] && isfixed[3 * i + 2];
  int32_t *agg = malloc(nn * sizeof(int32_t));
  int64_t na = crf_aggregate(nn, M->rp, M->ci, skipn, agg);
  double *cen = calloc(na * 4, sizeof(double));
  for (int64_t i = 0; i < nn; i++)
    if (agg[i] >= 0) {
      double *c = cen + 4 * agg[i];
      c[0] += M->xyz[3 * i]; c[1] += M->xyz[3 * i + 1]; c[2] += M->xyz[3 * i + 2]; c[3] += 1.0;
    }
  double *Bn = calloc(3 * nn * 6, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nn; i++) {
    if (agg[i] < 0) continue;
    const double *c = cen + 4 * agg[i];
    const double x = M->xyz[3 * i] - c[0] / c[3], y = M->xyz[3 * i + 1] - c[1] / c[3],
                 z = M->xyz[3 * i + 2] - c[2] / c[3];
    double *b = Bn + 18 * i;
    /* u, v, w translations; rotations about z, x, y */
    b[0 * 6 + 0] = 1; b[1 * 6 + 1] = 1; b[2 * 6 + 2] = 1;
    b[0 * 6 + 3] = -y; b[1 * 6 + 3] = x;
    b[1 * 6 + 4] = -z; b[2 * 6 + 4] = y;
    b[0 * 6 + 5] = z; b[2 * 6 + 5] = -x;
  }
  free(cen);
  csr_t Pt;
  double *Bc = crf_tentative(nn, 3, agg, na, Bn, isfixed, &Pt);
  free(Bn);
  free(skipn);
  free(agg);
  int lev = 0;
  int64_t nc = na * 6;
  while (1) {
    crf_level *L = &H->L[lev];
    L->lmax = crf_lmax(L);
    crf_smooth_p(L, &Pt, nc, &L->P);
    csr_release(&Pt);
    crf_level *Ln = &H->L[lev + 1];
    crf_galerkin(L, nc, &Ln->A);
    Ln->n = nc;
    crf_alloc_vecs(Ln);
    crf_diag(Ln);
    lev++;
    if (lev + 1 >= CRF_MAXLEV || nc <= 1200) break;
    /* units of 6: graph of the 6x6 block pattern */
    const int64_t nu = nc / 6;
    int64_t *grp = calloc(nu + 1, sizeof(int64_t));
    uint8_t *seen = calloc(nu, 1);
    int32_t *tmp = malloc(nu * sizeof(int32_t));
    int32_t *gci = NULL;
    int64_t cap = 0, w = 0;
    for (int64_t u = 0; u < nu; u++) {
      int t = 0;
      for (int c = 0; c < 6; c++) {
        const int64_t row = u * 6 + c;
        for (int64_t k = Ln->A.rp[row]; k < Ln->A.rp[row + 1]; k++) {
          const int32_t v = Ln->A.ci[k] / 6;
          if (!seen[v] && Ln->A.v[k] != 0.0) {
            seen[v] = 1;
            tmp[t++] = v;
          }
        }
      }
      if (w + t > cap) {
        cap = 2 * (w + t) + 1024;
        gci = realloc(gci, cap * sizeof(int32_t));
      }
      for (int q = 0; q < t; q++) {
        gci[w++] = tmp[q];
        seen[tmp[q]] = 0;
      }
      grp[u + 1] = w;
    }
    free(seen);
    free(tmp);
    int32_t *ag2 = malloc(nu * sizeof(int32_t));
    const int64_t na2 = crf_aggregate(nu, grp, gci, NULL, ag2);
    free(grp);
    free(gci);
    double *Bc2 = crf_tentative(nu, 6, ag2, na2, Bc, NULL, &Pt);
    free(Bc);
    free(ag2);
    Bc = Bc2;
    nc = na2 * 6;
  }
  free(Bc);
  H->nlev = lev + 1;
  /* coarsest: dense Cholesky */
  crf_level *Lc = &H->L[H->nlev - 1];
  const int64_t n = Lc->n;
  H->nc = n;
  double *D = calloc(n * n, sizeof(double));
  for (int64_t i = 0; i < n; i++)
    for (int64_t k = Lc->A.rp[i]; k < Lc->A.rp[i + 1]; k++) D[i * n + Lc->A.ci[k]] = Lc->A.v[k];
  for (int64_t j = 0; j < n; j++) {
    double s = D[j * n + j];
    for (int64_t k = 0; k < j; k++) s -= D[j * n + k] * D[j * n + k];
    const double dj = s > 0.0 ? sqrt(s) : 1.0;
    D[j * n + j] = dj;
#pragma omp parallel for schedule(static)
    for (int64_t i = j + 1; i < n; i++) {
      double t = D[i * n + j];
      for (int64_t k = 0; k < j; k++) t -= D[i * n + k] * D[j * n + k];
      D[i * n + j] = t / dj;
    }
  }
  H->cf = D;
}

/* Chebyshev(deg) with Jacobi on [0.1, 1.1] lambda_max for A x = b, x updated */
static void crf_cheb(crf_level *L, const double *b, double *x, int deg, int zero_guess) {
  const int64_t n = L->n;
  const double emax = 1.1 * L->lmax, emin = 0.1 * L->lmax;
  const double th = 0.5 * (emax + emin), de = 0.5 * (emax - emin), sg = th / de;
  double rho = 1.0 / sg;
  double *r = L->r, *d = L->d;
  if (zero_guess) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      r[i] = b[i];
      d[i] = L->dinv[i] * r[i] / th;
      x[i] = d[i];
    }
  } else {
    pmv(&L->A, x, L->t);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      r[i] = b[i] - L->t[i];
      d[i] = L->dinv[i] * r[i] / th;
      x[i] += d[i];
    }
  }
  for (int k = 1; k < deg; k++) {
    pmv(&L->A, d, L->t);
    const double rn = 1.0 / (2.0 * sg - rho);
    const double c1 = rn * rho, c2 = 2.0 * rn / de;
    rho = rn;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      r[i] -= L->t[i];
      d[i] = c1 * d[i] + c2 * L->dinv[i] * r[i];
      x[i] += d[i];
    }
  }
}

/* V-cycle on level l: x = B_l b */
static void crf_vcycle(crf_amg *H, int l, const double *b, double *x) {
  crf_level *L = &H->L[l];
  if (l == H->nlev - 1) {
    const int64_t n = H->nc;
    const double *D = H->cf;
    for (int64_t i = 0; i < n; i++) {
      double s = b[i];
      for (int64_t k = 0; k < i; k++) s -= D[i * n + k] * x[k];
      x[i] = s / D[i * n + i];
    }
    for (int64_t i = n - 1; i >= 0; i--) {
      double s = x[i];
      for (int64_t k = i + 1; k < n; k++) s -= D[k * n + i] * x[k];
      x[i] = s / D[i * n + i];
    }
    return;
  }
  crf_cheb(L, b, x, 2, 1);
  pmv(&L->A, x, L->t);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < L->n; i++) L->t[i] = b[i] - L->t[i];
  crf_level *Ln = &H->L[l + 1];
  pmv(&L->R, L->t, Ln->b);
  crf_vcycle(H, l + 1, Ln->b, Ln->x);
  pmv(&L->P, Ln->x, L->t);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < L->n; i++) x[i] += L->t[i];
  crf_cheb(L, b, x, 2, 0);
}

/* FGMRES(30), right preconditioning, x0 = 0; stops when the residual norm
 * of the Arnoldi recurrence reaches rtol ||b||; returns the iterations */
static int crf_fgmres(crf_amg *H, const csr_t *A, const double *b, double *x, double rtol, int maxit,
                      double *relres) {
  const int64_t n = A->n;
  const int m = 30;
  double *V = malloc((size_t)(m + 1) * n * sizeof(double));
  double *Z = malloc((size_t)m * n * sizeof(double));
  double Hh[31 * 30], cs[30], sn[30], g[31];
  memset(x, 0, n * sizeof(double));
  const double bn = sqrt(pdot(n, b, b));
  int it = 0;
  double rn = bn;
  if (bn == 0.0) {
    *relres = 0.0;
    free(V); free(Z);
    return 0;
  }
  while (it < maxit) {
    /* r = b - A x */
    pmv(A, x, V);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) V[i] = b[i] - V[i];
    rn = sqrt(pdot(n, V, V));
    if (rn <= rtol * bn) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) V[i] /= rn;
    memset(g, 0, sizeof(g));
    g[0] = rn;
    int j;
    for (j = 0; j < m && it < maxit; j++, it++) {
      double *vj = V + (size_t)j * n, *zj = Z + (size_t)j * n, *w = V + (size_t)(j + 1) * n;
      crf_vcycle(H, 0, vj, zj);
      pmv(A, zj, w);
      for (int i = 0; i <= j; i++) {
        const double *vi = V + (size_t)i * n;
        const double h = pdot(n, w, vi);
        Hh[i * m + j] = h;
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; k++) w[k] -= h * vi[k];
      }
      const double hn = sqrt(pdot(n, w, w));
      Hh[(j + 1) * m + j] = hn;
      if (hn > 0.0) {
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; k++) w[k] /= hn;
      }
      for (int i = 0; i < j; i++) {
        const double a = Hh[i * m + j], c = Hh[(i + 1) * m + j];
        Hh[i * m + j] = cs[i] * a + sn[i] * c;
        Hh[(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
      }
      const double a = Hh[j * m + j], c = Hh[(j + 1) * m + j];
      const double den = sqrt(a * a + c * c);
      cs[j] = den > 0.0 ? a / den : 1.0;
      sn[j] = den > 0.0 ? c / den : 0.0;
      Hh[j * m + j] = den;
      Hh[(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      rn = fabs(g[j + 1]);
      if (rn <= rtol * bn) {
        j++;
        it++;
        break;
      }
    }
    /* x += Z y, H y = g */
    double y[30];
    for (int i = j - 1; i >= 0; i--) {
      double s = g[i];
      for (int k = i + 1; k < j; k++) s -= Hh[i * m + k] * y[k];
      y[i] = s / Hh[i * m + i];
    }
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; k++) {
      double s = 0.0;
      for (int i = 0; i < j; i++) s += Z[(size_t)i * n + k] * y[i];
      x[k] += s;
    }
    if (rn <= rtol * bn) break;
  }
  /* true residual (SolveFem.cpp:142-152) */
  pmv(A, x, V);
  double t = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : t)
  for (int64_t i = 0; i < n; i++) t += (b[i] - V[i]) * (b[i] - V[i]);
  *relres = sqrt(t) / bn;
  free(V);
  free(Z);
  return it;
}

/* ------------------------------------------------------------------------- */
/* the reference-structure design iteration                                   */
/* ------------------------------------------------------------------------- */

typedef struct {
  orc_ctx *C;
  int nthreads;
  crf_mesh M;
  int ncolor;
  int64_t *cptr;     /* elements grouped by colour (no two share a node) */
  int32_t *celem;
  int64_t *kpos;     /* per element, CSR offsets of its nd x nd entries in K */
  int64_t *iptr;     /* node -> incident (element, local node) */
  int32_t *inc;
  double *work;
} crf_prob;

static int64_t kfind(const csr_t *A, int64_t row, int32_t col) {
  int64_t lo = A->rp[row], hi = A->rp[row + 1] - 1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (A->ci[mid] == col) return mid;
    if (A->ci[mid] < col) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* Setup (excluded from the timed iteration, like mesh input and
 * PrepareFilter): element colouring, CSR positions of every element entry
 * (MatSetValues' search, done once), node incidence, node graph. */
crf_prob *crf_create(orc_ctx *C, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  crf_prob *P = calloc(1, sizeof(crf_prob));
  P->C = C;
  P->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
  const int npe = C->npe, nd = C->nd;
  const int64_t ne = C->ne, nn = C->nn;
  /* greedy colouring */
  int32_t *col = malloc(ne * sizeof(int32_t));
  uint64_t *used = calloc(nn, sizeof(uint64_t));
  int ncolor = 0;
  for (int64_t e = 0; e < ne; e++) {
    uint64_t m = 0;
    for (int a = 0; a < npe; a++) m |= used[C->elem[e * npe + a]];
    int c = 0;
    while (c < 63 && ((m >> c) & 1u)) c++;
    col[e] = c;
    if (c + 1 > ncolor) ncolor = c + 1;
    for (int a = 0; a < npe; a++) used[C->elem[e * npe + a]] |= 1ull << c;
  }
  free(used);
  P->ncolor = ncolor;
  P->cptr = calloc(ncolor + 1, sizeof(int64_t));
  for (int64_t e = 0; e < ne; e++) P->cptr[col[e] + 1]++;
  for (int c = 0; c < ncolor; c++) P->cptr[c + 1] += P->cptr[c];
  P->celem = malloc(ne * sizeof(int32_t));
  int64_t *pos = malloc((ncolor + 1) * sizeof(int64_t));
  memcpy(pos, P->cptr, ncolor * sizeof(int64_t));
  for (int64_t e = 0; e < ne; e++) P->celem[pos[col[e]]++] = (int32_t)e;
  free(pos);
  free(col);
  P->kpos = malloc((size_t)ne * nd * nd * sizeof(int64_t));
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < ne; e++) {
    int32_t dof[24];
    for (int k = 0; k < nd; k++) dof[k] = 3 * C->elem[e * npe + (k % npe)] + k / npe;
    for (int p = 0; p < nd; p++)
      for (int q = 0; q < nd; q++) P->kpos[(e * nd + p) * nd + q] = kfind(&C->K, dof[p], dof[q]);
  }
  P->iptr = calloc(nn + 1, sizeof(int64_t));
  for (int64_t e = 0; e < ne; e++)
    for (int a = 0; a < npe; a++) P->iptr[C->elem[e * npe + a] + 1]++;
  for (int64_t i = 0; i < nn; i++) P->iptr[i + 1] += P->iptr[i];
  P->inc = malloc(ne * npe * sizeof(int32_t));
  int64_t *ip = malloc(nn * sizeof(int64_t));
  memcpy(ip, P->iptr, nn * sizeof(int64_t));
  for (int64_t e = 0; e < ne; e++)
    for (int a = 0; a < npe; a++) P->inc[ip[C->elem[e * npe + a]]++] = (int32_t)(e * npe + a);
  free(ip);
  /* node graph without self loops (the aggregation graph) */
  P->M.nn = nn;
  P->M.xyz = C->xyz;
  P->M.rp = malloc((nn + 1) * sizeof(int64_t));
  P->M.rp[0] = 0;
  int64_t w = 0;
  P->M.ci = malloc(C->KF.rp[nn] * sizeof(int32_t));
  for (int64_t i = 0; i < nn; i++) {
    for (int64_t k = C->KF.rp[i]; k < C->KF.rp[i + 1]; k++)
      if (C->KF.ci[k] != i) P->M.ci[w++] = C->KF.ci[k];
    P->M.rp[i + 1] = w;
  }
  P->work = malloc(4 * (C->ndof > nn ? C->ndof : nn) * sizeof(double));
  return P;
}

void crf_destroy(crf_prob *P) {
  if (!P) return;
  free(P->cptr); free(P->celem); free(P->kpos); free(P->iptr); free(P->inc);
  free(P->M.rp); free(P->M.ci); free(P->work);
  free(P);
}

static double crf_ke(const orc_ctx *C, int64_t e, double *KE) {
  if (C->KEc) {
    memcpy(KE, C->KEc + e * C->nd * C->nd, C->nd * C->nd * sizeof(double));
    return 1.0;
  }
  double X[24];
  for (int a = 0; a < C->npe; a++) {
    const int32_t n = C->elem[e * C->npe + a];
    X[3 * a] = C->xyz[3 * n]; X[3 * a + 1] = C->xyz[3 * n + 1]; X[3 * a + 2] = C->xyz[3 * n + 2];
  }
  return C->npe == 8 ? orc_hex8_ke(X, C->P.nu, KE) : orc_tet4_ke(X, C->P.nu, KE);
}

/* Jacobi-CG with the reference's KSPCG semantics (PDEFilter.cpp:117-137),
 * as the oracle's, with threaded kernels */
static int crf_cg_filter(orc_ctx *C, const double *b, double *x, double *wk) {
  const int64_t n = C->nn;
  double *r = wk, *z = wk + n, *p = wk + 2 * n, *q = wk + 3 * n;
  pmv(&C->KF, x, q);
  double snorm = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : snorm)
  for (int64_t i = 0; i < n; i++) {
    const double zb = b[i] / C->KFdiag[i];
    snorm += zb * zb;
    r[i] = b[i] - q[i];
    z[i] = r[i] / C->KFdiag[i];
  }
  snorm = sqrt(snorm);
  double dp = sqrt(pdot(n, z, z));
  if (snorm == 0.0) snorm = dp;
  double ttol = C->P.rtol_f * snorm;
  if (ttol < 1e-50) ttol = 1e-50;
  int it = 0;
  double beta = pdot(n, z, r), betaold = 0.0;
  if (!(dp <= ttol)) {
    for (it = 0; it < C->P.maxit_f;) {
      const double bb = it == 0 ? 0.0 : beta / betaold;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) p[i] = z[i] + bb * (it == 0 ? 0.0 : p[i]);
      pmv(&C->KF, p, q);
      const double al = beta / pdot(n, p, q);
      double zz = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : zz)
      for (int64_t i = 0; i < n; i++) {
        x[i] += al * p[i];
        r[i] -= al * q[i];
        z[i] = r[i] / C->KFdiag[i];
        zz += z[i] * z[i];
      }
      dp = sqrt(zz);
      it++;
      if (dp <= ttol || dp >= 1e5 * snorm) break;
      betaold = beta;
      beta = pdot(n, z, r);
    }
  }
  return it;
}

/* TX = TF x by node incidence, then RX = KF^-1 TX, then out = TFX^T RX */
static int crf_filter_apply(crf_prob *P, const double *x, double *out) {
  orc_ctx *C = P->C;
  const int npe = C->npe;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < C->nn; i++) {
    double s = 0.0;
    for (int64_t k = P->iptr[i]; k < P->iptr[i + 1]; k++) {
      const int32_t v = P->inc[k];
      s += C->Tel[v] * x[v / npe];
    }
    C->TX[i] = s;
  }
  const int it = crf_cg_filter(C, C->TX, C->RX, P->work);
  const double wgt = 1.0 / (double)npe;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < C->ne; e++) {
    double s = 0.0;
    for (int a = 0; a < npe; a++) s += wgt * C->RX[C->elem[e * npe + a]];
    out[e] = s;
  }
  return it;
}

/* One design iteration in the reference's structure.  times[0..6]: KE +
 * assembly, AMG setup, FGMRES solve, sensitivities, sensitivity filter, MMA,
 * density filter (seconds); iters[0] Krylov iterations, iters[1] filter CG
 * iterations of the three solves, iters[2] multigrid levels. */
int crf_design_iteration(crf_prob *P, int itr, double *fx, double *gx, double *ch, double *relres, double *times,
                         int *iters) {
  orc_ctx *C = P->C;
  const orc_params *Pp = &C->P;
  const int npe = C->npe, nd = C->nd;
  const int64_t ne = C->ne, ndof = C->ndof;
  double t0 = crf_clock();
  /* AssembleFem: KE recomputed for every element, E(x) KE scattered */
  memset(C->K.v, 0, C->K.rp[C->K.n] * sizeof(double));
  for (int c = 0; c < P->ncolor; c++) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = P->cptr[c]; q < P->cptr[c + 1]; q++) {
      const int64_t e = P->celem[q];
      double KE[576];
      crf_ke(C, e, KE);
      const double E = Pp->Emin + pow(C->XPhys[e], Pp->penal) * (Pp->Emax - Pp->Emin);
      const int64_t *kp = P->kpos + e * nd * nd;
      for (int k = 0; k < nd * nd; k++) C->K.v[kp[k]] += E * KE[k];
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < ndof; r++)
    for (int64_t k = C->K.rp[r]; k < C->K.rp[r + 1]; k++) {
      const int32_t cc = C->K.ci[k];
      if (C->isfixed[r] || C->isfixed[cc]) C->K.v[k] = (r == cc) ? 1.0 : 0.0;
    }
  memset(C->F, 0, ndof * sizeof(double));
  for (int lc = 0; lc < C->nlc; lc++)
    for (int i = 0; i < C->nload; i++) C->F[C->ldofs[lc * C->nload + i]] = C->lvals[lc * C->nload + i];
  for (int i = 0; i < C->nfixed; i++) C->F[C->fixed[i]] = 0.0;
  double t1 = crf_clock();
  times[0] = t1 - t0;
  /* ComputeFem: GAMG set up on the new matrix, FGMRES */
  crf_amg H;
  crf_amg_setup(&H, &C->K, &P->M, C->isfixed);
  double t2 = crf_clock();
  times[1] = t2 - t1;
  iters[2] = H.nlev;
  iters[0] = crf_fgmres(&H, &C->K, C->F, C->U, Pp->rtol_el, Pp->maxit_el, relres);
  crf_amg_free(&H);
  double t3 = crf_clock();
  times[2] = t3 - t2;
  /* ComputeObjectiveAndSensitivities: KE recomputed again */
  double f = 0.0, g = 0.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : f, g)
  for (int64_t e = 0; e < ne; e++) {
    double KE[576], ue[24];
    crf_ke(C, e, KE);
    for (int k = 0; k < nd; k++) ue[k] = C->U[3 * C->elem[e * npe + (k % npe)] + k / npe];
    double uKu = 0.0;
    for (int p = 0; p < nd; p++) {
      double s = 0.0;
      for (int q = 0; q < nd; q++) s += KE[p * nd + q] * ue[q];
      uKu += ue[p] * s;
    }
    f += (Pp->Emin + pow(C->XPhys[e], Pp->penal) * (Pp->Emax - Pp->Emin)) * uKu;
    C->dfdx[e] = -1.0 * Pp->penal * pow(C->XPhys[e], Pp->penal - 1) * (Pp->Emax - Pp->Emin) * uKu;
    g += C->XPhys[e] * C->vol[e];
    C->dgdx[e] = C->vol[e] / C->TV;
  }
  C->fx = f;
  C->gx[0] = g / C->TV - Pp->volfrac;
  double t4 = crf_clock();
  times[3] = t4 - t3;
  /* ComputeFilteredSens: forward filter operator on dfdx and dgdx, in place */
  int fit = crf_filter_apply(P, C->dfdx, C->dfdx);
  for (int j = 0; j < Pp->m; j++) fit += crf_filter_apply(P, C->dgdx + j * ne, C->dgdx + j * ne);
  double t5 = crf_clock();
  times[4] = t5 - t4;
  /* MMA */
  orc_mma_movelimit(C, C->X, C->xmin, C->xmax);
  orc_mma_update(C, C->X, C->dfdx, C->gx, C->dgdx, C->xmin, C->xmax);
  *ch = orc_design_change(C, C->X, C->Xold);
  double t6 = crf_clock();
  times[5] = t6 - t5;
  /* Filter + HeavisideFilter */
  fit += crf_filter_apply(P, C->X, C->XFiled);
  for (int64_t e = 0; e < ne; e++) {
    if (C->XFiled[e] < Pp->Xmin) C->XFiled[e] = Pp->Xmin;
    else if (C->XFiled[e] > 1.0) C->XFiled[e] = 1.0;
  }
  orc_heaviside(C, C->XFiled, C->XPhys, itr);
  times[6] = crf_clock() - t6;
  iters[1] = fit;
  *fx = C->fx;
  for (int j = 0; j < Pp->m; j++) gx[j] = C->gx[j];
  return *relres <= Pp->rtol_el ? 0 : 1;
}

int crf_threads(crf_prob *P) { return P->nthreads; }

/* Seconds to compute every element matrix once with the same threads (the
 * part of an iteration a KE-cached CPU code saves, twice per iteration). */
double crf_time_ke(crf_prob *P) {
  orc_ctx *C = P->C;
  const double t0 = crf_clock();
  double sink = 0.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : sink)
  for (int64_t e = 0; e < C->ne; e++) {
    double KE[576], X[24];
    for (int a = 0; a < C->npe; a++) {
      const int32_t n = C->elem[e * C->npe + a];
      X[3 * a] = C->xyz[3 * n]; X[3 * a + 1] = C->xyz[3 * n + 1]; X[3 * a + 2] = C->xyz[3 * n + 2];
    }
    if (C->npe == 8) orc_hex8_ke(X, C->P.nu, KE); else orc_tet4_ke(X, C->P.nu, KE);
    sink += KE[e % (C->nd * C->nd)];
  }
  const double t = crf_clock() - t0;
  return sink == sink ? t : -t;
}
