// gfx950 kernels of the aggregation multigrid V-cycle: Chebyshev-Jacobi
// smoothing, 6x6 block-sparse operators, transfer operators, the
// deterministic Galerkin products that rebuild the coarse operators for each
// new density field, and the dense coarsest-level inverse.
//
// Transfer: every unit u of level l (a node on level 0, an aggregate above)
// belongs to exactly one unit parent[u] of level l+1 with a bs x 6 block P_u
// (bs = 3 on level 0, 6 above), so
//   restriction  b_{l+1}[A] = sum_{u in children(A)} P_u^T r_l[u]   (fixed child order)
//   prolongation x_l[u]    += P_u x_{l+1}[parent(u)]
// Galerkin:  A_{l+1}[A,B] = sum P_a^T A_l[a,b] P_b over the precomputed,
// ordered contribution list of block (A,B): one wavefront per coarse block,
// lanes over contributions, fixed-order cross-lane reduction.  No atomics.
#include <type_traits>

#include "uft_internal.h"

namespace {

constexpr int TPB = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// ---------------------------------------------------------------- Chebyshev
// first step: r = b (x = 0) or r given; d = (1/theta) D^-1 r; x (+)= d
template <typename T>
__global__ __launch_bounds__(TPB) void k_cheb_first(int64_t n, const T *__restrict__ b, const T *__restrict__ dinv,
                                                    double inv_theta, T *__restrict__ r, T *__restrict__ d,
                                                    T *__restrict__ x, int x_is_zero) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  if (x_is_zero) {
    const T ri = b[i];
    r[i] = ri;
    const T di = (T)inv_theta * dinv[i] * ri;
    d[i] = di;
    x[i] = di;
  } else {
    const T di = (T)inv_theta * dinv[i] * r[i];
    d[i] = di;
    x[i] += di;
  }
}
// later steps: r -= t (t = A d); d = c1 d + c2 D^-1 r; x += d
template <typename T>
__global__ __launch_bounds__(TPB) void k_cheb_step(int64_t n, const T *__restrict__ t, const T *__restrict__ dinv,
                                                   double c1, double c2, T *__restrict__ r, T *__restrict__ d,
                                                   T *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const T ri = r[i] - t[i];
  r[i] = ri;
  const T di = (T)c1 * d[i] + (T)c2 * dinv[i] * ri;
  d[i] = di;
  x[i] += di;
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_sub_inplace(int64_t n, const T *__restrict__ t, T *__restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) r[i] -= t[i];
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_subv(int64_t n, const T *__restrict__ a, const T *__restrict__ b,
                                              T *__restrict__ c) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) c[i] = a[i] - b[i];
}

// ---------------------------------------------------------------- BSR 6x6
// LPR lanes per scalar row: y[6A+i] = sum_blocks sum_j V[blk][i][j] x[6B+j];
// lane l of a row takes blocks l, l+LPR, ...; the LPR partial sums are then
// combined with a fixed shuffle tree (deterministic).
constexpr int LPR = 8;
template <typename T>
__global__ __launch_bounds__(TPB) void k_bsr6_mv(int64_t nrows, const int32_t *__restrict__ rp,
                                                 const int32_t *__restrict__ ci, const T *__restrict__ val,
                                                 const T *__restrict__ x, T *__restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int64_t row = t / LPR;
  const int l = (int)(t - row * LPR);
  T acc = T(0);
  if (row < nrows * 6) {
    const int64_t A = row / 6;
    const int i = (int)(row - A * 6);
    for (int k = rp[A] + l; k < rp[A + 1]; k += LPR) {
      const int64_t B = ci[k];
      const T *v = val + (size_t)k * 36 + i * 6;
      const T *xb = x + B * 6;
#pragma unroll
      for (int j = 0; j < 6; ++j) acc += v[j] * xb[j];
    }
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, LPR);
  if (l == 0 && row < nrows * 6) y[row] = acc;
}

// zero diagonal entries of diagonal blocks (rank-deficient coarse columns,
// exactly decoupled) become 1 so the operator stays SPD
__global__ __launch_bounds__(TPB) void k_fixed_unit_diag(int64_t nrows, const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ ci, double *__restrict__ val) {
  const int64_t row = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (row >= nrows * 6) return;
  const int64_t A = row / 6;
  const int i = (int)(row - A * 6);
  for (int k = rp[A]; k < rp[A + 1]; ++k)
    if (ci[k] == A) {
      double *v = val + (size_t)k * 36 + i * 7;
      if (*v == 0.0) *v = 1.0;
    }
}

__global__ __launch_bounds__(TPB) void k_bsr6_dinv(int64_t nrows, const int32_t *__restrict__ rp,
                                                   const int32_t *__restrict__ ci, const double *__restrict__ val,
                                                   double *__restrict__ dinv, float *__restrict__ dinvf) {
  const int64_t row = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (row >= nrows * 6) return;
  const int64_t A = row / 6;
  const int i = (int)(row - A * 6);
  double d = 1.0;
  for (int k = rp[A]; k < rp[A + 1]; ++k)
    if (ci[k] == A) d = val[(size_t)k * 36 + i * 7];
  const double v = 1.0 / d;
  dinv[row] = v;
  if (dinvf) dinvf[row] = (float)v;
}

// ---------------------------------------------------------------- transfers
// one wavefront per coarse unit: lanes stride over the (child, row) pairs of
// the unit, then a fixed-order cross-lane sum of the 6 components
template <int BS, typename T>
__global__ __launch_bounds__(TPB) void k_restrict(int64_t nc, const int32_t *__restrict__ cptr,
                                                  const int32_t *__restrict__ child, const T *__restrict__ P,
                                                  const T *__restrict__ r, T *__restrict__ bc) {
  const int64_t A = ((int64_t)blockIdx.x * TPB + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (A >= nc) return;  // whole wavefront
  const int k0 = cptr[A], npair = (cptr[A + 1] - k0) * BS;
  T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int q = lane; q < npair; q += 64) {
    const int kk = q / BS, i = q - kk * BS;
    const int64_t u = child[k0 + kk];
    const T *Pu = P + (size_t)u * BS * 6 + i * 6;
    const T ri = r[u * BS + i];
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[j] += Pu[j] * ri;
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[j] += __shfl_down(acc[j], off, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < 6; ++j) bc[A * 6 + j] = acc[j];
  }
}
template <int BS, typename T>
__global__ __launch_bounds__(TPB) void k_prolong_add(int64_t nu, const int32_t *__restrict__ parent,
                                                     const T *__restrict__ P, const T *__restrict__ xc,
                                                     T *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= nu * BS) return;
  const int64_t u = t / BS;
  const int i = (int)(t - u * BS);
  const int64_t A = parent[u];
  if (A < 0) return;
  const T *Pu = P + (size_t)u * BS * 6 + i * 6;
  const T *xa = xc + A * 6;
  T acc = T(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) acc += Pu[j] * xa[j];
  x[t] += acc;
}

// ---------------------------------------------------------------- Galerkin
__device__ __forceinline__ int pk_index(int r, int c, int nd) {
  if (r > c) {
    int t = r;
    r = c;
    c = t;
  }
  return r * nd - (r * (r - 1)) / 2 + (c - r);
}

template <int K>
__device__ __forceinline__ void wave_sum(double (&v)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  }
}

// level 0 -> 1: contribution c = e*npe*npe + la*npe + lb adds
//   E(x_e) Q_{n(la)}^T KE_e[la-block, lb-block] Q_{n(lb)}   (Q: 3x6 per node)
template <int NPE>
__global__ __launch_bounds__(TPB) void k_galerkin_fine(int64_t nblk_c, int ne, const int32_t *__restrict__ glptr,
                                                       const int32_t *__restrict__ gl, const int32_t *__restrict__ elem,
                                                       const double *__restrict__ KE, const double *__restrict__ xphys,
                                                       SimpCoef sc, const double *__restrict__ Q,
                                                       double *__restrict__ out) {
  constexpr int ND = 3 * NPE;
  const int64_t blk = ((int64_t)blockIdx.x * TPB + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (blk >= nblk_c) return;  // whole wavefront exits together
  double acc[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) acc[k] = 0.0;
  const int k0 = glptr[blk], k1 = glptr[blk + 1];
  for (int k = k0 + lane; k < k1; k += 64) {
    const int c = gl[k];
    const int e = c / (NPE * NPE);
    const int rem = c - e * NPE * NPE;
    const int la = rem / NPE, lb = rem - la * NPE;
    const int na = elem[(size_t)la * ne + e], nb = elem[(size_t)lb * ne + e];
    const double E = sc.Emin + pow(xphys[e], sc.penal) * sc.dE;
    double Kb[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int p = pk_index(r * NPE + la, s * NPE + lb, ND);
        Kb[r][s] = E * KE[((size_t)(p >> 1) * ne + e) * 2 + (p & 1)];
      }
    const double *Qa = Q + (size_t)na * 18, *Qb = Q + (size_t)nb * 18;
    double W[3][6];  // Kb Qb
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int j = 0; j < 6; ++j) W[r][j] = Kb[r][0] * Qb[j] + Kb[r][1] * Qb[6 + j] + Kb[r][2] * Qb[12 + j];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) acc[i * 6 + j] += Qa[i] * W[0][j] + Qa[6 + i] * W[1][j] + Qa[12 + i] * W[2][j];
  }
  wave_sum<36>(acc);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 36; ++k) out[(size_t)blk * 36 + k] = acc[k];
  }
}

// level l -> l+1 (l >= 1): contribution = block index kb of level l with row
// brow[kb] and column ci[kb]; adds P_a^T A_l[kb] P_b (P: 6x6 per unit)
__global__ __launch_bounds__(TPB) void k_galerkin_coarse(int64_t nblk_c, const int32_t *__restrict__ glptr,
                                                         const int32_t *__restrict__ gl, const int32_t *__restrict__ brow,
                                                         const int32_t *__restrict__ ci, const double *__restrict__ P,
                                                         const double *__restrict__ val, double *__restrict__ out) {
  const int64_t blk = ((int64_t)blockIdx.x * TPB + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (blk >= nblk_c) return;
  double acc[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) acc[k] = 0.0;
  const int k0 = glptr[blk], k1 = glptr[blk + 1];
  for (int k = k0 + lane; k < k1; k += 64) {
    const int kb = gl[k];
    const double *Pa = P + (size_t)brow[kb] * 36, *Pb = P + (size_t)ci[kb] * 36;
    const double *A = val + (size_t)kb * 36;
    double W[6][6];  // A Pb
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) s += A[r * 6 + q] * Pb[q * 6 + j];
        W[r][j] = s;
      }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < 6; ++r) s += Pa[r * 6 + i] * W[r][j];
        acc[i * 6 + j] += s;
      }
  }
  wave_sum<36>(acc);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 36; ++k) out[(size_t)blk * 36 + k] = acc[k];
  }
}

// ---------------------------------------------------------------- dense coarsest level
__global__ __launch_bounds__(TPB) void k_bsr_to_dense(int64_t nrows, const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ ci, const double *__restrict__ val,
                                                      int64_t n, double *__restrict__ D) {
  const int64_t row = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (row >= nrows * 6) return;
  const int64_t A = row / 6;
  const int i = (int)(row - A * 6);
  double *Dr = D + row * n;
  for (int64_t c = 0; c < n; ++c) Dr[c] = 0.0;
  for (int k = rp[A]; k < rp[A + 1]; ++k) {
    const int64_t B = ci[k];
#pragma unroll
    for (int j = 0; j < 6; ++j) Dr[B * 6 + j] = val[(size_t)k * 36 + i * 6 + j];
  }
}

// In-place Gauss-Jordan inversion of an SPD matrix (no pivoting needed): for
// each pivot k the pivot row is scaled and every other row is eliminated in
// parallel; the pivot column is staged in LDS first.  One workgroup.
__global__ __launch_bounds__(1024) void k_dense_gj_inv(int n, double *__restrict__ A, double *__restrict__ Ainv,
                                                       float *__restrict__ Ainvf) {
  extern __shared__ double lds[];  // pivot column (n) and scaled pivot row (n)
  double *colk = lds, *rowk = lds + n;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = nt >> 6;
  for (int i = wave; i < n; i += nwave)
    for (int j = lane; j < n; j += 64) Ainv[(size_t)i * n + j] = A[(size_t)i * n + j];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    for (int i = tid; i < n; i += nt) colk[i] = Ainv[(size_t)i * n + k];
    __syncthreads();
    double piv = colk[k];
    if (!(piv > 0.0)) piv = 1.0;  // decoupled zero row/column of a rank-deficient aggregate
    const double ip = 1.0 / piv;
    for (int j = tid; j < n; j += nt) {
      const double v = (j == k ? 1.0 : Ainv[(size_t)k * n + j]) * ip;
      rowk[j] = v;
      Ainv[(size_t)k * n + j] = v;
    }
    __syncthreads();
    for (int i = wave; i < n; i += nwave) {
      if (i == k) continue;
      const double f = colk[i];
      double *Ai = Ainv + (size_t)i * n;
      for (int j = lane; j < n; j += 64) Ai[j] = (j == k ? 0.0 : Ai[j]) - f * rowk[j];
    }
    __syncthreads();
  }
  if (Ainvf)
    for (int64_t t = tid; t < (int64_t)n * n; t += nt) Ainvf[t] = (float)Ainv[t];
}

// In-place Cholesky A = L L^T (lower triangle), then Ainv = L^-T L^-1.
// One workgroup of 1024 threads; n is at most a few thousand.
__global__ __launch_bounds__(1024) void k_dense_chol_inv(int n, double *__restrict__ A, double *__restrict__ W,
                                                         double *__restrict__ Ainv, float *__restrict__ Ainvf) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      double akk = A[(size_t)k * n + k];
      A[(size_t)k * n + k] = akk > 0.0 ? sqrt(akk) : 1.0;
    }
    __syncthreads();
    const double lkk = A[(size_t)k * n + k];
    for (int i = k + 1 + tid; i < n; i += nt) A[(size_t)i * n + k] /= lkk;
    __syncthreads();
    const int m = n - k - 1;
    for (int64_t t = tid; t < (int64_t)m * m; t += nt) {
      const int i = k + 1 + (int)(t / m), j = k + 1 + (int)(t % m);
      if (j <= i) A[(size_t)i * n + j] -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
    }
    __syncthreads();
  }
  // W = L^-1, column by column (each thread owns columns)
  for (int c = tid; c < n; c += nt) {
    for (int i = 0; i < n; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) s -= A[(size_t)i * n + k] * W[(size_t)k * n + c];
      W[(size_t)i * n + c] = (i < c) ? 0.0 : s / A[(size_t)i * n + i];
    }
  }
  __syncthreads();
  // Ainv = W^T W
  for (int64_t t = tid; t < (int64_t)n * n; t += nt) {
    const int i = (int)(t / n), j = (int)(t % n);
    const int k0 = i > j ? i : j;
    double s = 0.0;
    for (int k = k0; k < n; ++k) s += W[(size_t)k * n + i] * W[(size_t)k * n + j];
    Ainv[t] = s;
    if (Ainvf) Ainvf[t] = (float)s;
  }
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_dense_mv(int64_t n, const T *__restrict__ A, const T *__restrict__ x,
                                                  T *__restrict__ y) {
  // one wavefront per row
  const int64_t row = ((int64_t)blockIdx.x * TPB + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double acc = 0.0;
  for (int64_t c = lane; c < n; c += 64) acc += (double)A[row * n + c] * (double)x[c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) y[row] = (T)acc;
}

}  // namespace

namespace uftk {

void cheb_first(hipStream_t s, int64_t n, const double *b, const double *dinv, double inv_theta, double *r, double *d,
                double *x, int x_is_zero) {
  k_cheb_first<double><<<nblk(n), TPB, 0, s>>>(n, b, dinv, inv_theta, r, d, x, x_is_zero);
}
void cheb_step(hipStream_t s, int64_t n, const double *t, const double *dinv, double c1, double c2, double *r,
               double *d, double *x) {
  k_cheb_step<double><<<nblk(n), TPB, 0, s>>>(n, t, dinv, c1, c2, r, d, x);
}
void cheb_firstf(hipStream_t s, int64_t n, const float *b, const float *dinv, double inv_theta, float *r, float *d,
                 float *x, int x_is_zero) {
  k_cheb_first<float><<<nblk(n), TPB, 0, s>>>(n, b, dinv, inv_theta, r, d, x, x_is_zero);
}
void cheb_stepf(hipStream_t s, int64_t n, const float *t, const float *dinv, double c1, double c2, float *r, float *d,
                float *x) {
  k_cheb_step<float><<<nblk(n), TPB, 0, s>>>(n, t, dinv, c1, c2, r, d, x);
}
void sub_inplace(hipStream_t s, int64_t n, const double *t, double *r) { k_sub_inplace<double><<<nblk(n), TPB, 0, s>>>(n, t, r); }
void sub_inplacef(hipStream_t s, int64_t n, const float *t, float *r) { k_sub_inplace<float><<<nblk(n), TPB, 0, s>>>(n, t, r); }
void subf(hipStream_t s, int64_t n, const float *a, const float *b, float *c) {
  k_subv<float><<<nblk(n), TPB, 0, s>>>(n, a, b, c);
}
void bsr6_mv(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const double *val, const double *x,
             double *y) {
  k_bsr6_mv<double><<<nblk(nb_rows * 6 * LPR), TPB, 0, s>>>(nb_rows, rp, ci, val, x, y);
}
void bsr6_mvf(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const float *val, const float *x,
              float *y) {
  k_bsr6_mv<float><<<nblk(nb_rows * 6 * LPR), TPB, 0, s>>>(nb_rows, rp, ci, val, x, y);
}
void bsr6_dinv(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, double *val, double *dinv,
               float *dinvf) {
  k_bsr6_dinv<<<nblk(nb_rows * 6), TPB, 0, s>>>(nb_rows, rp, ci, val, dinv, dinvf);
}
void fixed_unit_diag(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, double *val) {
  k_fixed_unit_diag<<<nblk(nb_rows * 6), TPB, 0, s>>>(nb_rows, rp, ci, val);
}
void restrict_(hipStream_t s, int bs, int64_t nc, const int32_t *cptr, const int32_t *child, const double *P,
               const double *r, double *bc) {
  if (bs == 3)
    k_restrict<3, double><<<nblk(nc * 64), TPB, 0, s>>>(nc, cptr, child, P, r, bc);
  else
    k_restrict<6, double><<<nblk(nc * 64), TPB, 0, s>>>(nc, cptr, child, P, r, bc);
}
void prolong_add(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const double *P, const double *xc,
                 double *x) {
  if (bs == 3)
    k_prolong_add<3, double><<<nblk(nu * 3), TPB, 0, s>>>(nu, parent, P, xc, x);
  else
    k_prolong_add<6, double><<<nblk(nu * 6), TPB, 0, s>>>(nu, parent, P, xc, x);
}
void restrictf(hipStream_t s, int bs, int64_t nc, const int32_t *cptr, const int32_t *child, const float *P,
               const float *r, float *bc) {
  if (bs == 3)
    k_restrict<3, float><<<nblk(nc * 64), TPB, 0, s>>>(nc, cptr, child, P, r, bc);
  else
    k_restrict<6, float><<<nblk(nc * 64), TPB, 0, s>>>(nc, cptr, child, P, r, bc);
}
void prolong_addf(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const float *P, const float *xc,
                  float *x) {
  if (bs == 3)
    k_prolong_add<3, float><<<nblk(nu * 3), TPB, 0, s>>>(nu, parent, P, xc, x);
  else
    k_prolong_add<6, float><<<nblk(nu * 6), TPB, 0, s>>>(nu, parent, P, xc, x);
}
void galerkin_fine(hipStream_t s, const uft_ctx *C, const double *xphys, int64_t nb, const int32_t *glptr,
                   const int32_t *gl, const double *Q, double *val) {
  SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
  const unsigned g = (unsigned)((nb * 64 + TPB - 1) / TPB);
  if (C->npe == 8)
    k_galerkin_fine<8><<<g, TPB, 0, s>>>(nb, (int)C->ne, glptr, gl, C->d_elem, C->d_KE, xphys, sc, Q, val);
  else
    k_galerkin_fine<4><<<g, TPB, 0, s>>>(nb, (int)C->ne, glptr, gl, C->d_elem, C->d_KE, xphys, sc, Q, val);
}
void galerkin_coarse(hipStream_t s, int64_t nb, const int32_t *glptr, const int32_t *gl, const int32_t *brow,
                     const int32_t *ci, const double *P, const double *val, double *valc) {
  const unsigned g = (unsigned)((nb * 64 + TPB - 1) / TPB);
  k_galerkin_coarse<<<g, TPB, 0, s>>>(nb, glptr, gl, brow, ci, P, val, valc);
}
void bsr_to_dense(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const double *val, int64_t n,
                  double *dense) {
  k_bsr_to_dense<<<nblk(nb_rows * 6), TPB, 0, s>>>(nb_rows, rp, ci, val, n, dense);
}
void dense_spd_inverse(hipStream_t s, int64_t n, double *A, double *Ainv, float *Ainvf) {
  k_dense_gj_inv<<<1, 1024, 2 * (size_t)n * sizeof(double), s>>>((int)n, A, Ainv, Ainvf);
}
void dense_spd_inverse_chol(hipStream_t s, int64_t n, double *A, double *Ainv, float *Ainvf) {
  // W (n x n) lives right after A in the same allocation
  k_dense_chol_inv<<<1, 1024, 0, s>>>((int)n, A, A + n * n, Ainv, Ainvf);
}
void dense_mv(hipStream_t s, int64_t n, const double *A, const double *x, double *y) {
  k_dense_mv<double><<<(unsigned)((n * 64 + TPB - 1) / TPB), TPB, 0, s>>>(n, A, x, y);
}
void dense_mvf(hipStream_t s, int64_t n, const float *A, const float *x, float *y) {
  k_dense_mv<float><<<(unsigned)((n * 64 + TPB - 1) / TPB), TPB, 0, s>>>(n, A, x, y);
}

}  // namespace uftk
