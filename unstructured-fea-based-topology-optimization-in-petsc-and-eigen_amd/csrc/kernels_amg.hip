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
#include <hip/hip_fp16.h>

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

// level 0 -> 1, once per mesh: entry k (element e, upper block (A,B)) gets
//   G_k = sum_{la: agg(n_la) = A} sum_{lb: agg(n_lb) = B} Q_{n_la}^T KE_e[la-block, lb-block] Q_{n_lb}
// (Q: 3x6 per node, KE_e the density-free element matrix).  One thread per entry.
template <int NPE>
__global__ __launch_bounds__(TPB) void k_galerkin_pre(int64_t nent, int ne, const int32_t *__restrict__ gblk,
                                                      const int32_t *__restrict__ gl, const int32_t *__restrict__ brow,
                                                      const int32_t *__restrict__ ci, const int32_t *__restrict__ parent,
                                                      const int32_t *__restrict__ elem, const double *__restrict__ KE,
                                                      const double *__restrict__ Q, double *__restrict__ G) {
  constexpr int ND = 3 * NPE;
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k >= nent) return;
  const int e = gl[k], kb = gblk[k];
  const int A = brow[kb], B = ci[kb];
  double acc[36];
#pragma unroll
  for (int q = 0; q < 36; ++q) acc[q] = 0.0;
  for (int la = 0; la < NPE; ++la) {
    const int na = elem[(size_t)la * ne + e];
    if (parent[na] != A) continue;
    const double *Qa = Q + (size_t)na * 18;
    for (int lb = 0; lb < NPE; ++lb) {
      const int nb = elem[(size_t)lb * ne + e];
      if (parent[nb] != B) continue;
      double Kb[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int p = pk_index(r * NPE + la, c * NPE + lb, ND);
          Kb[r][c] = KE[((size_t)(p >> 1) * ne + e) * 2 + (p & 1)];
        }
      const double *Qb = Q + (size_t)nb * 18;
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
  }
#pragma unroll
  for (int q = 0; q < 36; ++q) G[(size_t)k * 36 + q] = acc[q];
}

__global__ __launch_bounds__(TPB) void k_simp_E(int64_t ne, const double *__restrict__ x, SimpCoef sc,
                                                double *__restrict__ Ee) {
  const int64_t e = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (e < ne) Ee[e] = sc.Emin + pow(x[e], sc.penal) * sc.dE;
}

// level 0 -> 1 per density field: one thread per entry (kb, q) of the blocks
// on or above the diagonal, which also writes the entry into the transposed
// block below the diagonal (so off-diagonal blocks are exact transposes of
// each other and the entries are read once).  Fixed entry order.
__global__ __launch_bounds__(TPB) void k_galerkin_fine(int64_t nblk_c, const int32_t *__restrict__ glptr,
                                                       const int32_t *__restrict__ gl, const int32_t *__restrict__ gtr,
                                                       const int32_t *__restrict__ brow, const int32_t *__restrict__ ci,
                                                       const double *__restrict__ Ee, const double *__restrict__ G,
                                                       double *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= nblk_c * 36) return;
  const int64_t kb = t / 36;
  const int q = (int)(t - kb * 36);
  if (brow[kb] > ci[kb]) return;  // a lower block is written by its upper transpose below
  double acc = 0.0;
  for (int k = glptr[kb]; k < glptr[kb + 1]; ++k) acc += Ee[gl[k]] * G[(size_t)k * 36 + q];
  out[t] = acc;
  const int64_t kt = gtr[kb];
  if (kt != kb) out[(size_t)kt * 36 + (q % 6) * 6 + q / 6] = acc;  // the entries are read once
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

// Blocked Gauss-Jordan inversion of an SPD matrix (no pivoting needed: every
// pivot block is a Schur complement of an SPD matrix).  For each block K of
// nb <= GJ_NB consecutive indices:
//   Dinv = M[K,K]^-1;  R = Dinv M[K,:];  Cp = M[:,K]      (k_gj_diag, k_gj_panels)
//   M[i,j] -= Cp[i,:] R[:,j]   (i, j not in K)              (k_gj_update)
//   M[i,K]  = -Cp[i,:] Dinv,  M[K,j] = R[:,j],  M[K,K] = Dinv
// after the last block M holds the inverse.  Three launches per block, all
// entries updated in parallel, panels in a separate scratch area so that no
// kernel reads what another workgroup of the same launch writes.
constexpr int GJ_NB = 32;

constexpr int GJ_DIAG_T = TPB;
__global__ __launch_bounds__(GJ_DIAG_T) void k_gj_diag(int n, int k0, int nb, const double *__restrict__ M,
                                                       double *__restrict__ Dinv) {
  __shared__ double D[GJ_NB * GJ_NB];
  __shared__ double colk[GJ_NB], rowk[GJ_NB];
  const int tid = threadIdx.x;
  for (int t = tid; t < nb * nb; t += GJ_DIAG_T) D[t] = M[(size_t)(k0 + t / nb) * n + k0 + t % nb];
  __syncthreads();
  for (int k = 0; k < nb; ++k) {
    if (tid < nb) colk[tid] = D[tid * nb + k];
    __syncthreads();
    double piv = colk[k];
    if (!(piv > 0.0)) piv = 1.0;  // decoupled zero row/column of a rank-deficient aggregate
    const double ip = 1.0 / piv;
    if (tid < nb) {
      const double v = (tid == k ? 1.0 : D[k * nb + tid]) * ip;
      rowk[tid] = v;
      D[k * nb + tid] = v;
    }
    __syncthreads();
    for (int t = tid; t < nb * nb; t += GJ_DIAG_T) {
      const int i = t / nb, j = t - (t / nb) * nb;
      if (i != k) D[t] = (j == k ? 0.0 : D[t]) - colk[i] * rowk[j];
    }
    __syncthreads();
  }
  for (int t = tid; t < nb * nb; t += GJ_DIAG_T) Dinv[t] = D[t];
}

// thread (l, j), l < nb: R[l, j] = (Dinv M[K, j])_l (j not in K) and
// Cp[j, l] = M[j, k0 + l]
__global__ __launch_bounds__(TPB) void k_gj_panels(int n, int k0, int nb, const double *__restrict__ M,
                                                   const double *__restrict__ Dinv, double *__restrict__ R,
                                                   double *__restrict__ Cp) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= (int64_t)nb * n) return;
  const int l = (int)(t / n), j = (int)(t - (int64_t)l * n);
  Cp[(size_t)j * nb + l] = M[(size_t)j * n + k0 + l];
  if (j >= k0 && j < k0 + nb) return;
  double acc = 0.0;
  for (int m = 0; m < nb; ++m) acc += Dinv[l * nb + m] * M[(size_t)(k0 + m) * n + j];
  R[(size_t)l * n + j] = acc;
}

// one row i per blockIdx.y, threads over columns j
__global__ __launch_bounds__(TPB) void k_gj_update(int n, int k0, int nb, double *__restrict__ M,
                                                   const double *__restrict__ Dinv, const double *__restrict__ R,
                                                   const double *__restrict__ Cp) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * TPB + threadIdx.x;
  if (j >= n) return;
  const bool iK = i >= k0 && i < k0 + nb, jK = j >= k0 && j < k0 + nb;
  double *Mij = M + (size_t)i * n + j;
  if (iK) {
    *Mij = jK ? Dinv[(i - k0) * nb + (j - k0)] : R[(size_t)(i - k0) * n + j];
    return;
  }
  const double *ci = Cp + (size_t)i * nb;
  double acc = 0.0;
  if (jK) {
    for (int l = 0; l < nb; ++l) acc += ci[l] * Dinv[l * nb + (j - k0)];
    *Mij = -acc;
  } else {
    for (int l = 0; l < nb; ++l) acc += ci[l] * R[(size_t)l * n + j];
    *Mij -= acc;
  }
}

__global__ __launch_bounds__(TPB) void k_copy_n(int64_t n, const double *__restrict__ a, double *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) b[i] = a[i];
}
// S = (S + S^T) / 2 in place: the elimination's rounding leaves the inverse
// of a symmetric matrix only nearly symmetric; exact symmetry lets the solve
// read it by columns.  One thread per (i < j) pair.
__global__ __launch_bounds__(TPB) void k_symmetrize(int64_t n, double *__restrict__ S) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int64_t i = t / n, j = t - i * n;
  if (i >= n || j <= i) return;
  const double v = 0.5 * (S[i * n + j] + S[j * n + i]);
  S[i * n + j] = v;
  S[j * n + i] = v;
}
__global__ __launch_bounds__(TPB) void k_d2f_n(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
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

namespace {
// FP16 copy of a level's operator scaled to unit diagonal: Ah_ij =
// a_ij / sqrt(a_ii a_jj) (|Ah_ij| <= 1 for an SPD matrix), and the FP32 row
// scales dh_i = sqrt(a_ii) from the inverse diagonal.
__global__ __launch_bounds__(256) void k_bsr6_to_half(int64_t nblk, const int32_t *__restrict__ brow,
                                                     const int32_t *__restrict__ ci, const double *__restrict__ val,
                                                     const double *__restrict__ dinv, __half *__restrict__ valh) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nblk * 36) return;
  const int64_t k = t / 36;
  const int e = (int)(t - k * 36), i = e / 6, j = e - (e / 6) * 6;
  const int64_t A = brow[k], B = ci[k];
  valh[t] = __float2half_rn((float)(val[t] * sqrt(dinv[A * 6 + i] * dinv[B * 6 + j])));
}
__global__ __launch_bounds__(256) void k_sqrt_diag(int64_t n, const double *__restrict__ dinv, float *__restrict__ dh) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < n) dh[r] = (float)(1.0 / sqrt(dinv[r]));
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
void bsr6_to_half(hipStream_t s, int64_t nblk, int64_t nb_rows, const int32_t *brow, const int32_t *ci,
                  const double *val, const double *dinv, uint16_t *valh, float *dh) {
  k_bsr6_to_half<<<(unsigned)((nblk * 36 + 255) / 256), 256, 0, s>>>(nblk, brow, ci, val, dinv,
                                                                    reinterpret_cast<__half *>(valh));
  k_sqrt_diag<<<(unsigned)((nb_rows * 6 + 255) / 256), 256, 0, s>>>(nb_rows * 6, dinv, dh);
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
void galerkin_pre(hipStream_t s, const uft_ctx *C, int64_t ngl, const int32_t *gblk, const int32_t *gl,
                  const int32_t *brow, const int32_t *ci, const int32_t *parent, const double *Q, double *G) {
  if (C->npe == 8)
    k_galerkin_pre<8><<<nblk(ngl), TPB, 0, s>>>(ngl, (int)C->ne, gblk, gl, brow, ci, parent, C->d_elem, C->d_KE, Q, G);
  else
    k_galerkin_pre<4><<<nblk(ngl), TPB, 0, s>>>(ngl, (int)C->ne, gblk, gl, brow, ci, parent, C->d_elem, C->d_KE, Q, G);
}
void simp_moduli(hipStream_t s, const uft_ctx *C, const double *xphys, double *Ee) {
  SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
  k_simp_E<<<nblk(C->ne), TPB, 0, s>>>(C->ne, xphys, sc, Ee);
}
void galerkin_fine(hipStream_t s, const uft_ctx *C, const double *xphys, int64_t nb, const int32_t *glptr,
                   const int32_t *gl, const int32_t *gtr, const int32_t *brow, const int32_t *ci, const double *G,
                   double *Ee, double *val) {
  if (xphys) {  // else Ee already holds E(x_e)
    SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
    k_simp_E<<<nblk(C->ne), TPB, 0, s>>>(C->ne, xphys, sc, Ee);
  }
  k_galerkin_fine<<<nblk(nb * 36), TPB, 0, s>>>(nb, glptr, gl, gtr, brow, ci, Ee, G, val);
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
// A (2 n^2 doubles): A[0, n^2) holds the matrix on entry, the rest is scratch
// for the panels; the inverse is written to Ainv (and to Ainvf when given).
void dense_spd_inverse(hipStream_t s, int64_t n, double *A, double *Ainv, float *Ainvf) {
  if (n >= UFT_SWEEP_MIN) {  // large coarsest levels: the tiled sweep (scratch after the matrix)
    dense_spd_sweep(s, n, A, A + n * n, Ainv, Ainvf);
    return;
  }
  const int nn = (int)n;
  const int nb = std::max(1, std::min(GJ_NB, nn / 4));
  k_copy_n<<<nblk(n * n), TPB, 0, s>>>(n * n, A, Ainv);
  double *Dinv = A + n * n, *R = Dinv + nb * nb, *Cp = R + (size_t)nb * n;
  for (int k0 = 0; k0 < nn; k0 += nb) {
    const int b = std::min(nb, nn - k0);
    k_gj_diag<<<1, GJ_DIAG_T, 0, s>>>(nn, k0, b, Ainv, Dinv);
    k_gj_panels<<<nblk((int64_t)b * n), TPB, 0, s>>>(nn, k0, b, Ainv, Dinv, R, Cp);
    k_gj_update<<<dim3(nblk(n), (unsigned)n), TPB, 0, s>>>(nn, k0, b, Ainv, Dinv, R, Cp);
  }
  k_symmetrize<<<nblk(n * n), TPB, 0, s>>>(n, Ainv);
  if (Ainvf) k_d2f_n<<<nblk(n * n), TPB, 0, s>>>(n * n, Ainv, Ainvf);
}
void dense_mv(hipStream_t s, int64_t n, const double *A, const double *x, double *y) {
  k_dense_mv<double><<<(unsigned)((n * 64 + TPB - 1) / TPB), TPB, 0, s>>>(n, A, x, y);
}
void dense_mvf(hipStream_t s, int64_t n, const float *A, const float *x, float *y) {
  k_dense_mv<float><<<(unsigned)((n * 64 + TPB - 1) / TPB), TPB, 0, s>>>(n, A, x, y);
}

}  // namespace uftk
