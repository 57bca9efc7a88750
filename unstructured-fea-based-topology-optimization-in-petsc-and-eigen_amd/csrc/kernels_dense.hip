// gfx950 kernels of the dense inverse of the coarsest multigrid operator.
//
// The coarsest level is solved exactly with its inverse, applied as one
// dense matrix-vector product per visit (the product reads the inverse once:
// 59 MB in FP32 for 3,840 unknowns).  The inverse is recomputed at every
// hierarchy refresh, so it has to be fast for a few thousand unknowns:
// a blocked symmetric Gauss-Jordan sweep (Goodnight 1979) on 64x64 tiles of
// the lower triangle.  Sweeping pivot block K of the symmetric matrix M
// (D = M_KK, SPD because it is a Schur complement of an SPD matrix):
//   M_KK <- -D^-1,  M_iK <- M_iK D^-1,  M_ij <- M_ij - M_iK D^-1 M_Kj
// for i, j outside K; after every block has been swept M = -A^-1.  Per
// block two launches: the panel (C_i = M_iK and W_i = C_i D^-1 into
// buffers, so the update reads no tile another workgroup writes) and the
// update of every lower tile (W_i C_j^T, one 64x64x64 product per tile on
// the FP64 matrix cores), whose workgroup for the next pivot tile also
// inverts it.  No pivoting and
// no atomics: every sum runs in a fixed order, results are reproducible.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "uft_internal.h"

namespace {

constexpr int DB = 64;           // tile edge
constexpr int DT = DB * DB;      // entries per tile
constexpr int DTPB = 256;        // threads per workgroup (the update: four wavefronts, a 32x32 quarter each)
constexpr int LDP = DB + 4;      // LDS row pitch of transposed operands (doubles)
constexpr int LDQ = DB + 2;      // LDS row pitch of the update's row-major operands (doubles)
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int64_t tile_off(int i, int j) { return ((int64_t)i * (i + 1) / 2 + j) * DT; }

// In-place Gauss-Jordan inverse of the SPD 64x64 tile D (LDS, row-major) by
// one workgroup of DTPB threads; a non-positive pivot (a decoupled unknown)
// is taken as 1.  Thread t keeps entries (i, t % 64), i = t / 64 + 4u, in
// registers for all 64 pivots; per pivot only the pivot column and row go
// through LDS (cb, rb: two buffers of 64, so one barrier per pivot: a
// buffer is rewritten two pivots later, after everyone has passed the
// barrier of the pivot in between).  The caller synchronises before the
// call (D written) and after it (D read by other threads).
__device__ void tile_inverse(double *D, double *cb, double *rb) {
  const int tid = threadIdx.x, j = tid & 63, i0 = tid >> 6;
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) v[u] = D[(i0 + 4 * u) * DB + j];
#pragma unroll
  for (int k = 0; k < DB; ++k) {
    double *ck = cb + (k & 1) * DB, *rk = rb + (k & 1) * DB;
    if (j == k) {
#pragma unroll
      for (int u = 0; u < 16; ++u) ck[i0 + 4 * u] = v[u];
    }
    if (i0 == (k & 3)) rk[j] = v[k >> 2];
    __syncthreads();
    double piv = ck[k];
    if (!(piv > 0.0)) piv = 1.0;
    const double r = (j == k ? 1.0 : rk[j]) / piv;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int i = i0 + 4 * u;
      v[u] = (i == k) ? r : ((j == k ? 0.0 : v[u]) - ck[i] * r);
    }
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) D[(i0 + 4 * u) * DB + j] = v[u];
}

// lower tiles of the padded matrix from the row-major dense n x n matrix A
// (padding: unit diagonal)
__global__ __launch_bounds__(DTPB) void k_sweep_pack(int64_t n, int nt, const double *__restrict__ A,
                                                     double *__restrict__ T) {
  const int64_t q = (int64_t)blockIdx.x * DTPB + threadIdx.x;
  const int64_t ntile = (int64_t)nt * (nt + 1) / 2;
  if (q >= ntile * DT) return;
  const int64_t t = q / DT;
  const int e = (int)(q % DT);
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((int64_t)(i + 1) * (i + 2) / 2 <= t) ++i;
  while ((int64_t)i * (i + 1) / 2 > t) --i;
  const int j = (int)(t - (int64_t)i * (i + 1) / 2);
  const int64_t r = (int64_t)i * DB + e / DB, c = (int64_t)j * DB + e % DB;
  T[q] = (r < n && c < n) ? A[r * n + c] : (r == c ? 1.0 : 0.0);
}

// lower tiles from a 6x6-block sparse matrix (block rows brow, columns ci,
// values row-major per block): k_sweep_zero clears them (unit diagonal on
// the padding), k_sweep_scatter writes every entry of the tiles on or below
// the diagonal
__global__ __launch_bounds__(DTPB) void k_sweep_zero(int64_t n, int nt, double *__restrict__ T) {
  const int64_t q = (int64_t)blockIdx.x * DTPB + threadIdx.x;
  const int64_t ntile = (int64_t)nt * (nt + 1) / 2;
  if (q >= ntile * DT) return;
  const int64_t t = q / DT;
  const int e = (int)(q % DT);
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((int64_t)(i + 1) * (i + 2) / 2 <= t) ++i;
  while ((int64_t)i * (i + 1) / 2 > t) --i;
  const int j = (int)(t - (int64_t)i * (i + 1) / 2);
  const int64_t r = (int64_t)i * DB + e / DB, c = (int64_t)j * DB + e % DB;
  T[q] = (r == c && r >= n) ? 1.0 : 0.0;
}
__global__ __launch_bounds__(DTPB) void k_sweep_scatter(int64_t nblk, const int32_t *__restrict__ brow,
                                                        const int32_t *__restrict__ ci, const double *__restrict__ val,
                                                        double *__restrict__ T) {
  const int64_t q = (int64_t)blockIdx.x * DTPB + threadIdx.x;
  if (q >= nblk * 36) return;
  const int64_t b = q / 36;
  const int e = (int)(q % 36);
  const int64_t r = 6 * (int64_t)brow[b] + e / 6, c = 6 * (int64_t)ci[b] + e % 6;
  if (r / DB < c / DB) return;  // tiles above the diagonal (a diagonal tile is stored whole)
  T[tile_off((int)(r / DB), (int)(c / DB)) + (r % DB) * DB + c % DB] = val[q];
}

// the inverse of tile (0, 0), the first pivot
__global__ __launch_bounds__(DTPB) void k_sweep_first(const double *__restrict__ T, double *__restrict__ Dout) {
  __shared__ double D[DT];
  __shared__ double colk[2 * DB], rowk[2 * DB];
  for (int t = threadIdx.x; t < DT; t += DTPB) D[t] = T[t];
  __syncthreads();
  tile_inverse(D, colk, rowk);
  __syncthreads();
  for (int t = threadIdx.x; t < DT; t += DTPB) Dout[t] = D[t];
}

// panel of pivot block k, workgroup i != k: C_i = M_iK (tile (i, k), or tile
// (k, i) transposed for i < k) and W_i = C_i D^-1 (D^-1 symmetric)
__global__ __launch_bounds__(DTPB) void k_sweep_panel(int k, const double *__restrict__ T,
                                                      const double *__restrict__ Dk, double *__restrict__ Wb,
                                                      double *__restrict__ Cb) {
  const int i = blockIdx.x;
  if (i == k) return;
  __shared__ double Ct[DB * LDP];  // C_i transposed: Ct[m][r] = C_i[r][m]
  __shared__ double Ds[DB * LDP];  // D^-1 rows: Ds[m][c]
  const int tid = threadIdx.x;
  const double *src = i > k ? T + tile_off(i, k) : T + tile_off(k, i);
  {  // all loads in flight before the stores
    double cv[DT / DTPB], dv[DT / DTPB];
#pragma unroll
    for (int u = 0; u < DT / DTPB; ++u) {
      const int t = tid + u * DTPB, r = t / DB, m = t % DB;
      cv[u] = i > k ? src[t] : src[m * DB + r];  // C_i[r][m] = tile(i,k)[r][m] or tile(k,i)[m][r]
      dv[u] = Dk[t];
    }
#pragma unroll
    for (int u = 0; u < DT / DTPB; ++u) {
      const int t = tid + u * DTPB, r = t / DB, m = t % DB;
      Ct[m * LDP + r] = cv[u];
      Cb[(int64_t)i * DT + t] = cv[u];
      Ds[r * LDP + m] = dv[u];
    }
  }
  __syncthreads();
  const int ty = tid / 16, tx = tid % 16;
  double acc[4][4] = {};
  for (int m = 0; m < DB; ++m) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = Ct[m * LDP + 4 * ty + u];
      b[u] = Ds[m * LDP + 4 * tx + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
  }
  double *w = Wb + (int64_t)i * DT;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) w[(4 * ty + u) * DB + 4 * tx + v] = acc[u][v];
}

// update of every lower tile (i, j) for pivot block k (one workgroup per
// tile, blocks handed out row by row):
//   (k, k) <- -D^-1 ; (i, k), i > k <- W_i ; (k, j), j < k <- W_j^T ;
//   otherwise (i, j) -= W_i C_j^T
// the workgroup of tile (k+1, k+1) then inverts it into Dnext
__global__ __launch_bounds__(DTPB) void k_sweep_update(int k, int nt, double *__restrict__ T,
                                                       const double *__restrict__ Dk, const double *__restrict__ Wb,
                                                       const double *__restrict__ Cb, double *__restrict__ Dnext) {
  // workgroup b runs on XCD b mod 8: each XCD takes a contiguous range of
  // tiles (a band of tile rows), so the panel tiles W_i it reads stay in its
  // L2 (the C_j of one band, at most nt tiles, too)
  const int64_t nb = gridDim.x, xb = blockIdx.x % 8, q8 = nb / 8, r8 = nb % 8;
  int64_t t = (xb < r8 ? xb * (q8 + 1) : r8 * (q8 + 1) + (xb - r8) * q8) + blockIdx.x / 8;
  // the next pivot tile goes to workgroup 0 (dispatched first; that mapping
  // gives it tile 0, which takes the pivot tile's workgroup instead): its
  // serial inverse then runs beside the other tiles' updates
  if (Dnext) {
    const int64_t tp = (int64_t)(k + 1) * (k + 2) / 2 + (k + 1);
    if (blockIdx.x == 0)
      t = tp;
    else if (t == tp)
      t = 0;
  }
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((int64_t)(i + 1) * (i + 2) / 2 <= t) ++i;
  while ((int64_t)i * (i + 1) / 2 > t) --i;
  const int j = (int)(t - (int64_t)i * (i + 1) / 2);
  double *tile = T + t * DT;
  const int tid = threadIdx.x;
  if (i == k && j == k) {
    for (int q = tid; q < DT; q += DTPB) tile[q] = -Dk[q];
    return;
  }
  if (j == k) {  // i > k
    const double *w = Wb + (int64_t)i * DT;
    for (int q = tid; q < DT; q += DTPB) tile[q] = w[q];
    return;
  }
  if (i == k) {  // j < k: M_Kj = W_j^T
    const double *w = Wb + (int64_t)j * DT;
    for (int q = tid; q < DT; q += DTPB) tile[q] = w[(q % DB) * DB + q / DB];
    return;
  }
  // W_i and C_j row by row (pitch LDQ: the fragment reads below fall in
  // distinct banks); Wt is the next pivot tile's row-major copy afterwards
  __shared__ double Wt[DB * LDQ];
  __shared__ double Ct[DB * LDQ];
  __shared__ double colk[2 * DB], rowk[2 * DB];
  const double *w = Wb + (int64_t)i * DT;
  const double *cj = Cb + (int64_t)j * DT;
  {  // all 32 loads in flight before the LDS stores
    double wv[DT / DTPB], cv[DT / DTPB];
#pragma unroll
    for (int u = 0; u < DT / DTPB; ++u) {
      wv[u] = w[tid + u * DTPB];
      cv[u] = cj[tid + u * DTPB];
    }
#pragma unroll
    for (int u = 0; u < DT / DTPB; ++u) {
      const int q = tid + u * DTPB, r = q / DB, m = q % DB;
      Wt[r * LDQ + m] = wv[u];
      Ct[r * LDQ + m] = cv[u];
    }
  }
  __syncthreads();
  // W_i C_j^T on the FP64 matrix cores: wavefront w takes rows R0..R0+31 and
  // columns C0..C0+31 as 2x2 16x16 tiles, 16 steps of v_mfma_f64_16x16x4f64
  // over the 64 inner indices (lane l: A = W_i[row l % 16][k + l / 16],
  // B = C_j[col l % 16][k + l / 16]; result entry r of lane l is row
  // l / 16 + 4 r, column l % 16 of the 16x16 tile)
  const int lane = tid & 63, wv = tid >> 6;
  const int R0 = 32 * (wv >> 1), C0 = 32 * (wv & 1), lr = lane & 15, lk = lane >> 4;
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int kk = 0; kk < DB; kk += 4) {
    const int kc = kk + lk;
    const double a0 = Wt[(R0 + lr) * LDQ + kc], a1 = Wt[(R0 + 16 + lr) * LDQ + kc];
    const double b0 = Ct[(C0 + lr) * LDQ + kc], b1 = Ct[(C0 + 16 + lr) * LDQ + kc];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
  const bool next = (i == k + 1 && j == k + 1 && Dnext);
  __syncthreads();  // Wt is reused for the next pivot below
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = (R0 + 16 * a + lk + 4 * r) * DB + C0 + 16 * b + lr;
        const double nv = tile[q] - acc[a][b][r];
        tile[q] = nv;
        if (next) Wt[q] = nv;
      }
  if (!next) return;
  __syncthreads();
  tile_inverse(Wt, colk, rowk);
  __syncthreads();
  for (int q = tid; q < DT; q += DTPB) Dnext[q] = Wt[q];
}

// the inverse (-M after the sweep) as full row-major n x n (FP64, and FP32
// when Af is given)
__global__ __launch_bounds__(DTPB) void k_sweep_out(int64_t n, const double *__restrict__ T, double *__restrict__ A,
                                                    float *__restrict__ Af) {
  const int64_t q = (int64_t)blockIdx.x * DTPB + threadIdx.x;
  if (q >= n * n) return;
  const int64_t r = q / n, c = q % n;
  const int64_t hi = r >= c ? r : c, lo = r >= c ? c : r;
  const double v = -T[tile_off((int)(hi / DB), (int)(lo / DB)) + (hi % DB) * DB + lo % DB];
  A[q] = v;
  if (Af) Af[q] = (float)v;
}

// y = A x for a row-major n x n FP32 matrix, n a multiple of 4: one
// wavefront per row, 16-byte loads
__global__ __launch_bounds__(256) void k_dense_mv4(int64_t n, const float4 *__restrict__ A, const float4 *__restrict__ x,
                                                   float *__restrict__ y) {
  const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const int64_t n4 = n / 4;
  const float4 *a = A + row * n4;
  float acc = 0.f;
  for (int64_t c = lane; c < n4; c += 64) {
    const float4 av = a[c], xv = x[c];
    acc += av.x * xv.x + av.y * xv.y + av.z * xv.z + av.w * xv.w;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) y[row] = acc;
}

// Row-scaled FP16 copy of a row-major n x n FP32 matrix: s_i = max_j |A_ij|
// (1 for a zero row), H_ij = A_ij / s_i rounded to FP16, so every entry
// keeps 11 significant bits relative to its row's largest.  One wavefront
// per row.
__global__ __launch_bounds__(256) void k_dense_to_half_rows(int64_t n, const float *__restrict__ A,
                                                            __half *__restrict__ H, float *__restrict__ sc) {
  const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const float *a = A + row * n;
  float m = 0.f;
  for (int64_t c = lane; c < n; c += 64) m = fmaxf(m, fabsf(a[c]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  const float s = m > 0.f ? m : 1.f;
  const float inv = 1.f / s;
  for (int64_t c = lane; c < n; c += 64) H[row * n + c] = __float2half_rn(a[c] * inv);
  if (lane == 0) sc[row] = s;
}

// y = A x with A as the row-scaled FP16 copy (n a multiple of 8): one
// wavefront per row, eight entries per 16-byte load, FP32 sums in a fixed
// order, y_i = s_i sum_j H_ij x_j
__global__ __launch_bounds__(256) void k_dense_mvh(int64_t n, const uint4 *__restrict__ H, const float *__restrict__ sc,
                                                   const float4 *__restrict__ x, float *__restrict__ y) {
  const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const int64_t n8 = n / 8;
  const uint4 *h = H + row * n8;
  float acc = 0.f;
  for (int64_t c = lane; c < n8; c += 64) {
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    const v4u hv = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(h + c));
    const float4 x0 = x[2 * c], x1 = x[2 * c + 1];
    const unsigned w[4] = {hv.x, hv.y, hv.z, hv.w};
    float2 f[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) f[q] = __half22float2(*reinterpret_cast<const __half2 *>(&w[q]));
    acc += f[0].x * x0.x + f[0].y * x0.y + f[1].x * x0.z + f[1].y * x0.w + f[2].x * x1.x + f[2].y * x1.y +
           f[3].x * x1.z + f[3].y * x1.w;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) y[row] = sc[row] * acc;
}

}  // namespace

namespace uftk {

void dense_to_half_rows(hipStream_t s, int64_t n, const float *A, uint16_t *H, float *sc) {
  k_dense_to_half_rows<<<(unsigned)((n * 64 + 255) / 256), 256, 0, s>>>(n, A, (__half *)H, sc);
}
void dense_mvh(hipStream_t s, int64_t n, const uint16_t *H, const float *sc, const float *x, float *y) {
  k_dense_mvh<<<(unsigned)((n * 64 + 255) / 256), 256, 0, s>>>(n, (const uint4 *)H, sc, (const float4 *)x, y);
}

// scratch doubles needed by dense_spd_sweep for an n x n matrix
int64_t dense_sweep_scratch(int64_t n) {
  const int64_t nt = (n + DB - 1) / DB;
  return nt * (nt + 1) / 2 * DT + 2 * nt * DT + 2 * DT;
}

// A: the row-major dense matrix (n x n); S: dense_sweep_scratch(n) doubles;
// the inverse goes to Ainv (and its FP32 copy to Ainvf when given)
void dense_spd_sweep(hipStream_t s, int64_t n, const double *A, double *S, double *Ainv, float *Ainvf) {
  const int nt = (int)((n + DB - 1) / DB);
  const int64_t ntile = (int64_t)nt * (nt + 1) / 2;
  double *T = S, *Wb = T + ntile * DT, *Cb = Wb + (int64_t)nt * DT, *Dv = Cb + (int64_t)nt * DT;
  k_sweep_pack<<<(unsigned)((ntile * DT + DTPB - 1) / DTPB), DTPB, 0, s>>>(n, nt, A, T);
  k_sweep_first<<<1, DTPB, 0, s>>>(T, Dv);
  for (int k = 0; k < nt; ++k) {
    double *Dk = Dv + (k & 1) * DT, *Dn = k + 1 < nt ? Dv + ((k + 1) & 1) * DT : nullptr;
    k_sweep_panel<<<nt, DTPB, 0, s>>>(k, T, Dk, Wb, Cb);
    k_sweep_update<<<(unsigned)ntile, DTPB, 0, s>>>(k, nt, T, Dk, Wb, Cb, Dn);
  }
  k_sweep_out<<<(unsigned)((n * n + DTPB - 1) / DTPB), DTPB, 0, s>>>(n, T, Ainv, Ainvf);
}

// the same from the 6x6-block sparse matrix (nblk blocks; n = 6 x block rows)
void dense_spd_sweep_bsr(hipStream_t s, int64_t n, int64_t nblk, const int32_t *brow, const int32_t *ci,
                         const double *val, double *S, double *Ainv, float *Ainvf) {
  const int nt = (int)((n + DB - 1) / DB);
  const int64_t ntile = (int64_t)nt * (nt + 1) / 2;
  double *T = S, *Wb = T + ntile * DT, *Cb = Wb + (int64_t)nt * DT, *Dv = Cb + (int64_t)nt * DT;
  k_sweep_zero<<<(unsigned)((ntile * DT + DTPB - 1) / DTPB), DTPB, 0, s>>>(n, nt, T);
  k_sweep_scatter<<<(unsigned)((nblk * 36 + DTPB - 1) / DTPB), DTPB, 0, s>>>(nblk, brow, ci, val, T);
  k_sweep_first<<<1, DTPB, 0, s>>>(T, Dv);
  for (int k = 0; k < nt; ++k) {
    double *Dk = Dv + (k & 1) * DT, *Dn = k + 1 < nt ? Dv + ((k + 1) & 1) * DT : nullptr;
    k_sweep_panel<<<nt, DTPB, 0, s>>>(k, T, Dk, Wb, Cb);
    k_sweep_update<<<(unsigned)ntile, DTPB, 0, s>>>(k, nt, T, Dk, Wb, Cb, Dn);
  }
  k_sweep_out<<<(unsigned)((n * n + DTPB - 1) / DTPB), DTPB, 0, s>>>(n, T, Ainv, Ainvf);
}

void dense_mvf4(hipStream_t s, int64_t n, const float *A, const float *x, float *y) {
  k_dense_mv4<<<(unsigned)((n * 64 + 255) / 256), 256, 0, s>>>(n, (const float4 *)A, (const float4 *)x, y);
}

}  // namespace uftk
