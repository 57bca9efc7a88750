// gfx950 kernels for the SIMP design iteration: element operators, the
// matrix-free stiffness apply, sensitivities, the Helmholtz filter,
// deterministic reductions, vector updates and the MMA passes.
//
// Data layout (HBM):
//   elem   [npe][ne] int32           node ids, structure-of-arrays
//   KE     [np/2][ne] double2        packed upper triangle of the unscaled
//                                    element stiffness (np = 300 hex8, 78 tet4),
//                                    entries 2q and 2q+1 of the row-major packing
//                                    interleaved per element so one wavefront
//                                    load moves 64 x 16 B contiguous bytes
//   KF     [npf/2][ne] double2       Helmholtz element matrix, same packing
//   T      [npe][ne]                 filter transfer weights
//   Ye     [nd][ne]                  element-local results of the element phase
//   DOF vectors are node-interleaved (3n + c), as in the reference.
// The stiffness apply is two-phase and atomic-free: the element phase writes
// E(x_e) KE_e u_e into Ye, the node phase sums each node's incident entries in
// a fixed order.  Results are therefore bit-reproducible run to run.
#include <type_traits>

#include <hip/hip_fp16.h>

#include "uft_internal.h"
#include "uft_reduce.h"
#include "uft_elem.h"

namespace {

constexpr int TPB = 256;

using namespace uftd;

// ---------------------------------------------------------------- reductions
template <int K>
__device__ __forceinline__ void block_sum_store(double (&v)[K], double *part) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  }
  __shared__ double sm[K][TPB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sm[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    part[(size_t)blockIdx.x * K + k] = (sm[k][0] + sm[k][1]) + (sm[k][2] + sm[k][3]);
  }
}

__device__ __forceinline__ void block_max_store(double v, double *part) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  __shared__ double sm[TPB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

__global__ __launch_bounds__(TPB) void k_final_sum(const double *__restrict__ part, int nb, int K,
                                                   double *__restrict__ out) {
  __shared__ double sm[TPB];
  for (int k = 0; k < K; ++k) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += TPB) acc += part[(size_t)i * K + k];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sm[0];
    __syncthreads();
  }
}
// K in {1, 2, 3, 4, 8, 16}: shuffle-tree variant of the same fixed-order sum
template <int K>
__global__ __launch_bounds__(TPB) void k_final_sum_k(const double *__restrict__ part, int nb, double *__restrict__ out) {
  double res[K];
  uftr::block_final<K>(part, nb, res);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = res[k];
  }
}

__global__ __launch_bounds__(TPB) void k_final_max(const double *__restrict__ part, int nb, double *__restrict__ out) {
  __shared__ double sm[TPB];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += TPB) acc = fmax(acc, part[i]);
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// ---------------------------------------------------------------- element stiffness
// Hex8: SolveFem::Hex8Isoparametric (SolveFem.cpp:339-408).  One thread per
// (element, upper node pair a<=b); the 3x3 block of nodes (a,b) is
//   sum_gp w_gp detJ [lam g_a g_b^T + mu g_b g_a^T + mu (g_a.g_b) I]
// which is the isotropic B^T D B block with engineering shear strains.
__constant__ double c_gp[3] = {-0.775, 0.0, 0.775};
__constant__ double c_gw[3] = {0.556, 0.889, 0.556};

__global__ __launch_bounds__(TPB) void k_ke_hex8(int ne, const int32_t *__restrict__ elem,
                                                 const double *__restrict__ xyz, double lam, double mu,
                                                 double *__restrict__ KE, float *__restrict__ KEf,
                                                 double *__restrict__ detmin) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= (int64_t)ne * 36) return;
  const int e = (int)(t / 36);
  int pr = (int)(t - (int64_t)e * 36);
  int a = 0;
  while (pr >= 8 - a) {
    pr -= 8 - a;
    ++a;
  }
  const int b = a + pr;
  double X[8][3];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int id = elem[(size_t)n * ne + e];
    X[n][0] = xyz[3 * (size_t)id];
    X[n][1] = xyz[3 * (size_t)id + 1];
    X[n][2] = xyz[3 * (size_t)id + 2];
  }
  double K[3][3] = {};
  double dmin = 1e300;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        const double xi = c_gp[i], eta = c_gp[j], zeta = c_gp[k];
        double J[3][3] = {};
        double dNa[3], dNb[3];
#pragma unroll
        for (int n = 0; n < 8; ++n) {
          const double sx = (n & 1) ? 1.0 : -1.0, sy = (n & 2) ? 1.0 : -1.0, sz = (n & 4) ? 1.0 : -1.0;
          const double fx = 1.0 + sx * xi, fy = 1.0 + sy * eta, fz = 1.0 + sz * zeta;
          const double d0 = 0.125 * sx * fy * fz, d1 = 0.125 * sy * fx * fz, d2 = 0.125 * sz * fx * fy;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            J[0][c] += d0 * X[n][c];
            J[1][c] += d1 * X[n][c];
            J[2][c] += d2 * X[n][c];
          }
          if (n == a) { dNa[0] = d0; dNa[1] = d1; dNa[2] = d2; }
          if (n == b) { dNb[0] = d0; dNb[1] = d1; dNb[2] = d2; }
        }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        dmin = fmin(dmin, det);
        const double id = 1.0 / det;
        double Ji[3][3];
        Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
        Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
        Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
        Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
        Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
        Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
        double ga[3], gb[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          ga[r] = Ji[r][0] * dNa[0] + Ji[r][1] * dNa[1] + Ji[r][2] * dNa[2];
          gb[r] = Ji[r][0] * dNb[0] + Ji[r][1] * dNb[1] + Ji[r][2] * dNb[2];
        }
        const double w = c_gw[i] * c_gw[j] * c_gw[k] * det;
        const double gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            K[r][c] += w * (lam * ga[r] * gb[c] + mu * ga[c] * gb[r] + (r == c ? mu * gg : 0.0));
      }
  // scatter the block into the packed upper triangle (row (r,a) = r*8+a)
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int row = r * 8 + a, col = c * 8 + b;
      if (a == b && row > col) continue;
      const int p = pk_index(row, col, 24);
      const size_t off = ((size_t)(p >> 1) * ne + e) * 2 + (p & 1);
      KE[off] = K[r][c];
      if (KEf) KEf[off] = (float)K[r][c];
    }
  if (a == 0 && b == 0 && detmin) detmin[e] = dmin;
}

// Tet4: tet4/Tet4Isoparametric.cpp:1-41, KE = (1/6) B^T D B sqrt(det(J J^T)).
__global__ __launch_bounds__(TPB) void k_ke_tet4(int ne, const int32_t *__restrict__ elem,
                                                 const double *__restrict__ xyz, double lam, double mu,
                                                 double *__restrict__ KE, float *__restrict__ KEf,
                                                 double *__restrict__ detmin) {
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= ne) return;
  double X[4][3];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int id = elem[(size_t)n * ne + e];
    X[n][0] = xyz[3 * (size_t)id];
    X[n][1] = xyz[3 * (size_t)id + 1];
    X[n][2] = xyz[3 * (size_t)id + 2];
  }
  double J[3][3], JJ[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) J[r][c] = X[r + 1][c] - X[0][c];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) JJ[r][c] = J[r][0] * J[c][0] + J[r][1] * J[c][1] + J[r][2] * J[c][2];
  auto det3 = [](const double M[3][3]) {
    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  };
  const double det = det3(J);
  const double ad = sqrt(det3(JJ));
  const double id = 1.0 / det;
  double Ji[3][3];
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  const double dN[3][4] = {{-1, 1, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 0, 1}};
  double g[4][3];
#pragma unroll
  for (int n = 0; n < 4; ++n)
#pragma unroll
    for (int r = 0; r < 3; ++r) g[n][r] = Ji[r][0] * dN[0][n] + Ji[r][1] * dN[1][n] + Ji[r][2] * dN[2][n];
  const double w = ad / 6.0;
  for (int a = 0; a < 4; ++a)
    for (int b = a; b < 4; ++b) {
      const double gg = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          const int row = r * 4 + a, col = c * 4 + b;
          if (a == b && row > col) continue;
          const double v = w * (lam * g[a][r] * g[b][c] + mu * g[a][c] * g[b][r] + (r == c ? mu * gg : 0.0));
          const int p = pk_index(row, col, 12);
          const size_t off = ((size_t)(p >> 1) * ne + e) * 2 + (p & 1);
          KE[off] = v;
          if (KEf) KEf[off] = (float)v;
        }
    }
  if (detmin) detmin[e] = det;
}

// ---------------------------------------------------------------- Helmholtz element
// PDEFilter::KeHel (PDEFilter.cpp:185-240) and tet4/KeHel.cpp:1-47, plus the
// element volume (SolveFem::ComputeVolOfHex, SolveFem.cpp:423-467; tet4:
// |det J|/6).  One thread per element; partial sums of the volume.
__device__ __forceinline__ double tet6(const double *a, const double *b, const double *c, const double *o) {
  const double u0 = o[0] - a[0], u1 = o[1] - a[1], u2 = o[2] - a[2];
  const double v0 = o[0] - b[0], v1 = o[1] - b[1], v2 = o[2] - b[2];
  const double w0 = o[0] - c[0], w1 = o[1] - c[1], w2 = o[2] - c[2];
  return (u0 * (v1 * w2 - v2 * w1) + u1 * (v2 * w0 - v0 * w2) + u2 * (v0 * w1 - v1 * w0)) / 6.0;
}

// r2d (design batches): the squared filter length of each design
__global__ __launch_bounds__(TPB) void k_kf_hex8(int ne, const int32_t *__restrict__ elem,
                                                 const double *__restrict__ xyz, double r2,
                                                 double *__restrict__ KF, double *__restrict__ T,
                                                 double *__restrict__ vol, double *__restrict__ part, SegG sg,
                                                 const double *__restrict__ r2d) {
  int64_t e64;
  int dsg;
  const bool ok = uftr::seg_row(sg, ne, e64, dsg);
  const int e = (int)e64;
  if (r2d) r2 = r2d[dsg];
  double v[1] = {0.0};
  if (ok) {
    double X[8][3];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const int id = elem[(size_t)n * ne + e];
      X[n][0] = xyz[3 * (size_t)id];
      X[n][1] = xyz[3 * (size_t)id + 1];
      X[n][2] = xyz[3 * (size_t)id + 2];
    }
    double K[36] = {}, Te[8] = {};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const double xi = c_gp[i], eta = c_gp[j], zeta = c_gp[k];
          double N[8], dN[3][8], J[3][3] = {};
#pragma unroll
          for (int n = 0; n < 8; ++n) {
            const double sx = (n & 1) ? 1.0 : -1.0, sy = (n & 2) ? 1.0 : -1.0, sz = (n & 4) ? 1.0 : -1.0;
            const double fx = 1.0 + sx * xi, fy = 1.0 + sy * eta, fz = 1.0 + sz * zeta;
            N[n] = 0.125 * fx * fy * fz;
            dN[0][n] = 0.125 * sx * fy * fz;
            dN[1][n] = 0.125 * sy * fx * fz;
            dN[2][n] = 0.125 * sz * fx * fy;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              J[0][c] += dN[0][n] * X[n][c];
              J[1][c] += dN[1][n] * X[n][c];
              J[2][c] += dN[2][n] * X[n][c];
            }
          }
          const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                             J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
          const double id = 1.0 / det;
          double Ji[3][3];
          Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
          Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
          Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
          Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
          Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
          Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
          Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
          Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
          Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
          double G[3][8];
#pragma unroll
          for (int n = 0; n < 8; ++n)
#pragma unroll
            for (int r = 0; r < 3; ++r) G[r][n] = Ji[r][0] * dN[0][n] + Ji[r][1] * dN[1][n] + Ji[r][2] * dN[2][n];
          const double w = c_gw[i] * c_gw[j] * c_gw[k] * det;
          int p = 0;
#pragma unroll
          for (int a = 0; a < 8; ++a) {
#pragma unroll
            for (int b = a; b < 8; ++b, ++p)
              K[p] += w * (r2 * (G[0][a] * G[0][b] + G[1][a] * G[1][b] + G[2][a] * G[2][b]) + N[a] * N[b]);
            Te[a] += w * N[a];
          }
        }
#pragma unroll
    for (int p = 0; p < 36; ++p) KF[((size_t)(p >> 1) * ne + e) * 2 + (p & 1)] = K[p];
#pragma unroll
    for (int a = 0; a < 8; ++a) T[(size_t)a * ne + e] = Te[a];
    const double p1 = tet6(X[0], X[1], X[3], X[5]);
    const double p2 = tet6(X[0], X[3], X[2], X[6]);
    const double p3 = tet6(X[5], X[4], X[6], X[0]);
    const double p4 = tet6(X[5], X[6], X[7], X[3]);
    const double p5 = tet6(X[0], X[3], X[6], X[5]);
    const double V = fabs(p1) + fabs(p2) + fabs(p3) + fabs(p4) + fabs(p5);
    vol[e] = V;
    v[0] = V;
  }
  block_sum_store<1>(v, part);
}

__global__ __launch_bounds__(TPB) void k_kf_tet4(int ne, const int32_t *__restrict__ elem,
                                                 const double *__restrict__ xyz, double r2,
                                                 double *__restrict__ KF, double *__restrict__ T,
                                                 double *__restrict__ vol, double *__restrict__ part, SegG sg,
                                                 const double *__restrict__ r2d) {
  int64_t e64;
  int dsg;
  const bool ok = uftr::seg_row(sg, ne, e64, dsg);
  const int e = (int)e64;
  if (r2d) r2 = r2d[dsg];
  double v[1] = {0.0};
  if (ok) {
    double X[4][3];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int id = elem[(size_t)n * ne + e];
      X[n][0] = xyz[3 * (size_t)id];
      X[n][1] = xyz[3 * (size_t)id + 1];
      X[n][2] = xyz[3 * (size_t)id + 2];
    }
    double J[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) J[r][c] = X[r + 1][c] - X[0][c];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double id = 1.0 / det;
    double Ji[3][3];
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    const double dN[3][4] = {{-1, 1, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 0, 1}};
    double G[3][4];
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int r = 0; r < 3; ++r) G[r][n] = Ji[r][0] * dN[0][n] + Ji[r][1] * dN[1][n] + Ji[r][2] * dN[2][n];
    const double qa = (5.0 + 3.0 * sqrt(5.0)) / 20.0, qb = (5.0 - sqrt(5.0)) / 20.0;
    const double Q[4][3] = {{qa, qb, qb}, {qb, qa, qb}, {qb, qb, qa}, {qb, qb, qb}};
    double K[10] = {}, Te[4] = {};
    const double w = 1.0 / 24.0 * det;
    for (int g = 0; g < 4; ++g) {
      const double N[4] = {1.0 - Q[g][0] - Q[g][1] - Q[g][2], Q[g][0], Q[g][1], Q[g][2]};
      int p = 0;
      for (int a = 0; a < 4; ++a) {
        for (int b = a; b < 4; ++b, ++p)
          K[p] += w * (r2 * (G[0][a] * G[0][b] + G[1][a] * G[1][b] + G[2][a] * G[2][b]) + N[a] * N[b]);
        Te[a] += w * N[a];
      }
    }
#pragma unroll
    for (int p = 0; p < 10; ++p) KF[((size_t)(p >> 1) * ne + e) * 2 + (p & 1)] = K[p];
#pragma unroll
    for (int a = 0; a < 4; ++a) T[(size_t)a * ne + e] = Te[a];
    vol[e] = fabs(det) / 6.0;
    v[0] = vol[e];
  }
  block_sum_store<1>(v, part);
}

// ---------------------------------------------------------------- stiffness apply

// Element phase: Ye[k][e] = E(x_e) (KE_e u_e)_k with the fixed DOFs of u
// masked to zero (MatZeroRowsColumns zeroes the columns, SolveFem.cpp:324).
// Streaming load of one packed KE pair: optionally non-temporal (the element
// blocks are read once per apply and should not evict the gathered vectors).
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ double2 ld_pair(const double2 *p) {
  if constexpr (NT) {
    const v2d v = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(p));
    return make_double2(v.x, v.y);
  } else {
    return *p;
  }
}
template <bool NT>
__device__ __forceinline__ float2 ld_pair(const float2 *p) {
  if constexpr (NT) {
    const v2f v = __builtin_nontemporal_load(reinterpret_cast<const v2f *>(p));
    return make_float2(v.x, v.y);
  } else {
    return *p;
  }
}

// Element phase variants (picked at setup by uft_autotune): workgroup size BT,
// non-temporal KE streaming NT, and an occupancy floor of WPE waves per SIMD
// (0 = compiler's choice).  All variants perform the same arithmetic in the
// same order, so their results are bit-identical.
template <int NPE, typename T, typename T2, int BT, bool NT, int WPE>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void k_elem_apply_v(int ne, const int32_t *__restrict__ elem, const T2 *__restrict__ KE,
                    const double *__restrict__ xphys, SimpCoef sc, const uint8_t *__restrict__ fixm,
                    const T *__restrict__ u, T *__restrict__ Ye, const double *__restrict__ gate,
                    const double *__restrict__ dgate, const int32_t *__restrict__ edes) {
  constexpr int ND = 3 * NPE, NP = ND * (ND + 1) / 2, NQ = NP / 2;
  if (gate && gate[0] == 0.0) return;  // a device-controlled solve did not ask for this product
  const int e = blockIdx.x * BT + threadIdx.x;
  if (e >= ne) return;
  if (edes && dgate[(size_t)edes[e] * UFT_BSC] == 0.0) return;  // a batch design that did not ask
  T ue[ND], ye[ND];
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    const int n = elem[(size_t)a * ne + e];
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T v = u[3 * (size_t)n + c];
      ue[c * NPE + a] = ((m >> c) & 1u) ? T(0) : v;
    }
  }
#pragma unroll
  for (int k = 0; k < ND; ++k) ye[k] = T(0);
  static_for<0, NQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const T2 kk = ld_pair<NT>(&KE[(size_t)q * ne + e]);
    kacc<pk_row(2 * q, ND), pk_col(2 * q, ND)>(kk.x, ue, ye);
    kacc<pk_row(2 * q + 1, ND), pk_col(2 * q + 1, ND)>(kk.y, ue, ye);
  });
  const T E = (T)simpE(sc, xphys[e]);
#pragma unroll
  for (int k = 0; k < ND; ++k) Ye[(size_t)k * ne + e] = E * ye[k];
}

// FP16 element blocks for the multigrid smoother of the mixed form: the packed
// triangle of KE_e / s_e (s_e = max |KE_e|, so every entry lies in [-1, 1]
// and keeps FP16's full relative precision down to 6e-5 of the largest
// entry) in groups of four halves, [NG][ne] x 8 bytes; FP32 arithmetic with
// E(x_e) s_e applied at the end.  Entries past the triangle are zero.
template <int NPE>
__global__ __launch_bounds__(TPB) void k_ke_to_half(int ne, const double2 *__restrict__ KE, uint4 *__restrict__ KEh,
                                                    float *__restrict__ KEs) {
  constexpr int ND = 3 * NPE, NP = ND * (ND + 1) / 2, NQ = NP / 2, NG = (NP + 7) / 8;
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= ne) return;
  double mx = 0.0;
  for (int q = 0; q < NQ; ++q) {
    const double2 v = KE[(size_t)q * ne + e];
    mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
  }
  const double sc = mx > 0.0 ? mx : 1.0;
  KEs[e] = (float)sc;
  for (int g = 0; g < NG; ++g) {
    unsigned w[4];
    for (int j = 0; j < 4; ++j) {
      unsigned short hb[2];
      for (int t = 0; t < 2; ++t) {
        const int p = 8 * g + 2 * j + t;
        double v = 0.0;
        if (p < NP) {
          const double2 pr = KE[(size_t)(p >> 1) * ne + e];
          v = (p & 1) ? pr.y : pr.x;
        }
        hb[t] = __half_as_ushort(__float2half_rn((float)(v / sc)));
      }
      w[j] = (unsigned)hb[0] | ((unsigned)hb[1] << 16);
    }
    KEh[(size_t)g * ne + e] = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

__device__ __forceinline__ float h_lo(unsigned w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu))); }
__device__ __forceinline__ float h_hi(unsigned w) { return __half2float(__ushort_as_half((unsigned short)(w >> 16))); }

// FP16 element phase: one 16-byte load moves eight packed entries, so a
// wavefront load instruction reads 1 KB contiguous (as the FP64 kernel's
// double2 loads do).
template <int NPE, int BT, bool NT, int WPE>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void k_elem_apply_h(int ne, const int32_t *__restrict__ elem, const uint4 *__restrict__ KEh,
                    const float *__restrict__ KEs, const double *__restrict__ xphys, SimpCoef sc,
                    const uint8_t *__restrict__ fixm, const float *__restrict__ u, float *__restrict__ Ye) {
  constexpr int ND = 3 * NPE, NP = ND * (ND + 1) / 2, NG = (NP + 7) / 8;
  const int e = blockIdx.x * BT + threadIdx.x;
  if (e >= ne) return;
  float ue[ND], ye[ND];
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    const int n = elem[(size_t)a * ne + e];
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float v = u[3 * (size_t)n + c];
      ue[c * NPE + a] = ((m >> c) & 1u) ? 0.f : v;
    }
  }
#pragma unroll
  for (int k = 0; k < ND; ++k) ye[k] = 0.f;
  static_for<0, NG>([&](auto gc) {
    constexpr int g = decltype(gc)::value;
    unsigned w[4];
    if constexpr (NT) {
      typedef unsigned v4u __attribute__((ext_vector_type(4)));
      const v4u v = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(&KEh[(size_t)g * ne + e]));
      w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    } else {
      const uint4 v = KEh[(size_t)g * ne + e];
      w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    }
    static_for<0, 8>([&](auto tc) {
      constexpr int t = decltype(tc)::value;
      constexpr int p = 8 * g + t;
      if constexpr (p < NP) kacc<pk_row(p, ND), pk_col(p, ND)>((t & 1) ? h_hi(w[t >> 1]) : h_lo(w[t >> 1]), ue, ye);
    });
  });
  const float E = (float)simpE(sc, xphys[e]) * KEs[e];
#pragma unroll
  for (int k = 0; k < ND; ++k) Ye[(size_t)k * ne + e] = E * ye[k];
}

template <int NPE, typename T, typename T2>
__global__ __launch_bounds__(TPB) void k_elem_apply(int ne, const int32_t *__restrict__ elem,
                                                    const T2 *__restrict__ KE, const double *__restrict__ xphys,
                                                    SimpCoef sc, const uint8_t *__restrict__ fixm,
                                                    const T *__restrict__ u, T *__restrict__ Ye,
                                                    const double *__restrict__ gate,
                                                    const double *__restrict__ dgate,
                                                    const int32_t *__restrict__ edes) {
  constexpr int ND = 3 * NPE, NP = ND * (ND + 1) / 2, NQ = NP / 2;
  if (gate && gate[0] == 0.0) return;  // a device-controlled solve did not ask for this product
  for (int e = blockIdx.x * TPB + threadIdx.x; e < ne; e += gridDim.x * TPB) {
    if (edes && dgate[(size_t)edes[e] * UFT_BSC] == 0.0) continue;  // a batch design that did not ask
    T ue[ND], ye[ND];
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
      const int n = elem[(size_t)a * ne + e];
      const unsigned m = fixm[n];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T v = u[3 * (size_t)n + c];
        ue[c * NPE + a] = ((m >> c) & 1u) ? T(0) : v;
      }
    }
#pragma unroll
    for (int k = 0; k < ND; ++k) ye[k] = T(0);
    static_for<0, NQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const T2 kk = KE[(size_t)q * ne + e];
      kacc<pk_row(2 * q, ND), pk_col(2 * q, ND)>(kk.x, ue, ye);
      kacc<pk_row(2 * q + 1, ND), pk_col(2 * q + 1, ND)>(kk.y, ue, ye);
    });
    const T E = (T)simpE(sc, xphys[e]);
#pragma unroll
    for (int k = 0; k < ND; ++k) Ye[(size_t)k * ne + e] = E * ye[k];
  }
}

// Element phase of the diagonal: Ye[k][e] = E(x_e) KE_e[k][k].
template <int NPE>
__global__ __launch_bounds__(TPB) void k_elem_diag(int ne, const double *__restrict__ KE,
                                                   const double *__restrict__ xphys, SimpCoef sc,
                                                   double *__restrict__ Ye) {
  constexpr int ND = 3 * NPE;
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= ne) return;
  const double E = simpE(sc, xphys[e]);
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const int p = k * ND - (k * (k - 1)) / 2;
    Ye[(size_t)k * ne + e] = E * KE[((size_t)(p >> 1) * ne + e) * 2 + (p & 1)];
  }
}

// Node phase: y[3n+c] = sum over incident (e,l) of Ye[c*npe+l][e] in a fixed
// order; fixed DOFs take FIX (1: copy ufix, i.e. unit diagonal times u;
// 2: the value 1, for the diagonal).  Optionally accumulates y.w per block.
template <int NPE, int NC, int FIX, typename T>
__global__ __launch_bounds__(TPB) void k_node_gather(int nn, int ne, const int32_t *__restrict__ iptr,
                                                     const int32_t *__restrict__ inc, const T *__restrict__ Ye,
                                                     const uint8_t *__restrict__ fixm, const T *__restrict__ ufix,
                                                     T *__restrict__ y, const T *__restrict__ dotw,
                                                     double *__restrict__ part, double *__restrict__ out) {
  const int n = blockIdx.x * TPB + threadIdx.x;
  double ds[1] = {0.0};
  if (n < nn) {
    T acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = T(0);
    const int k0 = iptr[n], k1 = iptr[n + 1];
    for (int k = k0; k < k1; ++k) {
      const int v = inc[k];
      const int e = v / NPE, l = v - e * NPE;
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += Ye[(size_t)(c * NPE + l) * ne + e];
    }
    if constexpr (NC == 3) {
      const unsigned m = fixm ? fixm[n] : 0u;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T val = acc[c];
        if ((m >> c) & 1u) {
          if constexpr (FIX == 1) val = ufix[3 * (size_t)n + c];
          if constexpr (FIX == 2) val = T(1);
        }
        y[3 * (size_t)n + c] = val;
        if (dotw) ds[0] += (double)val * (double)dotw[3 * (size_t)n + c];
      }
    } else {
      y[n] = acc[0];
      if (dotw) ds[0] += (double)acc[0] * (double)dotw[n];
    }
  }
  if (part) uftr::block_partials<1>(ds, part);
}

// ---- device control of the single-design flexible CG of the mixed forms.
// ctl (in the context's scalar array): [0] stop (0 running, 1 converged,
// 2 diverged, 3 iteration limit), [1] a reliable update runs this iteration,
// [2] iterations done, [3] reliable updates done, [4] last FP64 residual
// norm, [5] absolute tolerance, [6] iteration limit.  Every launch of an
// iteration is queued whatever the outcome; the kernels that change the
// iterate return at once when ctl[0] is set, those of the reliable update
// when ctl[1] is not, so the host reads ctl once per round of queued
// iterations instead of once per iteration.
__global__ void k_ksp_ctl_init(double *__restrict__ ctl, double rtrue, double tol, int maxit) {
  ctl[0] = 0.0;
  ctl[1] = 0.0;
  ctl[2] = 0.0;
  ctl[3] = 0.0;
  ctl[4] = rtrue;
  ctl[5] = tol;
  ctl[6] = (double)maxit;
}
// after an iteration's update (rr = its recursive r.r): the reliable update
// when the recursive residual has met the tolerance or fallen 100-fold below
// the last FP64 residual; otherwise the iteration limit
__global__ void k_ksp_ctl_iter(const double *__restrict__ rr, double *__restrict__ ctl) {
  if (ctl[0] != 0.0) return;
  const double it = ctl[2] + 1.0;
  ctl[2] = it;
  const double rn = sqrt(rr[0]);
  if (!isfinite(rn)) {
    ctl[0] = 2.0;
  } else if (rn <= ctl[5] || rn < 1e-2 * ctl[4]) {
    ctl[1] = 1.0;
  } else if (it >= ctl[6]) {
    ctl[0] = 3.0;
  }
}
// reliable update, node phase (ctl[1] set): q = K U summed from the element
// results Ye in incidence order (fixed DOFs: U), r = F - q, rf = FP32 r, the
// next fine cycle's first Chebyshev step from a zero guess on rf (the
// arithmetic of k_cg_update_lo: cr = rf, cd = cx = it dinvf rf), partials of
// r.r
template <int NPE>
__global__ __launch_bounds__(TPB) void k_reliable_node(int nn, int ne, const double *__restrict__ ctl,
                                                       const int32_t *__restrict__ iptr,
                                                       const int32_t *__restrict__ inc, const double *__restrict__ Ye,
                                                       const uint8_t *__restrict__ fixm, const double *__restrict__ U,
                                                       const double *__restrict__ F, double *__restrict__ r,
                                                       float *__restrict__ rf, const float *__restrict__ dinvf,
                                                       double it, float *__restrict__ cr, float *__restrict__ cd,
                                                       float *__restrict__ cx, double *__restrict__ part) {
  if (ctl[1] == 0.0) return;
  double ds[1] = {0.0};
  for (int n = blockIdx.x * TPB + threadIdx.x; n < nn; n += gridDim.x * TPB) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = iptr[n]; k < iptr[n + 1]; ++k) {
      const int v = inc[k];
      const int e = v / NPE, l = v - e * NPE;
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[c] += Ye[(size_t)(c * NPE + l) * ne + e];
    }
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t i = 3 * (size_t)n + c;
      const double q = ((m >> c) & 1u) ? U[i] : acc[c];
      const double rr = F[i] - q;
      r[i] = rr;
      const float rfi = (float)rr;
      rf[i] = rfi;
      cr[i] = rfi;
      const float di = (float)it * dinvf[i] * rfi;
      cd[i] = di;
      cx[i] = di;
      ds[0] += rr * rr;
    }
  }
  uftr::block_partials<1>(ds, part);
}
// q = K U from the element results Ye in incidence order (fixed DOFs: U),
// only when *gate != 0 (design batches: the FP64 residual pass that some
// design asked for)
template <int NPE>
__global__ __launch_bounds__(TPB) void k_node_q_gated(int nn, int ne, const double *__restrict__ gate,
                                                      const double *__restrict__ dgate,
                                                      const int32_t *__restrict__ ndes,
                                                      const int32_t *__restrict__ iptr,
                                                      const int32_t *__restrict__ inc, const double *__restrict__ Ye,
                                                      const uint8_t *__restrict__ fixm, const double *__restrict__ U,
                                                      double *__restrict__ q) {
  if (*gate == 0.0) return;
  const int n = blockIdx.x * TPB + threadIdx.x;
  if (n >= nn) return;
  if (ndes && dgate[(size_t)ndes[n] * UFT_BSC] == 0.0) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int k = iptr[n]; k < iptr[n + 1]; ++k) {
    const int v = inc[k];
    const int e = v / NPE, l = v - e * NPE;
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] += Ye[(size_t)(c * NPE + l) * ne + e];
  }
  const unsigned m = fixm[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t i = 3 * (size_t)n + c;
    q[i] = ((m >> c) & 1u) ? U[i] : acc[c];
  }
}
// one block after k_reliable_node: r.r (into rr_out), the stopping test on
// the FP64 residual norm
__global__ __launch_bounds__(TPB) void k_ksp_ctl_reliable(const double *__restrict__ part, int nb,
                                                          double *__restrict__ rr_out, double *__restrict__ ctl) {
  if (ctl[1] == 0.0) return;
  double res[1];
  uftr::block_final<1>(part, nb, res);
  if (threadIdx.x != 0) return;
  rr_out[0] = res[0];
  const double rt = sqrt(res[0]);
  ctl[4] = rt;
  ctl[3] += 1.0;
  ctl[1] = 0.0;
  if (rt <= ctl[5])
    ctl[0] = 1.0;
  else if (!isfinite(rt))
    ctl[0] = 2.0;
  else if (ctl[2] >= ctl[6])
    ctl[0] = 3.0;
}

// Sensitivities (SolveFem.cpp:208-237): uKu = u_e^T KE u_e from the packed
// triangle, fx += E uKu, df = -p x^(p-1) (Emax-Emin) uKu, gx += x V, dg = V/TV.
template <int NPE>
__global__ __launch_bounds__(TPB) void k_sens(int ne, const int32_t *__restrict__ elem,
                                              const double2 *__restrict__ KE, const double *__restrict__ xphys,
                                              SimpCoef sc, double Emax_minus_Emin, const double *__restrict__ vol,
                                              double invTV, const double *__restrict__ u, double *__restrict__ dfdx,
                                              double *__restrict__ dgdx0, double *__restrict__ part, SegG sg,
                                              const double *__restrict__ invTVd) {
  constexpr int ND = 3 * NPE, NP = ND * (ND + 1) / 2, NQ = NP / 2;
  int64_t e64;
  int dsg;
  const bool ok = uftr::seg_row(sg, ne, e64, dsg);
  const int e = (int)e64;
  if (invTVd) invTV = invTVd[dsg];
  double v[2] = {0.0, 0.0};
  if (ok) {
    double ue[ND];
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
      const int n = elem[(size_t)a * ne + e];
#pragma unroll
      for (int c = 0; c < 3; ++c) ue[c * NPE + a] = u[3 * (size_t)n + c];
    }
    double dg = 0.0, off = 0.0;
    static_for<0, NQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const double2 kk = KE[(size_t)q * ne + e];
      constexpr int i0 = pk_row(2 * q, ND), j0 = pk_col(2 * q, ND);
      constexpr int i1 = pk_row(2 * q + 1, ND), j1 = pk_col(2 * q + 1, ND);
      if constexpr (i0 == j0) dg += kk.x * ue[i0] * ue[j0]; else off += kk.x * ue[i0] * ue[j0];
      if constexpr (i1 == j1) dg += kk.y * ue[i1] * ue[j1]; else off += kk.y * ue[i1] * ue[j1];
    });
    const double uKu = dg + 2.0 * off;
    const double x = xphys[e];
    v[0] = (sc.Emin + pow(x, sc.penal) * Emax_minus_Emin) * uKu;
    dfdx[e] = -1.0 * sc.penal * pow(x, sc.penal - 1) * Emax_minus_Emin * uKu;
    v[1] = x * vol[e];
    if (dgdx0) dgdx0[e] = vol[e] * invTV;
  }
  block_sum_store<2>(v, part);
}

// ---------------------------------------------------------------- Helmholtz filter
template <int NPE>
__global__ __launch_bounds__(TPB) void k_helm_elem(int ne, const int32_t *__restrict__ elem,
                                                   const double2 *__restrict__ KF, const double *__restrict__ x,
                                                   double *__restrict__ Ye) {
  constexpr int NP = NPE * (NPE + 1) / 2, NQ = NP / 2;
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= ne) return;
  double xe[NPE], ye[NPE];
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    xe[a] = x[elem[(size_t)a * ne + e]];
    ye[a] = 0.0;
  }
  static_for<0, NQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const double2 kk = KF[(size_t)q * ne + e];
    kacc<pk_row(2 * q, NPE), pk_col(2 * q, NPE)>(kk.x, xe, ye);
    kacc<pk_row(2 * q + 1, NPE), pk_col(2 * q + 1, NPE)>(kk.y, xe, ye);
  });
#pragma unroll
  for (int a = 0; a < NPE; ++a) Ye[(size_t)a * ne + e] = ye[a];
}

template <int NPE>
__global__ __launch_bounds__(TPB) void k_helm_diag_elem(int ne, const double *__restrict__ KF, double *__restrict__ Ye) {
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= ne) return;
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    const int p = a * NPE - (a * (a - 1)) / 2;
    Ye[(size_t)a * ne + e] = KF[((size_t)(p >> 1) * ne + e) * 2 + (p & 1)];
  }
}

// TF x (PDEFilter.cpp:106,147): element phase of TX_n = sum T_e[a] x_e
template <int NPE>
__global__ __launch_bounds__(TPB) void k_tf_elem(int ne, const double *__restrict__ T, const double *__restrict__ x,
                                                 double *__restrict__ Ye) {
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= ne) return;
  const double xe = x[e];
#pragma unroll
  for (int a = 0; a < NPE; ++a) Ye[(size_t)a * ne + e] = T[(size_t)a * ne + e] * xe;
}

// TFX^T RX (PDEFilter.cpp:97,149): out_e = sum_a (1/npe) RX[node_a], with the
// optional [Xmin, 1] clamp of Filter (PDEFilter.cpp:155-163).
template <int NPE>
__global__ __launch_bounds__(TPB) void k_tfx_t(int ne, const int32_t *__restrict__ elem, const double *__restrict__ RX,
                                               double *__restrict__ out, int clamp, double Xmin) {
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= ne) return;
  const double w = 1.0 / (double)NPE;
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < NPE; ++a) s += w * RX[elem[(size_t)a * ne + e]];
  if (clamp) {
    if (s < Xmin) s = Xmin;
    else if (s > 1.0) s = 1.0;
  }
  out[e] = s;
}

// ---------------------------------------------------------------- vector kernels
template <typename T>
__global__ __launch_bounds__(TPB) void k_fill(int64_t n, T v, T *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) x[i] = v;
}
__global__ __launch_bounds__(TPB) void k_copy(int64_t n, const double *__restrict__ a, double *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) b[i] = a[i];
}
__global__ __launch_bounds__(TPB) void k_d2f(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ __launch_bounds__(TPB) void k_f2d_add(int64_t n, const float *__restrict__ a, double *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) b[i] += (double)a[i];
}
__global__ __launch_bounds__(TPB) void k_scale(int64_t n, const double *__restrict__ a, const double *__restrict__ dinv,
                                               double *__restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) z[i] = a[i] * dinv[i];
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_dot(int64_t n, const T *__restrict__ a, const T *__restrict__ b,
                                             double *__restrict__ part, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) v[0] = (double)a[i] * (double)b[i];
  uftr::block_partials<1>(v, part);
}
__global__ __launch_bounds__(TPB) void k_dot2(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                                              const double *__restrict__ c, double *__restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    v[0] = a[i] * b[i];
    v[1] = c[i] * c[i];
  }
  block_sum_store<2>(v, part);
}
__global__ void k_scalar_div(double *sc, int num, int den, int dst) { sc[dst] = sc[num] / sc[den]; }

// u += alpha p ; r -= alpha q ; partial r.r
template <typename T>
__global__ __launch_bounds__(TPB) void k_cg_update(int64_t n, const double *__restrict__ sc, int ia,
                                                   T *__restrict__ u, const T *__restrict__ p, T *__restrict__ r,
                                                   const T *__restrict__ q, double *__restrict__ part,
                                                   double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const double al = sc[ia];
  double v[1] = {0.0};
  if (i < n) {
    u[i] += (T)al * p[i];
    const T rr = r[i] - (T)al * q[i];
    r[i] = rr;
    v[0] = (double)rr * (double)rr;
  }
  uftr::block_partials<1>(v, part);
}
// FP64 iterate and residual updated with a lower-precision operator product:
// u += a p ; r -= a q (a = sc[ia] / sc[ib]) ; rf = FP32 copy of r for the
// preconditioner ; partials of r.r
// with cr given, also the next fine cycle's first Chebyshev step from a zero
// guess on b = rf (the arithmetic of k_cheb_first): cr = rf, cd = cx =
// it dinvf rf
__global__ __launch_bounds__(TPB) void k_cg_update_lo(int64_t n, const double *__restrict__ sc, int ia, int ib,
                                                      double *__restrict__ u, const double *__restrict__ p,
                                                      double *__restrict__ r, const float *__restrict__ q,
                                                      float *__restrict__ rf, double *__restrict__ part,
                                                      const float *__restrict__ dinvf, double it,
                                                      float *__restrict__ cr, float *__restrict__ cd,
                                                      float *__restrict__ cx, const double *__restrict__ stop) {
  if (stop && stop[0] != 0.0) return;  // a device-controlled solve has stopped: the iterate stays
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const double al = sc[ia] / sc[ib];
  double v[1] = {0.0};
  if (i < n) {
    u[i] += al * p[i];
    const double rr = r[i] - al * (double)q[i];
    r[i] = rr;
    const float rfi = (float)rr;
    rf[i] = rfi;
    v[0] = rr * rr;
    if (cr) {
      cr[i] = rfi;
      const float di = (float)it * dinvf[i] * rfi;
      cd[i] = di;
      cx[i] = di;
    }
  }
  uftr::block_partials<1>(v, part);
}
// after the FP32 cycle: zo = z ; z = zf ; partials of r.z, r.zo (flexible CG)
__global__ __launch_bounds__(TPB) void k_z_from_f32(int64_t n, const float *__restrict__ zf, const double *__restrict__ r,
                                                    double *__restrict__ z, double *__restrict__ zo,
                                                    double *__restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    const double zold = z[i], znew = (double)zf[i], ri = r[i];
    zo[i] = zold;
    z[i] = znew;
    v[0] = ri * znew;
    v[1] = ri * zold;
  }
  uftr::block_partials<2>(v, part);
}
// the same without the copy: the new z goes to znew (the old zo buffer, the
// caller swaps the two), the old z is only read
__global__ __launch_bounds__(TPB) void k_z_from_f32_swap(int64_t n, const float *__restrict__ zf,
                                                         const double *__restrict__ r, const double *__restrict__ zold,
                                                         double *__restrict__ znew, double *__restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    const double zn = (double)zf[i], ri = r[i];
    znew[i] = zn;
    v[0] = ri * zn;
    v[1] = ri * zold[i];
  }
  uftr::block_partials<2>(v, part);
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_xpby(int64_t n, const T *__restrict__ z, const double *__restrict__ sc, int ib,
                                              T *__restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const double be = sc[ib];
  if (i < n) p[i] = z[i] + (T)be * p[i];
}
// p = z + beta p and its FP32 copy (the mixed outer operator's input)
__global__ __launch_bounds__(TPB) void k_xpby_lo(int64_t n, const double *__restrict__ z, const double *__restrict__ sc,
                                                 int ib, double *__restrict__ p, float *__restrict__ pf) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const double be = sc[ib];
  if (i < n) {
    const double v = z[i] + be * p[i];
    p[i] = v;
    pf[i] = (float)v;
  }
}
__global__ __launch_bounds__(TPB) void k_axpy(int64_t n, double a, const double *__restrict__ x, double *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}
__global__ __launch_bounds__(TPB) void k_sub(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                                             double *__restrict__ c) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) c[i] = a[i] - b[i];
}
__global__ __launch_bounds__(TPB) void k_set_dofs(int n, const int32_t *__restrict__ idx, const double *__restrict__ val,
                                                  double *__restrict__ x) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i < n) x[idx[i]] = val ? val[i] : 0.0;
}

// ---------------------------------------------------------------- MMA
__global__ __launch_bounds__(TPB) void k_movelimit(int64_t n, const double *__restrict__ x, double Xmin, double Xmax,
                                                   double movlim, double *__restrict__ xmin,
                                                   double *__restrict__ xmax) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const double xi = x[i];
  const double hi = xi + movlim, lo = xi - movlim;
  xmax[i] = Xmax < hi ? Xmax : hi;   // Min(Xmax, x+movlim)
  xmin[i] = Xmin > lo ? Xmin : lo;   // Max(Xmin, x-movlim)
}

struct PtrArr {
  const double *p[UFT_MAXM];
};

// GenSub (MMA.cpp:506-637): asymptotes, move bounds, approximation
// coefficients; partial sums of sum_i pij/(U-x) + qij/(x-L) per constraint.
template <int M>
__global__ __launch_bounds__(TPB) void k_gensub(int64_t n, int k, double asyminit, double asymdec, double asyminc,
                                                int robust, int conmod,
                                                const double *__restrict__ x, const double *__restrict__ xo1,
                                                const double *__restrict__ xo2, const double *__restrict__ xmin,
                                                const double *__restrict__ xmax, const double *__restrict__ dfdx,
                                                PtrArr dg, double *__restrict__ L, double *__restrict__ U,
                                                double *__restrict__ alpha, double *__restrict__ beta,
                                                double *__restrict__ p0, double *__restrict__ q0,
                                                double *__restrict__ pij, double *__restrict__ qij,
                                                double *__restrict__ part, SegG sg) {
  int64_t i;
  int dsg;
  const bool ok = uftr::seg_row(sg, n, i, dsg);
  double v[M];
#pragma unroll
  for (int j = 0; j < M; ++j) v[j] = 0.0;
  if (ok) {
    const double xi = x[i], xmi = xmin[i], xma = xmax[i];
    double Li, Ui;
    if (k < 3) {
      Li = xi - asyminit * (xma - xmi);
      Ui = xi + asyminit * (xma - xmi);
    } else {
      const double x1 = xo1[i], x2 = xo2[i];
      const double hv = (xi - x1) * (x1 - x2);
      const double gm = hv < 0.0 ? asymdec : (hv > 0.0 ? asyminc : 1.0);
      Li = xi - gm * (x1 - L[i]);
      Ui = xi + gm * (U[i] - x1);
      const double w = fmax(1.0e-5, xma - xmi);
      if (robust == 0) {
        Li = fmax(Li, xi - 10.0 * w);
        Li = fmin(Li, xi - 0.01 * w);
        Ui = fmax(Ui, xi + 0.01 * w);
        Ui = fmin(Ui, xi + 10.0 * w);
      } else {  // RobustAsymptotesType 1 (MMA.cpp:557-575)
        Li = fmax(Li, xi - 100.0 * w);
        Li = fmin(Li, xi - 1.0e-4 * w);
        Ui = fmax(Ui, xi + 1.0e-4 * w);
        Ui = fmin(Ui, xi + 100.0 * w);
        const double lo = xmi - 1.0e-5, hi = xma + 1.0e-5;
        if (xi < lo) {
          Li = xi - (hi - xi) / 0.9;
          Ui = xi + (hi - xi) / 0.9;
        }
        if (xi > hi) {
          Li = xi - (xi - lo) / 0.9;
          Ui = xi + (xi - lo) / 0.9;
        }
      }
    }
    L[i] = Li;
    U[i] = Ui;
    const double al = 0.9 * Li + 0.1 * xi, be = 0.9 * Ui + 0.1 * xi;
    alpha[i] = xmi > al ? xmi : al;
    beta[i] = xma < be ? xma : be;
    const double df = dfdx[i];
    const double dfp = fmax(0.0, df), dfm = fmax(0.0, -1.0 * df);
    const double ux = Ui - xi, xl = xi - Li, ul = Ui - Li;
    const double feps = 1.0e-6;
    p0[i] = ux * ux * (dfp + 0.001 * fabs(df) + 0.5 * feps / ul);
    q0[i] = xl * xl * (dfm + 0.001 * fabs(df) + 0.5 * feps / ul);
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const double g = dg.p[j][i];
      // constraintModification adds the objective's regularisation (MMA.cpp:588-597)
      const double reg = conmod ? 0.001 * fabs(g) + 0.5 * feps / ul : 0.0;
      const double pj = ux * ux * (fmax(0.0, g) + reg), qj = xl * xl * (fmax(0.0, -1.0 * g) + reg);
      pij[(size_t)j * n + i] = pj;
      qij[(size_t)j * n + i] = qj;
      v[j] = pj / ux + qj / xl;
    }
  }
  block_sum_store<M>(v, part);
}

// One dual pass at fixed lambda (MMA.cpp:678-868,890-934): x(lambda) written
// in place, then sum_i pij/(U-x)+qij/(x-L) (shared by DualGrad and
// DualResidual) and the Hessian sum_i PQ_j df2 PQ_k (upper triangle).
struct LamArr {
  double v[UFT_MAXM];
};
template <int M>
__global__ __launch_bounds__(TPB) void k_dual(int64_t n, LamArr lam, const double *__restrict__ L,
                                              const double *__restrict__ U, const double *__restrict__ alpha,
                                              const double *__restrict__ beta, const double *__restrict__ p0,
                                              const double *__restrict__ q0, const double *__restrict__ pij,
                                              const double *__restrict__ qij, double *__restrict__ x,
                                              double *__restrict__ part, SegG sg, const double *__restrict__ lamd,
                                              const double *__restrict__ gate) {
  if (gate && *gate == 0.0) return;
  constexpr int K = M + M * (M + 1) / 2;
  int64_t i;
  int dsg;
  const bool ok = uftr::seg_row(sg, n, i, dsg);
  if (lamd) {  // per design (design batches, M = 1)
#pragma unroll
    for (int j = 0; j < M; ++j) lam.v[j] = lamd[dsg * M + j];
  }
  double v[K];
#pragma unroll
  for (int j = 0; j < K; ++j) v[j] = 0.0;
  if (ok) {
    double pl = p0[i], ql = q0[i];
    double pj[M], qj[M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      pj[j] = pij[(size_t)j * n + i];
      qj[j] = qij[(size_t)j * n + i];
      pl += pj[j] * lam.v[j];
      ql += qj[j] * lam.v[j];
    }
    const double Li = L[i], Ui = U[i];
    const double sp = sqrt(pl), sq = sqrt(ql);
    const double xu = (sp * Li + sq * Ui) / (sp + sq);
    double xi = xu;
    if (xi < alpha[i]) xi = alpha[i];
    if (xi > beta[i]) xi = beta[i];
    x[i] = xi;
    const double ux = Ui - xi, xl = xi - Li;
    double df2 = -1.0 / (2.0 * pl / (ux * ux * ux) + 2.0 * ql / (xl * xl * xl));
    if (xu < alpha[i]) df2 = 0.0;
    if (xu > beta[i]) df2 = 0.0;
    double PQ[M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      v[j] = pj[j] / ux + qj[j] / xl;
      PQ[j] = pj[j] / (ux * ux) - qj[j] / (xl * xl);
    }
    int h = M;
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
      for (int kk = j; kk < M; ++kk, ++h) v[h] = PQ[j] * df2 * PQ[kk];
  }
  block_sum_store<K>(v, part);
}

__global__ __launch_bounds__(TPB) void k_design_change(int64_t n, const double *__restrict__ x, double *__restrict__ xold,
                                                       double *__restrict__ part, SegG sg) {
  int64_t i;
  int dsg;
  const bool ok = uftr::seg_row(sg, n, i, dsg);
  double v = 0.0;
  if (ok) {
    v = fabs(x[i] - xold[i]);
    xold[i] = x[i];
  }
  block_max_store(v, part);
}

__global__ __launch_bounds__(TPB) void k_heaviside(int64_t n, const double *__restrict__ xf, double *__restrict__ xp,
                                                   double bt, double x0, double Xmin) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  double v = (tanh(bt * x0) + tanh(bt * (xf[i] - x0))) / (tanh(bt * x0) + tanh(bt * (1 - x0)));
  if (v < Xmin) v = Xmin;
  xp[i] = v;
}
__global__ __launch_bounds__(TPB) void k_projsens(int64_t n, const double *__restrict__ xf, double bt, double x0,
                                                  double *__restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const double th = tanh(bt * (xf[i] - x0));
  v[i] *= bt * (1.0 - th * th) / (tanh(bt * x0) + tanh(bt * (1 - x0)));
}

}  // namespace


// ---- element-phase variant dispatch
template <int NPE, typename T, typename T2>
static void elem_phase_t(hipStream_t s, int variant, int ne, const int32_t *elem, const T2 *KE, const double *xphys,
                         SimpCoef sc, const uint8_t *fixm, const T *u, T *Ye, const double *gate = nullptr,
                         const double *dgate = nullptr, const int32_t *edes = nullptr) {
  auto g = [&](int bt) { return (unsigned)((ne + bt - 1) / bt); };
  switch (variant) {
    case 1: k_elem_apply_v<NPE, T, T2, 256, true, 0><<<g(256), 256, 0, s>>>(ne, elem, KE, xphys, sc, fixm, u, Ye, gate, dgate, edes); break;
    case 2: k_elem_apply_v<NPE, T, T2, 256, true, 4><<<g(256), 256, 0, s>>>(ne, elem, KE, xphys, sc, fixm, u, Ye, gate, dgate, edes); break;
    case 3: k_elem_apply_v<NPE, T, T2, 128, true, 0><<<g(128), 128, 0, s>>>(ne, elem, KE, xphys, sc, fixm, u, Ye, gate, dgate, edes); break;
    case 4: k_elem_apply_v<NPE, T, T2, 512, true, 0><<<g(512), 512, 0, s>>>(ne, elem, KE, xphys, sc, fixm, u, Ye, gate, dgate, edes); break;
    case 5: k_elem_apply_v<NPE, T, T2, 256, false, 4><<<g(256), 256, 0, s>>>(ne, elem, KE, xphys, sc, fixm, u, Ye, gate, dgate, edes); break;
    case 6: k_elem_apply_v<NPE, T, T2, 64, true, 0><<<g(64), 64, 0, s>>>(ne, elem, KE, xphys, sc, fixm, u, Ye, gate, dgate, edes); break;
    default: k_elem_apply<NPE, T, T2><<<g(TPB), TPB, 0, s>>>(ne, elem, KE, xphys, sc, fixm, u, Ye, gate, dgate, edes); break;
  }
}

// ==========================================================================
// launch wrappers
// ==========================================================================
namespace uftk {

int elem_variant_count() { return 7; }

void elem_phase(hipStream_t s, const uft_ctx *C, int variant, const double *xphys, const double *u, double *Ye) {
  SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
  if (C->npe == 8)
    elem_phase_t<8, double, double2>(s, variant, (int)C->ne, C->d_elem, (const double2 *)C->d_KE, xphys, sc, C->d_fixm,
                                     u, Ye);
  else
    elem_phase_t<4, double, double2>(s, variant, (int)C->ne, C->d_elem, (const double2 *)C->d_KE, xphys, sc, C->d_fixm,
                                     u, Ye);
}
void elem_phase_f(hipStream_t s, const uft_ctx *C, int variant, const double *xphys, const float *u, float *Ye) {
  SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
  if (C->npe == 8)
    elem_phase_t<8, float, float2>(s, variant, (int)C->ne, C->d_elem, (const float2 *)C->d_KEf, xphys, sc, C->d_fixm, u,
                                   Ye);
  else
    elem_phase_t<4, float, float2>(s, variant, (int)C->ne, C->d_elem, (const float2 *)C->d_KEf, xphys, sc, C->d_fixm, u,
                                   Ye);
}

template <int NPE>
static void elem_phase_h_t(hipStream_t s, int variant, int ne, const int32_t *elem, const uint4 *KEh, const float *KEs,
                           const double *xphys, SimpCoef sc, const uint8_t *fixm, const float *u, float *Ye) {
  auto g = [ne](int bt) { return (unsigned)((ne + bt - 1) / bt); };
  switch (variant) {
    case 1: k_elem_apply_h<NPE, 256, true, 0><<<g(256), 256, 0, s>>>(ne, elem, KEh, KEs, xphys, sc, fixm, u, Ye); break;
    case 2: k_elem_apply_h<NPE, 256, true, 4><<<g(256), 256, 0, s>>>(ne, elem, KEh, KEs, xphys, sc, fixm, u, Ye); break;
    case 3: k_elem_apply_h<NPE, 128, true, 0><<<g(128), 128, 0, s>>>(ne, elem, KEh, KEs, xphys, sc, fixm, u, Ye); break;
    case 4: k_elem_apply_h<NPE, 512, true, 0><<<g(512), 512, 0, s>>>(ne, elem, KEh, KEs, xphys, sc, fixm, u, Ye); break;
    case 5: k_elem_apply_h<NPE, 256, false, 4><<<g(256), 256, 0, s>>>(ne, elem, KEh, KEs, xphys, sc, fixm, u, Ye); break;
    case 6: k_elem_apply_h<NPE, 64, true, 0><<<g(64), 64, 0, s>>>(ne, elem, KEh, KEs, xphys, sc, fixm, u, Ye); break;
    default: k_elem_apply_h<NPE, 256, false, 0><<<g(256), 256, 0, s>>>(ne, elem, KEh, KEs, xphys, sc, fixm, u, Ye); break;
  }
}
void elem_phase_h(hipStream_t s, const uft_ctx *C, int variant, const double *xphys, const float *u, float *Ye) {
  const SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
  if (C->npe == 8)
    elem_phase_h_t<8>(s, variant, (int)C->ne, C->d_elem, C->d_KEh, C->d_KEs, xphys, sc, C->d_fixm, u, Ye);
  else
    elem_phase_h_t<4>(s, variant, (int)C->ne, C->d_elem, C->d_KEh, C->d_KEs, xphys, sc, C->d_fixm, u, Ye);
}
void ke_to_half(hipStream_t s, const uft_ctx *C) {
  if (C->npe == 8)
    k_ke_to_half<8><<<nblk(C->ne), TPB, 0, s>>>((int)C->ne, (const double2 *)C->d_KE, C->d_KEh, C->d_KEs);
  else
    k_ke_to_half<4><<<nblk(C->ne), TPB, 0, s>>>((int)C->ne, (const double2 *)C->d_KE, C->d_KEh, C->d_KEs);
}
int ke_half_groups(int npe) {  // 16-byte groups of eight halves per element
  const int nd = 3 * npe, np = nd * (nd + 1) / 2;
  return (np + 7) / 8;
}

int red_blocks(int64_t n) { return (int)nblk(n); }

void gen_ke(hipStream_t s, int npe, int64_t ne, const int32_t *elem, const double *xyz, double nu, double *KE,
            float *KEf, double *detmin) {
  const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = 1.0 / (2.0 * (1.0 + nu));
  if (npe == 8)
    k_ke_hex8<<<nblk(ne * 36), TPB, 0, s>>>((int)ne, elem, xyz, lam, mu, KE, KEf, detmin);
  else
    k_ke_tet4<<<nblk(ne), TPB, 0, s>>>((int)ne, elem, xyz, lam, mu, KE, KEf, detmin);
}

void gen_kf(hipStream_t s, int npe, int64_t ne, const int32_t *elem, const double *xyz, double rmin, double *KF,
            double *T, double *vol, double *part, SegG sg, const double *r2d) {
  const double r = 1.0 / (2.0 * sqrt(3.0)) * rmin;
  const unsigned g = uftr::seg_grid(sg, ne);
  if (npe == 8)
    k_kf_hex8<<<g, TPB, 0, s>>>((int)ne, elem, xyz, r * r, KF, T, vol, part, sg, r2d);
  else
    k_kf_tet4<<<g, TPB, 0, s>>>((int)ne, elem, xyz, r * r, KF, T, vol, part, sg, r2d);
}

static SimpCoef simp(const uft_ctx *C) { return SimpCoef{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal}; }

void apply_K(hipStream_t s, const uft_ctx *C, const double *xphys, const double *u, double *y, const double *dotw,
             double *dot_out, hipEvent_t *ev) {
  double *part = dotw ? C->d_part : nullptr;
  const int ne = (int)C->ne, nn = (int)C->nn;
  if (ev) (void)hipEventRecord(ev[0], s);
  elem_phase(s, C, C->elem_variant, xphys, u, C->d_Ye);
  if (ev) (void)hipEventRecord(ev[1], s);
  if (C->npe == 8)
    k_node_gather<8, 3, 1, double><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Ye, C->d_fixm, u, y,
                                                            dotw, part, dot_out);
  else
    k_node_gather<4, 3, 1, double><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Ye, C->d_fixm, u, y,
                                                            dotw, part, dot_out);
  if (ev) (void)hipEventRecord(ev[2], s);
  if (dotw) final_sum(s, part, (int)nblk(nn), 1, dot_out);
}

void apply_Kf(hipStream_t s, const uft_ctx *C, const double *xphys, const float *u, float *y, hipEvent_t *ev) {
  const int ne = (int)C->ne, nn = (int)C->nn;
  if (ev) (void)hipEventRecord(ev[0], s);
  elem_phase_f(s, C, C->elem_variant_f, xphys, u, C->d_Yef);
  if (ev) (void)hipEventRecord(ev[1], s);
  if (C->npe == 8)
    k_node_gather<8, 3, 1, float><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Yef, C->d_fixm, u, y,
                                                           nullptr, nullptr, nullptr);
  else
    k_node_gather<4, 3, 1, float><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Yef, C->d_fixm, u, y,
                                                           nullptr, nullptr, nullptr);
  if (ev) (void)hipEventRecord(ev[2], s);
}

void ksp_ctl_init(hipStream_t s, double *ctl, double rtrue, double tol, int maxit) {
  k_ksp_ctl_init<<<1, 1, 0, s>>>(ctl, rtrue, tol, maxit);
}
void ksp_ctl_iter(hipStream_t s, const double *rr, double *ctl) { k_ksp_ctl_iter<<<1, 1, 0, s>>>(rr, ctl); }
void reliable_update(hipStream_t s, const uft_ctx *C, const double *xphys, const double *U, double *r, float *rf,
                     const float *dinvf, double it, float *cr, float *cd, float *cx, double *rr_out, double *ctl) {
  const SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
  const int ne = (int)C->ne, nn = (int)C->nn;
  const double *gate = ctl + 1;
  // (an idle launch costs little: every workgroup returns at once)
  const unsigned gn = nblk(nn);
  if (C->npe == 8) {
    elem_phase_t<8, double, double2>(s, C->elem_variant, ne, C->d_elem, (const double2 *)C->d_KE, xphys, sc, C->d_fixm,
                                     U, C->d_Ye, gate);
    k_reliable_node<8><<<gn, TPB, 0, s>>>(nn, ne, ctl, C->d_iptr, C->d_inc, C->d_Ye, C->d_fixm, U, C->d_F, r, rf,
                                          dinvf, it, cr, cd, cx, C->d_part);
  } else {
    elem_phase_t<4, double, double2>(s, C->elem_variant, ne, C->d_elem, (const double2 *)C->d_KE, xphys, sc, C->d_fixm,
                                     U, C->d_Ye, gate);
    k_reliable_node<4><<<gn, TPB, 0, s>>>(nn, ne, ctl, C->d_iptr, C->d_inc, C->d_Ye, C->d_fixm, U, C->d_F, r, rf,
                                          dinvf, it, cr, cd, cx, C->d_part);
  }
  k_ksp_ctl_reliable<<<1, TPB, 0, s>>>(C->d_part, (int)gn, rr_out, ctl);
}

void apply_K_gated(hipStream_t s, const uft_ctx *C, const double *xphys, const double *U, double *q,
                   const double *gate, const double *dgate) {
  const SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
  const int ne = (int)C->ne, nn = (int)C->nn;
  const int32_t *edes = dgate ? C->bt.d_edes : nullptr, *ndes = dgate ? C->bt.d_ndes : nullptr;
  if (C->npe == 8) {
    elem_phase_t<8, double, double2>(s, C->elem_variant, ne, C->d_elem, (const double2 *)C->d_KE, xphys, sc, C->d_fixm,
                                     U, C->d_Ye, gate, dgate, edes);
    k_node_q_gated<8><<<nblk(nn), TPB, 0, s>>>(nn, ne, gate, dgate, ndes, C->d_iptr, C->d_inc, C->d_Ye, C->d_fixm, U,
                                               q);
  } else {
    elem_phase_t<4, double, double2>(s, C->elem_variant, ne, C->d_elem, (const double2 *)C->d_KE, xphys, sc, C->d_fixm,
                                     U, C->d_Ye, gate, dgate, edes);
    k_node_q_gated<4><<<nblk(nn), TPB, 0, s>>>(nn, ne, gate, dgate, ndes, C->d_iptr, C->d_inc, C->d_Ye, C->d_fixm, U,
                                               q);
  }
}

__global__ void k_inv(int64_t n, const double *__restrict__ d, double *__restrict__ dinv, float *__restrict__ dinvf) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const double v = 1.0 / d[i];
  dinv[i] = v;
  if (dinvf) dinvf[i] = (float)v;
}

void diag_K(hipStream_t s, const uft_ctx *C, const double *xphys, double *dinv, float *dinvf) {
  const int ne = (int)C->ne, nn = (int)C->nn;
  double *tmp = C->d_q;  // scratch ndof vector
  if (C->npe == 8) {
    k_elem_diag<8><<<nblk(ne), TPB, 0, s>>>(ne, C->d_KE, xphys, simp(C), C->d_Ye);
    k_node_gather<8, 3, 2, double><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Ye, C->d_fixm,
                                                            nullptr, tmp, nullptr, nullptr, nullptr);
  } else {
    k_elem_diag<4><<<nblk(ne), TPB, 0, s>>>(ne, C->d_KE, xphys, simp(C), C->d_Ye);
    k_node_gather<4, 3, 2, double><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Ye, C->d_fixm,
                                                            nullptr, tmp, nullptr, nullptr, nullptr);
  }
  k_inv<<<nblk(C->ndof), TPB, 0, s>>>(C->ndof, tmp, dinv, dinvf);
}

void sens(hipStream_t s, const uft_ctx *C, const double *xphys, const double *u, double *dfdx, double *dgdx0,
          double *part, SegG sg, const double *invTVd) {
  const int ne = (int)C->ne;
  const unsigned g = uftr::seg_grid(sg, ne);
  if (C->npe == 8)
    k_sens<8><<<g, TPB, 0, s>>>(ne, C->d_elem, (const double2 *)C->d_KE, xphys, simp(C), C->P.Emax - C->P.Emin,
                                C->vec[UFT_VOL], 1.0 / C->TV, u, dfdx, dgdx0, part, sg, invTVd);
  else
    k_sens<4><<<g, TPB, 0, s>>>(ne, C->d_elem, (const double2 *)C->d_KE, xphys, simp(C), C->P.Emax - C->P.Emin,
                                C->vec[UFT_VOL], 1.0 / C->TV, u, dfdx, dgdx0, part, sg, invTVd);
}

void apply_KF(hipStream_t s, const uft_ctx *C, const double *x, double *y, double *dot_out) {
  if (C->fell.d_hv) {  // assembled on the node pattern: one launch
    helm_ell_apply(s, C, x, y, dot_out);
    return;
  }
  const int ne = (int)C->ne, nn = (int)C->nn;
  const double *dotw = dot_out ? x : nullptr;  // with dot_out: also x.y in the node phase
  double *part = dot_out ? C->d_part : nullptr;
  if (C->npe == 8) {
    k_helm_elem<8><<<nblk(ne), TPB, 0, s>>>(ne, C->d_elem, (const double2 *)C->d_KF, x, C->d_Ye);
    k_node_gather<8, 1, 0, double><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Ye, nullptr, nullptr, y,
                                                            dotw, part, nullptr);
  } else {
    k_helm_elem<4><<<nblk(ne), TPB, 0, s>>>(ne, C->d_elem, (const double2 *)C->d_KF, x, C->d_Ye);
    k_node_gather<4, 1, 0, double><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Ye, nullptr, nullptr, y,
                                                            dotw, part, nullptr);
  }
  if (dot_out) final_sum(s, part, (int)nblk(nn), 1, dot_out);
}

void diag_KF(hipStream_t s, const uft_ctx *C, double *dinv) {
  const int ne = (int)C->ne, nn = (int)C->nn;
  if (C->npe == 8) {
    k_helm_diag_elem<8><<<nblk(ne), TPB, 0, s>>>(ne, C->d_KF, C->d_Ye);
    k_node_gather<8, 1, 0, double><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Ye, nullptr, nullptr,
                                                            C->d_fq, nullptr, nullptr, nullptr);
  } else {
    k_helm_diag_elem<4><<<nblk(ne), TPB, 0, s>>>(ne, C->d_KF, C->d_Ye);
    k_node_gather<4, 1, 0, double><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Ye, nullptr, nullptr,
                                                            C->d_fq, nullptr, nullptr, nullptr);
  }
  k_inv<<<nblk(nn), TPB, 0, s>>>(nn, C->d_fq, dinv, nullptr);
}

void tf_mult(hipStream_t s, const uft_ctx *C, const double *xe, double *TX) {
  const int ne = (int)C->ne, nn = (int)C->nn;
  if (C->npe == 8) {
    k_tf_elem<8><<<nblk(ne), TPB, 0, s>>>(ne, C->d_T, xe, C->d_Ye);
    k_node_gather<8, 1, 0, double><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Ye, nullptr, nullptr, TX,
                                                            nullptr, nullptr, nullptr);
  } else {
    k_tf_elem<4><<<nblk(ne), TPB, 0, s>>>(ne, C->d_T, xe, C->d_Ye);
    k_node_gather<4, 1, 0, double><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, C->d_Ye, nullptr, nullptr, TX,
                                                            nullptr, nullptr, nullptr);
  }
}

void tfx_t(hipStream_t s, const uft_ctx *C, const double *RX, double *xe, int clamp, double Xmin) {
  const int ne = (int)C->ne;
  if (C->npe == 8)
    k_tfx_t<8><<<nblk(ne), TPB, 0, s>>>(ne, C->d_elem, RX, xe, clamp, Xmin);
  else
    k_tfx_t<4><<<nblk(ne), TPB, 0, s>>>(ne, C->d_elem, RX, xe, clamp, Xmin);
}

void final_sum(hipStream_t s, const double *part, int nb, int K, double *out) {
  switch (K) {
    case 1: k_final_sum_k<1><<<1, TPB, 0, s>>>(part, nb, out); break;
    case 2: k_final_sum_k<2><<<1, TPB, 0, s>>>(part, nb, out); break;
    case 3: k_final_sum_k<3><<<1, TPB, 0, s>>>(part, nb, out); break;
    case 4: k_final_sum_k<4><<<1, TPB, 0, s>>>(part, nb, out); break;
    case 8: k_final_sum_k<8><<<1, TPB, 0, s>>>(part, nb, out); break;
    case 16: k_final_sum_k<16><<<1, TPB, 0, s>>>(part, nb, out); break;
    default: k_final_sum<<<1, TPB, 0, s>>>(part, nb, K, out); break;
  }
}
void final_max(hipStream_t s, const double *part, int nb, double *out) { k_final_max<<<1, TPB, 0, s>>>(part, nb, out); }
void dot(hipStream_t s, int64_t n, const double *a, const double *b, double *part, double *out) {
  k_dot<double><<<nblk(n), TPB, 0, s>>>(n, a, b, part, out);
  final_sum(s, part, (int)nblk(n), 1, out);
}
void dotf(hipStream_t s, int64_t n, const float *a, const float *b, double *part, double *out) {
  k_dot<float><<<nblk(n), TPB, 0, s>>>(n, a, b, part, out);
  final_sum(s, part, (int)nblk(n), 1, out);
}
void dot2(hipStream_t s, int64_t n, const double *a, const double *b, const double *c, double *part, double *out) {
  k_dot2<<<nblk(n), TPB, 0, s>>>(n, a, b, c, part);
  final_sum(s, part, (int)nblk(n), 2, out);
}
void scalar_div(hipStream_t s, double *sc, int num, int den, int dst) { k_scalar_div<<<1, 1, 0, s>>>(sc, num, den, dst); }
void fill(hipStream_t s, int64_t n, double v, double *x) {
  if (n > 0) k_fill<double><<<nblk(n), TPB, 0, s>>>(n, v, x);
}
void fillf(hipStream_t s, int64_t n, float v, float *x) {
  if (n > 0) k_fill<float><<<nblk(n), TPB, 0, s>>>(n, v, x);
}
void copy(hipStream_t s, int64_t n, const double *a, double *b) {
  if (n > 0) k_copy<<<nblk(n), TPB, 0, s>>>(n, a, b);
}
void d2f(hipStream_t s, int64_t n, const double *a, float *b) { k_d2f<<<nblk(n), TPB, 0, s>>>(n, a, b); }
void f2d_add(hipStream_t s, int64_t n, const float *a, double *b) { k_f2d_add<<<nblk(n), TPB, 0, s>>>(n, a, b); }
void scale(hipStream_t s, int64_t n, const double *a, const double *dinv, double *z) {
  k_scale<<<nblk(n), TPB, 0, s>>>(n, a, dinv, z);
}
void cg_update(hipStream_t s, int64_t n, const double *sc, int ia, double *u, const double *p, double *r,
               const double *q, double *part, double *out) {
  k_cg_update<double><<<nblk(n), TPB, 0, s>>>(n, sc, ia, u, p, r, q, part, out);
  final_sum(s, part, (int)nblk(n), 1, out);
}
void cg_update_lo(hipStream_t s, int64_t n, const double *sc, int ia, int ib, double *u, const double *p, double *r,
                  const float *q, float *rf, double *part, double *out) {
  k_cg_update_lo<<<nblk(n), TPB, 0, s>>>(n, sc, ia, ib, u, p, r, q, rf, part, nullptr, 0.0, nullptr, nullptr,
                                           nullptr, nullptr);
  final_sum(s, part, (int)nblk(n), 1, out);
}
void cg_update_lo_cheb(hipStream_t s, int64_t n, const double *sc, int ia, int ib, double *u, const double *p,
                       double *r, const float *q, float *rf, double *part, double *out, const float *dinvf, double it,
                       float *cr, float *cd, float *cx, const double *stop) {
  k_cg_update_lo<<<nblk(n), TPB, 0, s>>>(n, sc, ia, ib, u, p, r, q, rf, part, dinvf, it, cr, cd, cx, stop);
  final_sum(s, part, (int)nblk(n), 1, out);
}
void z_from_f32(hipStream_t s, int64_t n, const float *zf, const double *r, double *z, double *zo, double *part,
                double *out) {
  k_z_from_f32<<<nblk(n), TPB, 0, s>>>(n, zf, r, z, zo, part);
  final_sum(s, part, (int)nblk(n), 2, out);
}
void z_from_f32_swap_beta(hipStream_t s, int64_t n, const float *zf, const double *r, const double *zold,
                          double *znew, double *part, double *sc, int irz, int iprev, int ibeta) {
  k_z_from_f32_swap<<<nblk(n), TPB, 0, s>>>(n, zf, r, zold, znew, part);
  final_beta_flex(s, part, (int)nblk(n), sc, irz, iprev, ibeta);
}
void z_from_f32_beta(hipStream_t s, int64_t n, const float *zf, const double *r, double *z, double *zo, double *part,
                     double *sc, int irz, int iprev, int ibeta) {
  k_z_from_f32<<<nblk(n), TPB, 0, s>>>(n, zf, r, z, zo, part);
  final_beta_flex(s, part, (int)nblk(n), sc, irz, iprev, ibeta);
}
void cg_updatef(hipStream_t s, int64_t n, const double *sc, int ia, float *u, const float *p, float *r, const float *q,
                double *part, double *out) {
  k_cg_update<float><<<nblk(n), TPB, 0, s>>>(n, sc, ia, u, p, r, q, part, out);
  final_sum(s, part, (int)nblk(n), 1, out);
}
void xpby(hipStream_t s, int64_t n, const double *z, const double *sc, int ib, double *p) {
  k_xpby<double><<<nblk(n), TPB, 0, s>>>(n, z, sc, ib, p);
}
void xpby_lo(hipStream_t s, int64_t n, const double *z, const double *sc, int ib, double *p, float *pf) {
  k_xpby_lo<<<nblk(n), TPB, 0, s>>>(n, z, sc, ib, p, pf);
}
void xpbyf(hipStream_t s, int64_t n, const float *z, const double *sc, int ib, float *p) {
  k_xpby<float><<<nblk(n), TPB, 0, s>>>(n, z, sc, ib, p);
}
void axpy_inplace(hipStream_t s, int64_t n, double a, const double *x, double *y) {
  k_axpy<<<nblk(n), TPB, 0, s>>>(n, a, x, y);
}
void sub(hipStream_t s, int64_t n, const double *a, const double *b, double *c) { k_sub<<<nblk(n), TPB, 0, s>>>(n, a, b, c); }
void set_dofs(hipStream_t s, int n, const int32_t *idx, const double *val, double *x) {
  if (n > 0) k_set_dofs<<<nblk(n), TPB, 0, s>>>(n, idx, val, x);
}

void mma_movelimit(hipStream_t s, int64_t n, const double *x, double Xmin, double Xmax, double movlim, double *xmin,
                   double *xmax) {
  k_movelimit<<<nblk(n), TPB, 0, s>>>(n, x, Xmin, Xmax, movlim, xmin, xmax);
}

template <int M>
static void gensub_m(hipStream_t s, uft_ctx *C, int64_t n, int k, const double *x, const double *xmin,
                     const double *xmax, const double *dfdx, const double *const *dgdx, double *part, SegG sg) {
  PtrArr pa{};
  for (int j = 0; j < M; ++j) pa.p[j] = dgdx[j];
  k_gensub<M><<<uftr::seg_grid(sg, n), TPB, 0, s>>>(n, k, C->asyminit, C->asymdec, C->asyminc, C->mma_robust, C->mma_conmod, x,
                                      C->d_xo1, C->d_xo2, xmin, xmax,
                                      dfdx, pa, C->d_L, C->d_U, C->d_alpha, C->d_beta, C->d_p0, C->d_q0, C->d_pij,
                                      C->d_qij, part, sg);
}
void mma_gensub(hipStream_t s, uft_ctx *C, int64_t n, int m, int k, const double *x, const double *xmin,
                const double *xmax, const double *dfdx, const double *const *dgdx, double *part, SegG sg) {
  switch (m) {
    case 1: gensub_m<1>(s, C, n, k, x, xmin, xmax, dfdx, dgdx, part, sg); break;
    case 2: gensub_m<2>(s, C, n, k, x, xmin, xmax, dfdx, dgdx, part, sg); break;
    case 3: gensub_m<3>(s, C, n, k, x, xmin, xmax, dfdx, dgdx, part, sg); break;
    case 4: gensub_m<4>(s, C, n, k, x, xmin, xmax, dfdx, dgdx, part, sg); break;
    case 5: gensub_m<5>(s, C, n, k, x, xmin, xmax, dfdx, dgdx, part, sg); break;
    case 6: gensub_m<6>(s, C, n, k, x, xmin, xmax, dfdx, dgdx, part, sg); break;
    case 7: gensub_m<7>(s, C, n, k, x, xmin, xmax, dfdx, dgdx, part, sg); break;
    default: gensub_m<8>(s, C, n, k, x, xmin, xmax, dfdx, dgdx, part, sg); break;
  }
}

template <int M>
static void dual_m(hipStream_t s, uft_ctx *C, int64_t n, const double *lam, double *x, double *part, SegG sg,
                   const double *lamd, const double *gate) {
  LamArr la{};
  if (lam)
    for (int j = 0; j < M; ++j) la.v[j] = lam[j];
  k_dual<M><<<uftr::seg_grid(sg, n), TPB, 0, s>>>(n, la, C->d_L, C->d_U, C->d_alpha, C->d_beta, C->d_p0, C->d_q0,
                                                  C->d_pij, C->d_qij, x, part, sg, lamd, gate);
}
void mma_dual_pass(hipStream_t s, uft_ctx *C, int64_t n, int m, const double *lam, double *x, double *part, SegG sg,
                   const double *lamd, const double *gate) {
  switch (m) {
    case 1: dual_m<1>(s, C, n, lam, x, part, sg, lamd, gate); break;
    case 2: dual_m<2>(s, C, n, lam, x, part, sg, lamd, gate); break;
    case 3: dual_m<3>(s, C, n, lam, x, part, sg, lamd, gate); break;
    case 4: dual_m<4>(s, C, n, lam, x, part, sg, lamd, gate); break;
    case 5: dual_m<5>(s, C, n, lam, x, part, sg, lamd, gate); break;
    case 6: dual_m<6>(s, C, n, lam, x, part, sg, lamd, gate); break;
    case 7: dual_m<7>(s, C, n, lam, x, part, sg, lamd, gate); break;
    default: dual_m<8>(s, C, n, lam, x, part, sg, lamd, gate); break;
  }
}

void design_change(hipStream_t s, int64_t n, const double *x, double *xold, double *part, SegG sg) {
  k_design_change<<<uftr::seg_grid(sg, n), TPB, 0, s>>>(n, x, xold, part, sg);
}
void heaviside(hipStream_t s, int64_t n, const double *xf, double *xp, double beta, double x0, double Xmin) {
  k_heaviside<<<nblk(n), TPB, 0, s>>>(n, xf, xp, beta, x0, Xmin);
}
void projected_sens(hipStream_t s, int64_t n, const double *xf, double beta, double x0, double *v) {
  k_projsens<<<nblk(n), TPB, 0, s>>>(n, xf, beta, x0, v);
}

}  // namespace uftk

// ---------------------------------------------------------------- MMA KKT residual
// MMA::KKTresidual (MMA.cpp:412-480): per variable the stationarity residual
// with the bound multipliers it implies, and the two complementarity terms;
// partials (sum of squares, max abs) per block.
namespace {
template <int M>
__global__ __launch_bounds__(TPB) void k_kkt(int64_t n, LamArr lam, const double *__restrict__ x,
                                             const double *__restrict__ dfdx, PtrArr dg,
                                             const double *__restrict__ xmin, const double *__restrict__ xmax,
                                             double *__restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    double ri = dfdx[i];
#pragma unroll
    for (int j = 0; j < M; ++j) ri += lam.v[j] * dg.p[j][i];
    const double xi = x[i], lo = xmin[i], hi = xmax[i];
    const double mumin = (xi < lo + 1.0e-5 && ri > 0.0) ? ri : 0.0;
    const double mumax = (xi > hi - 1.0e-5 && ri < 0.0) ? -ri : 0.0;
    ri += -mumin + mumax;
    const double r1 = mumin * (xi - lo), r2 = mumax * (hi - xi);
    v[0] = ri * ri + r1 * r1 + r2 * r2;
    v[1] = fmax(fabs(ri), fmax(fabs(r1), fabs(r2)));
  }
  // fixed-order block combine: sum for the squares, max for the abs values
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v[0] += __shfl_down(v[0], off, 64);
    v[1] = fmax(v[1], __shfl_down(v[1], off, 64));
  }
  __shared__ double sm[2][TPB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sm[0][w] = v[0];
    sm[1][w] = v[1];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * (size_t)blockIdx.x] = (sm[0][0] + sm[0][1]) + (sm[0][2] + sm[0][3]);
    part[2 * (size_t)blockIdx.x + 1] = fmax(fmax(sm[1][0], sm[1][1]), fmax(sm[1][2], sm[1][3]));
  }
}
__global__ __launch_bounds__(TPB) void k_final_sum_max(const double *__restrict__ part, int nb,
                                                       double *__restrict__ out) {
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += TPB) {
    v[0] += part[2 * (size_t)i];
    v[1] = fmax(v[1], part[2 * (size_t)i + 1]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v[0] += __shfl_down(v[0], off, 64);
    v[1] = fmax(v[1], __shfl_down(v[1], off, 64));
  }
  __shared__ double sm[2][TPB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sm[0][w] = v[0];
    sm[1][w] = v[1];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = (sm[0][0] + sm[0][1]) + (sm[0][2] + sm[0][3]);
    out[1] = fmax(fmax(sm[1][0], sm[1][1]), fmax(sm[1][2], sm[1][3]));
  }
}

// ---------------------------------------------------------------- von Mises stress
// SolveFem::ComputeStress (stress/ComputeStress.cpp:53-104): strain at the
// element centroid (hex8: B(0,0,0); tet4: the constant strain), stress
// Emax D eps scaled by sqrt(x_e), von Mises of the six components.
template <int NPE>
__global__ __launch_bounds__(TPB) void k_stress(int ne, const int32_t *__restrict__ elem,
                                                const double *__restrict__ xyz, const double *__restrict__ xphys,
                                                const double *__restrict__ u, double Emax, double lam, double mu,
                                                double *__restrict__ vms) {
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= ne) return;
  double dNr[3][NPE];
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    if constexpr (NPE == 8) {
      dNr[0][a] = (a & 1) ? 0.125 : -0.125;
      dNr[1][a] = (a & 2) ? 0.125 : -0.125;
      dNr[2][a] = (a & 4) ? 0.125 : -0.125;
    } else {
#pragma unroll
      for (int r = 0; r < 3; ++r) dNr[r][a] = a == 0 ? -1.0 : (a == r + 1 ? 1.0 : 0.0);
    }
  }
  int nid[NPE];
  double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    nid[a] = elem[(size_t)a * ne + e];
    const double X0 = xyz[3 * (size_t)nid[a]], X1 = xyz[3 * (size_t)nid[a] + 1], X2 = xyz[3 * (size_t)nid[a] + 2];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      J[r * 3 + 0] += dNr[r][a] * X0;
      J[r * 3 + 1] += dNr[r][a] * X1;
      J[r * 3 + 2] += dNr[r][a] * X2;
    }
  }
  const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                     J[2] * (J[3] * J[7] - J[4] * J[6]);
  double Ji[9];
  Ji[0] = (J[4] * J[8] - J[5] * J[7]) / det;
  Ji[1] = (J[2] * J[7] - J[1] * J[8]) / det;
  Ji[2] = (J[1] * J[5] - J[2] * J[4]) / det;
  Ji[3] = (J[5] * J[6] - J[3] * J[8]) / det;
  Ji[4] = (J[0] * J[8] - J[2] * J[6]) / det;
  Ji[5] = (J[2] * J[3] - J[0] * J[5]) / det;
  Ji[6] = (J[3] * J[7] - J[4] * J[6]) / det;
  Ji[7] = (J[1] * J[6] - J[0] * J[7]) / det;
  Ji[8] = (J[0] * J[4] - J[1] * J[3]) / det;
  double eps[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    const double g0 = Ji[0] * dNr[0][a] + Ji[1] * dNr[1][a] + Ji[2] * dNr[2][a];
    const double g1 = Ji[3] * dNr[0][a] + Ji[4] * dNr[1][a] + Ji[5] * dNr[2][a];
    const double g2 = Ji[6] * dNr[0][a] + Ji[7] * dNr[1][a] + Ji[8] * dNr[2][a];
    const double ux = u[3 * (size_t)nid[a]], uy = u[3 * (size_t)nid[a] + 1], uz = u[3 * (size_t)nid[a] + 2];
    eps[0] += g0 * ux;
    eps[1] += g1 * uy;
    eps[2] += g2 * uz;
    eps[3] += g1 * ux + g0 * uy;
    eps[4] += g2 * uy + g1 * uz;
    eps[5] += g2 * ux + g0 * uz;
  }
  const double tr = eps[0] + eps[1] + eps[2], sc = Emax * sqrt(xphys[e]);
  double sg[6];
#pragma unroll
  for (int k = 0; k < 3; ++k) sg[k] = sc * (lam * tr + 2.0 * mu * eps[k]);
#pragma unroll
  for (int k = 3; k < 6; ++k) sg[k] = sc * mu * eps[k];
  const double d01 = sg[0] - sg[1], d02 = sg[0] - sg[2], d12 = sg[1] - sg[2];
  vms[e] = (1.0 / sqrt(2.0)) *
           sqrt(d01 * d01 + d02 * d02 + d12 * d12 + 6.0 * (sg[3] * sg[3] + sg[4] * sg[4] + sg[5] * sg[5]));
}
}  // namespace

namespace uftk {
template <int M>
static void kkt_m(hipStream_t s, int64_t n, const double *lam, const double *x, const double *dfdx,
                  const double *const *dgdx, const double *xmin, const double *xmax, double *part) {
  PtrArr pa{};
  LamArr la{};
  for (int j = 0; j < M; ++j) {
    pa.p[j] = dgdx[j];
    la.v[j] = lam[j];
  }
  k_kkt<M><<<nblk(n), TPB, 0, s>>>(n, la, x, dfdx, pa, xmin, xmax, part);
}
void mma_kkt(hipStream_t s, int64_t n, int m, const double *lam, const double *x, const double *dfdx,
             const double *const *dgdx, const double *xmin, const double *xmax, double *part, double *out) {
  switch (m) {
    case 1: kkt_m<1>(s, n, lam, x, dfdx, dgdx, xmin, xmax, part); break;
    case 2: kkt_m<2>(s, n, lam, x, dfdx, dgdx, xmin, xmax, part); break;
    case 3: kkt_m<3>(s, n, lam, x, dfdx, dgdx, xmin, xmax, part); break;
    case 4: kkt_m<4>(s, n, lam, x, dfdx, dgdx, xmin, xmax, part); break;
    case 5: kkt_m<5>(s, n, lam, x, dfdx, dgdx, xmin, xmax, part); break;
    case 6: kkt_m<6>(s, n, lam, x, dfdx, dgdx, xmin, xmax, part); break;
    case 7: kkt_m<7>(s, n, lam, x, dfdx, dgdx, xmin, xmax, part); break;
    default: kkt_m<8>(s, n, lam, x, dfdx, dgdx, xmin, xmax, part); break;
  }
  k_final_sum_max<<<1, TPB, 0, s>>>(part, (int)nblk(n), out);
}
void stress(hipStream_t s, const uft_ctx *C, const double *xphys, const double *u, double *vms) {
  const double nu = C->P.nu;
  const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 1.0 / (2.0 * (1.0 + nu));
  const int ne = (int)C->ne;
  if (C->npe == 8)
    k_stress<8><<<nblk(ne), TPB, 0, s>>>(ne, C->d_elem, C->d_xyz, xphys, u, C->P.Emax, lam, mu, vms);
  else
    k_stress<4><<<nblk(ne), TPB, 0, s>>>(ne, C->d_elem, C->d_xyz, xphys, u, C->P.Emax, lam, mu, vms);
}
}  // namespace uftk
