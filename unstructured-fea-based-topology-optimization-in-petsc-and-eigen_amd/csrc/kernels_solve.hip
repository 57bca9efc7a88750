// gfx950 kernels of the Krylov loops: fused vector updates of the Helmholtz
// filter's Jacobi-CG (PETSc KSPCG semantics, preconditioned residual norm)
// and the device-side scalar recurrences of CG, so a Krylov step needs no
// host round trip except its convergence test.
#include "uft_internal.h"
#include "uft_reduce.h"

namespace {
constexpr int TPB = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

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

// r = b - q ; z = D^-1 r ; partials ||D^-1 b||^2, ||z||^2, z.r
__global__ __launch_bounds__(TPB) void k_fcg_init(int64_t n, const double *__restrict__ b, const double *__restrict__ q,
                                                  const double *__restrict__ dinv, double *__restrict__ r,
                                                  double *__restrict__ z, double *__restrict__ part,
                                                  double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const double di = dinv[i], bi = b[i];
    const double ri = bi - q[i], zi = ri * di, zb = bi * di;
    r[i] = ri;
    z[i] = zi;
    v[0] = zb * zb;
    v[1] = zi * zi;
    v[2] = zi * ri;
  }
  uftr::block_partials<3>(v, part);
}

// x += alpha p ; r -= alpha q ; z = D^-1 r ; partials ||z||^2, z.r
__global__ __launch_bounds__(TPB) void k_fcg_update(int64_t n, const double *__restrict__ sc, int ia,
                                                    double *__restrict__ x, const double *__restrict__ p,
                                                    double *__restrict__ r, const double *__restrict__ q,
                                                    const double *__restrict__ dinv, double *__restrict__ z,
                                                    double *__restrict__ part, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const double al = sc[ia];
  double v[2] = {0.0, 0.0};
  if (i < n) {
    x[i] += al * p[i];
    const double ri = r[i] - al * q[i];
    r[i] = ri;
    const double zi = ri * dinv[i];
    z[i] = zi;
    v[0] = zi * zi;
    v[1] = zi * ri;
  }
  uftr::block_partials<2>(v, part);
}

// the same with alpha = sc[ia] / sc[ib] formed in place; nothing once *stop
// is set (the device-side convergence flag of the filter CG)
__global__ __launch_bounds__(TPB) void k_fcg_update_ab(int64_t n, const double *__restrict__ sc, int ia, int ib,
                                                       double *__restrict__ x, const double *__restrict__ p,
                                                       double *__restrict__ r, const double *__restrict__ q,
                                                       const double *__restrict__ dinv, double *__restrict__ z,
                                                       double *__restrict__ part, const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // the same for every block
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const double al = sc[ia] / sc[ib];
  double v[2] = {0.0, 0.0};
  if (i < n) {
    x[i] += al * p[i];
    const double ri = r[i] - al * q[i];
    r[i] = ri;
    const double zi = ri * dinv[i];
    z[i] = zi;
    v[0] = zi * zi;
    v[1] = zi * ri;
  }
  uftr::block_partials<2>(v, part);
}
// final sums of ||z||^2 and z.r into sc[izz], and the CG bookkeeping
// beta_old = beta (sc[ibo] = sc[ib]), beta = z.r (sc[ib])
__global__ __launch_bounds__(TPB) void k_final_fcg(const double *__restrict__ part, int nb, double *__restrict__ sc,
                                                   int izz, int ib, int ibo) {
  double res[2];
  uftr::block_final<2>(part, nb, res);
  if (threadIdx.x == 0) {
    sc[izz] = res[0];
    sc[ibo] = sc[ib];
    sc[ib] = res[1];
  }
}

// The filter CG's control on the device (slots ic .. ic + 4: state, tolerance,
// divergence bound, iteration count, last ||z||), so that the host queues
// several iterations per round trip.  State 0 runs, 1 converged, 2 diverged.
// Start, after the sums of the first residual (sc[i0] = ||D^-1 b||^2,
// sc[i0 + 1] = ||z||^2): the tests of PETSc's KSPCG with the preconditioned
// norm, rtol relative to ||D^-1 b|| (||z|| when b = 0), atol 1e-50, dtol 1e5.
__global__ void k_fcg_start(double *sc, int i0, int ic, double rtol) {
  const double sn0 = sqrt(sc[i0]), dp = sqrt(sc[i0 + 1]);
  const double snorm = sn0 == 0.0 ? dp : sn0;
  const double a = rtol * snorm;
  const double ttol = a < 1e-50 ? 1e-50 : a;
  sc[ic + 1] = ttol;
  sc[ic + 2] = 1e5 * snorm;
  sc[ic + 3] = 0.0;
  sc[ic + 4] = dp;
  sc[ic] = dp <= ttol ? 1.0 : 0.0;
}
// k_final_fcg with the iteration's tests: count the iteration, then stop on
// ||z|| <= tolerance or on divergence
__global__ __launch_bounds__(TPB) void k_final_fcg_ctl(const double *__restrict__ part, int nb, double *__restrict__ sc,
                                                       int izz, int ib, int ibo, int ic) {
  if (sc[ic] != 0.0) return;
  double res[2];
  uftr::block_final<2>(part, nb, res);
  if (threadIdx.x == 0) {
    sc[izz] = res[0];
    sc[ibo] = sc[ib];
    sc[ib] = res[1];
    sc[ic + 3] += 1.0;
    const double dp = sqrt(res[0]);
    sc[ic + 4] = dp;
    if (dp <= sc[ic + 1])
      sc[ic] = 1.0;
    else if (!(dp < sc[ic + 2]))
      sc[ic] = 2.0;
  }
}

// p = z + (sc[ib] / sc[ibo]) p, or p = z when first; nothing once *stop is set
__global__ __launch_bounds__(TPB) void k_fcg_p(int64_t n, const double *__restrict__ z, const double *__restrict__ sc,
                                               int ib, int ibo, int first, double *__restrict__ p,
                                               const double *__restrict__ stop) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n || (stop && *stop != 0.0)) return;
  if (first) {
    p[i] = z[i];
  } else {
    const double bb = sc[ib] / sc[ibo];
    p[i] = z[i] + bb * p[i];
  }
}

// Lanczos (Jacobi-CG) step without the iterate: alpha = sc[irz] / sc[ipq];
// r -= alpha q; z = D^-1 r; block partials of r.z.  q in FP64 or FP32.
template <typename TQ>
__global__ __launch_bounds__(TPB) void k_lanczos_update(int64_t n, const double *__restrict__ sc, int irz, int ipq,
                                                        double *__restrict__ r, const TQ *__restrict__ q,
                                                        const double *__restrict__ dinv, double *__restrict__ z,
                                                        double *__restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const double al = sc[irz] / sc[ipq];
  double v[1] = {0.0};
  if (i < n) {
    const double ri = r[i] - al * (double)q[i];
    r[i] = ri;
    const double zi = ri * dinv[i];
    z[i] = zi;
    v[0] = ri * zi;
  }
  uftr::block_partials<1>(v, part);
}
// the step's scalars: rz' = sum of the partials; log (alpha, beta = rz' / rz);
// sc[irz] = rz', sc[ibeta] = beta
__global__ __launch_bounds__(TPB) void k_lanczos_final(const double *__restrict__ part, int nb, double *__restrict__ sc,
                                                       int irz, int ipq, int ibeta, double *__restrict__ log2) {
  double res[1];
  uftr::block_final<1>(part, nb, res);
  if (threadIdx.x == 0) {
    const double al = sc[irz] / sc[ipq], be = res[0] / sc[irz];
    log2[0] = al;
    log2[1] = be;
    sc[ibeta] = be;
    sc[irz] = res[0];
  }
}

// CG scalars: sc[dst] = sc[a] / sc[b]
__global__ void k_div(double *sc, int a, int b, int dst) { sc[dst] = sc[a] / sc[b]; }
// beta step: sc[ibeta] = sc[inew] / sc[iold]; sc[iold] = sc[inew]; optional log
__global__ void k_beta(double *sc, int inew, int iold, int ibeta, double *log2) {
  const double b = sc[inew] / sc[iold];
  if (log2) {
    log2[0] = sc[iold];
    log2[1] = sc[inew];
  }
  sc[ibeta] = b;
  sc[iold] = sc[inew];
}
__global__ void k_copy_scalar(const double *src, double *dst) { *dst = *src; }
}  // namespace

namespace uftk {
void fcg_init(hipStream_t s, int64_t n, const double *b, const double *q, const double *dinv, double *r, double *z,
              double *part, double *out) {
  k_fcg_init<<<nblk(n), TPB, 0, s>>>(n, b, q, dinv, r, z, part, out);
  final_sum(s, part, (int)nblk(n), 3, out);
}
void fcg_update(hipStream_t s, int64_t n, const double *sc, int ia, double *x, const double *p, double *r,
                const double *q, const double *dinv, double *z, double *part, double *out) {
  k_fcg_update<<<nblk(n), TPB, 0, s>>>(n, sc, ia, x, p, r, q, dinv, z, part, out);
  final_sum(s, part, (int)nblk(n), 2, out);
}
void fcg_update_ab(hipStream_t s, int64_t n, double *sc, int ia, int ib, double *x, const double *p, double *r,
                   const double *q, const double *dinv, double *z, double *part, int izz, int ibeta, int ibo) {
  k_fcg_update_ab<<<nblk(n), TPB, 0, s>>>(n, sc, ia, ib, x, p, r, q, dinv, z, part, nullptr);
  k_final_fcg<<<1, TPB, 0, s>>>(part, (int)nblk(n), sc, izz, ibeta, ibo);
}
void fcg_start(hipStream_t s, double *sc, int i0, int ic, double rtol) { k_fcg_start<<<1, 1, 0, s>>>(sc, i0, ic, rtol); }
void fcg_update_ctl(hipStream_t s, int64_t n, double *sc, int ia, int ib, double *x, const double *p, double *r,
                    const double *q, const double *dinv, double *z, double *part, int izz, int ibeta, int ibo,
                    int ic) {
  k_fcg_update_ab<<<nblk(n), TPB, 0, s>>>(n, sc, ia, ib, x, p, r, q, dinv, z, part, sc + ic);
  k_final_fcg_ctl<<<1, TPB, 0, s>>>(part, (int)nblk(n), sc, izz, ibeta, ibo, ic);
}
void fcg_p(hipStream_t s, int64_t n, const double *z, const double *sc, int ib, int ibo, int first, double *p) {
  k_fcg_p<<<nblk(n), TPB, 0, s>>>(n, z, sc, ib, ibo, first, p, nullptr);
}
void fcg_p_ctl(hipStream_t s, int64_t n, const double *z, const double *sc, int ib, int ibo, int first,
               const double *stop, double *p) {
  k_fcg_p<<<nblk(n), TPB, 0, s>>>(n, z, sc, ib, ibo, first, p, stop);
}
void lanczos_step(hipStream_t s, int64_t n, double *sc, int irz, int ipq, int ibeta, double *r, const double *q,
                  const double *dinv, double *z, double *part, double *log2) {
  k_lanczos_update<double><<<nblk(n), TPB, 0, s>>>(n, sc, irz, ipq, r, q, dinv, z, part);
  k_lanczos_final<<<1, TPB, 0, s>>>(part, (int)nblk(n), sc, irz, ipq, ibeta, log2);
}
void lanczos_step(hipStream_t s, int64_t n, double *sc, int irz, int ipq, int ibeta, double *r, const float *q,
                  const double *dinv, double *z, double *part, double *log2) {
  k_lanczos_update<float><<<nblk(n), TPB, 0, s>>>(n, sc, irz, ipq, r, q, dinv, z, part);
  k_lanczos_final<<<1, TPB, 0, s>>>(part, (int)nblk(n), sc, irz, ipq, ibeta, log2);
}
void sc_div(hipStream_t s, double *sc, int a, int b, int dst) { k_div<<<1, 1, 0, s>>>(sc, a, b, dst); }
void sc_beta(hipStream_t s, double *sc, int inew, int iold, int ibeta, double *log2) {
  k_beta<<<1, 1, 0, s>>>(sc, inew, iold, ibeta, log2);
}
void sc_copy(hipStream_t s, const double *src, double *dst) { k_copy_scalar<<<1, 1, 0, s>>>(src, dst); }
}  // namespace uftk

namespace {
__global__ __launch_bounds__(TPB) void k_f2d(int64_t n, const float *__restrict__ a, double *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
// partials of r.z and r.zo (flexible CG, Polak-Ribiere form)
__global__ __launch_bounds__(TPB) void k_dot_pr(int64_t n, const double *__restrict__ r, const double *__restrict__ z,
                                                const double *__restrict__ zo, double *__restrict__ part,
                                                double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    v[0] = r[i] * z[i];
    v[1] = r[i] * zo[i];
  }
  uftr::block_partials<2>(v, part);
}
// beta = (rz - rzo) / rz_prev ; rz_prev = rz
__global__ void k_beta_flex(double *sc, int irz, int irzo, int iprev, int ibeta) {
  sc[ibeta] = (sc[irz] - sc[irzo]) / sc[iprev];
  sc[iprev] = sc[irz];
}
// the final sums of r.z and r.zo into sc[irz], sc[irz + 1] followed by
// k_beta_flex in the same launch
__global__ __launch_bounds__(TPB) void k_final_beta_flex(const double *__restrict__ part, int nb, double *sc, int irz,
                                                         int iprev, int ibeta) {
  double res[2];
  uftr::block_final<2>(part, nb, res);
  if (threadIdx.x == 0) {
    sc[irz] = res[0];
    sc[irz + 1] = res[1];
    sc[ibeta] = (res[0] - res[1]) / sc[iprev];
    sc[iprev] = res[0];
  }
}

// x_i uniform in [-1, 1) from a SplitMix64 hash of the index (a fixed,
// non-smooth start vector for the eigenvalue estimates)
__global__ __launch_bounds__(TPB) void k_fill_hash(int64_t n, double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  x[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
}  // namespace

namespace uftk {
void fill_hash(hipStream_t s, int64_t n, double *x) { k_fill_hash<<<nblk(n), TPB, 0, s>>>(n, x); }
void f2d(hipStream_t s, int64_t n, const float *a, double *b) { k_f2d<<<nblk(n), TPB, 0, s>>>(n, a, b); }
void dot_pr(hipStream_t s, int64_t n, const double *r, const double *z, const double *zo, double *part, double *out) {
  k_dot_pr<<<nblk(n), TPB, 0, s>>>(n, r, z, zo, part, out);
  final_sum(s, part, (int)nblk(n), 2, out);
}
void sc_beta_flex(hipStream_t s, double *sc, int irz, int irzo, int iprev, int ibeta) {
  k_beta_flex<<<1, 1, 0, s>>>(sc, irz, irzo, iprev, ibeta);
}
void final_beta_flex(hipStream_t s, const double *part, int nb, double *sc, int irz, int iprev, int ibeta) {
  k_final_beta_flex<<<1, TPB, 0, s>>>(part, nb, sc, irz, iprev, ibeta);
}
void dot_pr_beta(hipStream_t s, int64_t n, const double *r, const double *z, const double *zo, double *part,
                 double *sc, int irz, int iprev, int ibeta) {
  k_dot_pr<<<nblk(n), TPB, 0, s>>>(n, r, z, zo, part, sc + irz);
  final_beta_flex(s, part, (int)nblk(n), sc, irz, iprev, ibeta);
}
}  // namespace uftk

// ---- K-cycle helpers: linear combinations with device-resident coefficients
namespace {
template <typename T>
__global__ __launch_bounds__(TPB) void k_lincomb(int64_t n, const double *__restrict__ sc, int ia,
                                                 const T *__restrict__ a, int ib, const T *__restrict__ b,
                                                 T *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const T ca = ia < 0 ? T(1) : (T)sc[ia];
  const T cb = ib < 0 ? T(1) : (T)sc[ib];
  y[i] = ca * a[i] + cb * b[i];
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_addto(int64_t n, const T *__restrict__ d, T *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) x[i] += d[i];
}
// sc: 0 rho1 = c.v, 1 alpha1 = c.b  ->  7 = -alpha1/rho1
__global__ void k_kc1(double *sc) { sc[7] = sc[0] != 0.0 ? -sc[1] / sc[0] : 0.0; }
// 2 gamma = d.v, 3 beta = d.w, 4 alpha2 = d.r  ->  5, 6 coefficients of c and d
__global__ void k_kc2(double *sc) {
  const double rho1 = sc[0], a1 = sc[1], g = sc[2], be = sc[3], a2 = sc[4];
  const double t1 = rho1 != 0.0 ? a1 / rho1 : 0.0;
  const double rho2 = rho1 != 0.0 ? be - g * g / rho1 : 0.0;
  if (rho2 > 0.0) {
    sc[5] = t1 - g * a2 / (rho1 * rho2);
    sc[6] = a2 / rho2;
  } else {
    sc[5] = t1;
    sc[6] = 0.0;
  }
}
}  // namespace

namespace uftk {
void lincomb(hipStream_t s, int64_t n, const double *sc, int ia, const double *a, int ib, const double *b, double *y) {
  k_lincomb<double><<<nblk(n), TPB, 0, s>>>(n, sc, ia, a, ib, b, y);
}
void lincombf(hipStream_t s, int64_t n, const double *sc, int ia, const float *a, int ib, const float *b, float *y) {
  k_lincomb<float><<<nblk(n), TPB, 0, s>>>(n, sc, ia, a, ib, b, y);
}
void axpy_inplace_t(hipStream_t s, int64_t n, const double *d, double *x) {
  k_addto<double><<<nblk(n), TPB, 0, s>>>(n, d, x);
}
void axpy_inplace_t(hipStream_t s, int64_t n, const float *d, float *x) {
  k_addto<float><<<nblk(n), TPB, 0, s>>>(n, d, x);
}
void kcycle_coef1(hipStream_t s, double *sc) { k_kc1<<<1, 1, 0, s>>>(sc); }
void kcycle_coef2(hipStream_t s, double *sc) { k_kc2<<<1, 1, 0, s>>>(sc); }
}  // namespace uftk
