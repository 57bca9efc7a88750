// gfx950 kernels of design batches: many independent designs held in one
// context, every launch covering all of them.  Rows of every vector space are
// grouped by design, and a segment grid (SegG) cuts each design's rows into
// blocks of at most 256 that never span two designs, so a block's design is
// uniform: block partial sums are per design, and a per-design final kernel
// sums its own blocks in a fixed order (deterministic, like the single-design
// reductions).  Each design keeps its own Krylov scalars at sc + d * UFT_BSC;
// one thread per design runs the scalar recurrences, including the stopping
// tests, so a design that has converged keeps zero step lengths while the
// others continue.
#include "uft_internal.h"
#include "uft_reduce.h"

namespace {

constexpr int TPB = 256;
using uftr::seg_row;

// ---------------------------------------------------------------- per-design final reductions
template <int K>
__global__ __launch_bounds__(TPB) void k_seg_final(const double *__restrict__ part, const int32_t *__restrict__ dfirst,
                                                   double *__restrict__ sc, int base, const double *__restrict__ gate) {
  if (gate && *gate == 0.0) return;
  const int d = blockIdx.x;
  const int b0 = dfirst[d], nb = dfirst[d + 1] - b0;
  double res[K];
  uftr::block_final<K>(part + (size_t)b0 * K, nb, res);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sc[(size_t)d * UFT_BSC + base + k] = res[k];
  }
}

__global__ __launch_bounds__(TPB) void k_seg_final_max(const double *__restrict__ part,
                                                       const int32_t *__restrict__ dfirst, double *__restrict__ sc,
                                                       int base) {
  const int d = blockIdx.x;
  double acc = 0.0;
  for (int i = dfirst[d] + threadIdx.x; i < dfirst[d + 1]; i += TPB) acc = fmax(acc, part[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc = fmax(acc, __shfl_down(acc, off, 64));
  __shared__ double sm[TPB / 64];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sc[(size_t)d * UFT_BSC + base] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

// ---------------------------------------------------------------- vector kernels
__global__ __launch_bounds__(TPB) void k_seg_dot(SegG g, int64_t n, const double *__restrict__ a,
                                                 const double *__restrict__ b, double *__restrict__ part) {
  int64_t i;
  int d;
  double v[1] = {0.0};
  if (seg_row(g, n, i, d)) v[0] = a[i] * b[i];
  uftr::block_partials<1>(v, part);
}

__global__ __launch_bounds__(TPB) void k_seg_dot_pr(SegG g, int64_t n, const double *__restrict__ r,
                                                    const double *__restrict__ z, const double *__restrict__ zo,
                                                    double *__restrict__ part) {
  int64_t i;
  int d;
  double v[2] = {0.0, 0.0};
  if (seg_row(g, n, i, d)) {
    v[0] = r[i] * z[i];
    v[1] = r[i] * zo[i];
  }
  uftr::block_partials<2>(v, part);
}

// x += a_d p ; r -= a_d q ; partials of r.r
__global__ __launch_bounds__(TPB) void k_seg_cg_update(SegG g, int64_t n, const double *__restrict__ sc, int ia,
                                                       double *__restrict__ x, const double *__restrict__ p,
                                                       double *__restrict__ r, const double *__restrict__ q,
                                                       double *__restrict__ part) {
  int64_t i;
  int d;
  double v[1] = {0.0};
  if (seg_row(g, n, i, d)) {
    const double al = sc[(size_t)d * UFT_BSC + ia];
    x[i] += al * p[i];
    const double ri = r[i] - al * q[i];
    r[i] = ri;
    v[0] = ri * ri;
  }
  uftr::block_partials<1>(v, part);
}

// mixed forms: the outer product q in FP32 -- partials of p.q, and
// x += a_d p ; r -= a_d q with partials of r.r
__global__ __launch_bounds__(TPB) void k_seg_dot_lo(SegG g, int64_t n, const double *__restrict__ a,
                                                    const float *__restrict__ b, double *__restrict__ part) {
  int64_t i;
  int d;
  double v[1] = {0.0};
  if (seg_row(g, n, i, d)) v[0] = a[i] * (double)b[i];
  uftr::block_partials<1>(v, part);
}
__global__ __launch_bounds__(TPB) void k_seg_cg_update_lo(SegG g, int64_t n, const double *__restrict__ sc, int ia,
                                                          double *__restrict__ x, const double *__restrict__ p,
                                                          double *__restrict__ r, const float *__restrict__ q,
                                                          double *__restrict__ part) {
  int64_t i;
  int d;
  double v[1] = {0.0};
  if (seg_row(g, n, i, d)) {
    const double al = sc[(size_t)d * UFT_BSC + ia];
    x[i] += al * p[i];
    const double ri = r[i] - al * (double)q[i];
    r[i] = ri;
    v[0] = ri * ri;
  }
  uftr::block_partials<1>(v, part);
}

// reliable update of the designs that asked for one (B_NEED): r = F - q
// with q = K U in FP64; partials of r.r for every design
__global__ __launch_bounds__(TPB) void k_seg_true_res(SegG g, int64_t n, const double *__restrict__ sc,
                                                      const double *__restrict__ F, const double *__restrict__ q,
                                                      double *__restrict__ r, double *__restrict__ part,
                                                      const double *__restrict__ gate) {
  if (gate && *gate == 0.0) return;
  int64_t i;
  int d;
  double v[1] = {0.0};
  if (seg_row(g, n, i, d)) {
    double ri = r[i];
    if (sc[(size_t)d * UFT_BSC + B_NEED] != 0.0) {
      ri = F[i] - q[i];
      r[i] = ri;
    }
    v[0] = ri * ri;
  }
  uftr::block_partials<1>(v, part);
}

// p = z + b_d p
// ---- K-cycle of design batches: per-design inner products and coefficients
// (the single-design K-cycle's formulas, k_final_kc, per design)
// K inner products a_k . b_k per design (FP64 accumulation of T products)
template <int K, typename T>
__global__ __launch_bounds__(TPB) void k_seg_dots(SegG g, int64_t n, const T *__restrict__ a0, const T *__restrict__ b0,
                                                  const T *__restrict__ a1, const T *__restrict__ b1,
                                                  const T *__restrict__ a2, const T *__restrict__ b2,
                                                  double *__restrict__ part) {
  int64_t i;
  int d;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  if (seg_row(g, n, i, d)) {
    v[0] = (double)a0[i] * (double)b0[i];
    if constexpr (K > 1) v[1] = (double)a1[i] * (double)b1[i];
    if constexpr (K > 2) v[2] = (double)a2[i] * (double)b2[i];
  }
  uftr::block_partials<K>(v, part);
}
// first inner step: sc[base + 7] = -alpha1 / rho1 (rho1 = sc[base], alpha1 = sc[base + 1])
__global__ void k_bsc_kc1(int nseg, double *__restrict__ sc, int base) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nseg) return;
  double *s = sc + (size_t)d * UFT_BSC + base;
  s[7] = s[0] != 0.0 ? -s[1] / s[0] : 0.0;
}
// second inner step: sc[base + 5], sc[base + 6] from gamma, beta, alpha2 (sc[base + 2..4])
__global__ void k_bsc_kc2(int nseg, double *__restrict__ sc, int base) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nseg) return;
  double *s = sc + (size_t)d * UFT_BSC + base;
  const double rho1 = s[0], a1 = s[1], g = s[2], be = s[3], a2 = s[4];
  const double t1 = rho1 != 0.0 ? a1 / rho1 : 0.0;
  const double rho2 = rho1 != 0.0 ? be - g * g / rho1 : 0.0;
  if (rho2 > 0.0) {
    s[5] = t1 - g * a2 / (rho1 * rho2);
    s[6] = a2 / rho2;
  } else {
    s[5] = t1;
    s[6] = 0.0;
  }
}
// second start: kr = b + c7_d v ; r = kr ; d = it D^-1 kr ; kd = d
template <typename T>
__global__ __launch_bounds__(TPB) void k_seg_lincomb_first(SegG g, int64_t n, const double *__restrict__ sc, int ic,
                                                           const T *__restrict__ b, const T *__restrict__ v,
                                                           T *__restrict__ kr, T *__restrict__ r,
                                                           const T *__restrict__ dinv, double it, T *__restrict__ dd,
                                                           T *__restrict__ kd) {
  int64_t i;
  int d;
  if (!seg_row(g, n, i, d)) return;
  const T c = (T)sc[(size_t)d * UFT_BSC + ic];
  const T ri = b[i] + c * v[i];
  kr[i] = ri;
  r[i] = ri;
  const T di = (T)it * dinv[i] * ri;
  dd[i] = di;
  kd[i] = di;
}
// prolongation of a K-cycle result into the rows of the level above (segment
// grid of those rows): x_u += P_u (c5_d xa + c6_d xb)_parent(u)
template <int BS, typename T>
__global__ __launch_bounds__(TPB) void k_seg_prolong_kc(SegG g, int64_t n, const int32_t *__restrict__ parent,
                                                        const T *__restrict__ P, const double *__restrict__ sc, int i5,
                                                        const T *__restrict__ xa, const T *__restrict__ xb,
                                                        T *__restrict__ x) {
  int64_t t;
  int d;
  if (!seg_row(g, n, t, d)) return;
  const int64_t u = t / BS;
  const int i = (int)(t - u * BS);
  const int64_t A = parent[u];
  if (A < 0) return;
  const double *s = sc + (size_t)d * UFT_BSC + i5;
  const T ca = (T)s[0], cb = (T)s[1];
  const T *Pu = P + (size_t)u * BS * 6 + i * 6;
  T acc = T(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) acc += Pu[j] * (ca * xa[A * 6 + j] + cb * xb[A * 6 + j]);
  x[t] += acc;
}

__global__ __launch_bounds__(TPB) void k_seg_xpby(SegG g, int64_t n, const double *__restrict__ z,
                                                  const double *__restrict__ sc, int ib, double *__restrict__ p) {
  int64_t i;
  int d;
  if (!seg_row(g, n, i, d)) return;
  p[i] = z[i] + sc[(size_t)d * UFT_BSC + ib] * p[i];
}

__global__ __launch_bounds__(TPB) void k_seg_fill(SegG g, int64_t n, const double *__restrict__ vals,
                                                  double *__restrict__ x) {
  int64_t i;
  int d;
  if (seg_row(g, n, i, d)) x[i] = vals[d];
}

// Chebyshev scaling folded into the inverse diagonal: D^-1 / lambda_max of
// the row's design (and its FP32 copy)
__global__ __launch_bounds__(TPB) void k_seg_scale_dinv(SegG g, int64_t n, const double *__restrict__ fac,
                                                        double *__restrict__ dinv, float *__restrict__ dinvf) {
  int64_t i;
  int d;
  if (!seg_row(g, n, i, d)) return;
  const double v = dinv[i] * fac[d];
  dinv[i] = v;
  if (dinvf) dinvf[i] = (float)v;
}

// Helmholtz filter CG (PETSc KSPCG semantics, preconditioned residual norm):
// r = b - q ; z = D^-1 r ; partials ||D^-1 b||^2, ||z||^2, z.r
__global__ __launch_bounds__(TPB) void k_seg_fcg_init(SegG g, int64_t n, const double *__restrict__ b,
                                                      const double *__restrict__ q, const double *__restrict__ dinv,
                                                      double *__restrict__ r, double *__restrict__ z,
                                                      double *__restrict__ part) {
  int64_t i;
  int d;
  double v[3] = {0.0, 0.0, 0.0};
  if (seg_row(g, n, i, d)) {
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

// x += a_d p ; r -= a_d q ; z = D^-1 r ; partials ||z||^2, z.r
__global__ __launch_bounds__(TPB) void k_seg_fcg_update(SegG g, int64_t n, const double *__restrict__ sc, int ia,
                                                        double *__restrict__ x, const double *__restrict__ p,
                                                        double *__restrict__ r, const double *__restrict__ q,
                                                        const double *__restrict__ dinv, double *__restrict__ z,
                                                        double *__restrict__ part) {
  int64_t i;
  int d;
  double v[2] = {0.0, 0.0};
  if (seg_row(g, n, i, d)) {
    const double al = sc[(size_t)d * UFT_BSC + ia];
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

// p = z (first step) or p = z + c_d p
__global__ __launch_bounds__(TPB) void k_seg_fcg_p(SegG g, int64_t n, const double *__restrict__ z,
                                                   const double *__restrict__ sc, int ic, int first,
                                                   double *__restrict__ p) {
  int64_t i;
  int d;
  if (!seg_row(g, n, i, d)) return;
  p[i] = first ? z[i] : z[i] + sc[(size_t)d * UFT_BSC + ic] * p[i];
}

// ---------------------------------------------------------------- per-design scalar recurrences
#define FOR_DESIGN(nseg)                 \
  const int d = threadIdx.x;             \
  if (d >= nseg) return;                 \
  double *S = sc + (size_t)d * UFT_BSC;

// after r = F and ||r||^2 in B_RR: ||F||, tolerance, active flag
__global__ void k_bsc_el_init(int nseg, double *sc, double rtol) {
  FOR_DESIGN(nseg)
  const double fn = sqrt(S[B_RR]);
  S[B_FN] = fn;
  S[B_TOL] = rtol * fn;
  S[B_ACT] = fn > 0.0 ? 1.0 : 0.0;
  S[B_IT] = 0.0;
  S[B_RT] = fn;
  S[B_NEED] = 0.0;
}
__global__ void k_bsc_el_alpha(int nseg, double *sc) {
  FOR_DESIGN(nseg)
  S[B_ALPHA] = S[B_ACT] != 0.0 ? S[B_RZ] / S[B_PQ] : 0.0;
}
// after the update: count the step; once the recursive residual meets
// ||r|| <= rtol ||F|| ask for the true residual F - K U (B_NEED), on which the
// stopping test itself is taken (k_bsc_el_true; a non-finite residual also
// stops the design, and the host reports it from the true residual)
// The per-design tests below also leave batch-wide flags (one workgroup of
// 64 threads, a design per thread): after the check whether some design asks
// for its FP64 residual (flags[BF_NEED], gating that pass, and the count of
// such passes in flags[BF_NRES]); after that pass whether some design is
// still active (flags[BF_ACT]), read by the host once per round of queued
// iterations.
__global__ void k_bsc_el_check(int nseg, double *sc, double *flags) {
  const int d = threadIdx.x;
  bool need = false;
  if (d < nseg) {
    double *S = sc + (size_t)d * UFT_BSC;
    S[B_NEED] = 0.0;
    if (S[B_ACT] != 0.0) {
      S[B_IT] += 1.0;
      if (!(sqrt(S[B_RR]) > S[B_TOL])) S[B_NEED] = 1.0;
    }
    need = S[B_NEED] != 0.0;
  }
  need = __syncthreads_or(need);
  if (d == 0) {
    flags[BF_NEED] = need ? 1.0 : 0.0;
    if (need) flags[BF_NRES] += 1.0;
  }
}
// lower-precision outer product: count the step; ask for the FP64 residual
// (B_NEED) when the recursive one has met the tolerance or fallen 100-fold
// below the last FP64 residual (B_RT); the stopping test itself is taken on
// the FP64 residual (k_bsc_el_true)
__global__ void k_bsc_el_check_lo(int nseg, double *sc, double *flags) {
  const int d = threadIdx.x;
  bool need = false;
  if (d < nseg) {
    double *S = sc + (size_t)d * UFT_BSC;
    S[B_NEED] = 0.0;
    if (S[B_ACT] != 0.0) {
      S[B_IT] += 1.0;
      const double rn = sqrt(S[B_RR]);
      if (!(rn > S[B_TOL]) || rn < 1e-2 * S[B_RT]) S[B_NEED] = 1.0;
    }
    need = S[B_NEED] != 0.0;
  }
  need = __syncthreads_or(need);
  if (d == 0) {
    flags[BF_NEED] = need ? 1.0 : 0.0;
    if (need) flags[BF_NRES] += 1.0;
  }
}
// after the FP64 residual of the designs that asked for it
__global__ void k_bsc_el_true(int nseg, double *sc, double *flags) {
  const int d = threadIdx.x;
  bool act = false;
  if (d < nseg) {
    double *S = sc + (size_t)d * UFT_BSC;
    if (S[B_NEED] != 0.0) {
      const double rt = sqrt(S[B_RR]);
      S[B_RT] = rt;
      if (S[B_ACT] != 0.0 && !(rt > S[B_TOL])) S[B_ACT] = 0.0;
    }
    act = S[B_ACT] != 0.0;
  }
  act = __syncthreads_or(act);
  if (d == 0) flags[BF_ACT] = act ? 1.0 : 0.0;
}

// flexible CG (Polak-Ribiere): beta = (r.z - r.zo) / rz_prev
__global__ void k_bsc_el_beta(int nseg, double *sc) {
  FOR_DESIGN(nseg)
  if (S[B_ACT] != 0.0) {
    S[B_BETA] = (S[B_RZN] - S[B_RZO]) / S[B_RZ];
    S[B_RZ] = S[B_RZN];
  } else {
    S[B_BETA] = 0.0;
  }
}

// filter CG start (PDEFilter.cpp:117-137): ttol = max(rtol ||D^-1 b||, 1e-50)
// (||z|| when b = 0); a design already below it takes no step
__global__ void k_bsc_f_init(int nseg, double *sc, double rtol) {
  FOR_DESIGN(nseg)
  double snorm = sqrt(S[B_FB]);
  const double dp = sqrt(S[B_FZZ]);
  if (snorm == 0.0) snorm = dp;
  S[B_FSN] = snorm;
  S[B_FTOL] = fmax(rtol * snorm, 1e-50);
  S[B_FACT] = dp > S[B_FTOL] ? 1.0 : 0.0;
  S[B_FIT] = 0.0;
  S[B_FDIV] = 0.0;
  S[B_FPC] = 0.0;
}
__global__ void k_bsc_f_alpha(int nseg, double *sc) {
  FOR_DESIGN(nseg)
  S[B_FAL] = S[B_FACT] != 0.0 ? S[B_FBETA] / S[B_FPQ] : 0.0;
}
// after the update: beta_old = beta, beta = z.r, stopping and divergence
// tests (dtol 1e5), and the next direction's coefficient beta / beta_old
__global__ void k_bsc_f_after(int nseg, double *sc) {
  FOR_DESIGN(nseg)
  if (S[B_FACT] == 0.0) {
    S[B_FPC] = 0.0;
    return;
  }
  S[B_FIT] += 1.0;
  S[B_FBO] = S[B_FBETA];
  S[B_FBETA] = S[B_FZRN];
  const double dp = sqrt(S[B_FZZN]);
  if (dp <= S[B_FTOL]) {
    S[B_FACT] = 0.0;
  } else if (!(dp < 1e5 * S[B_FSN])) {
    S[B_FACT] = 0.0;
    S[B_FDIV] = 1.0;
  }
  S[B_FPC] = S[B_FACT] != 0.0 ? S[B_FBETA] / S[B_FBO] : 0.0;
}

// Lanczos from Jacobi-CG coefficients: alpha = rz / pq, logged per step
__global__ void k_bsc_l_alpha(int nseg, double *sc, double *log, int step) {
  FOR_DESIGN(nseg)
  const double a = S[B_LRZ] / S[B_LPQ];
  S[B_LAL] = a;
  log[((size_t)d * 64 + step) * 2] = a;
}
__global__ void k_bsc_l_beta(int nseg, double *sc, double *log, int step) {
  FOR_DESIGN(nseg)
  const double b = S[B_LRZN] / S[B_LRZ];
  S[B_LBE] = b;
  S[B_LRZ] = S[B_LRZN];
  log[((size_t)d * 64 + step) * 2 + 1] = b;
}

// ---- MMA dual interior point of a design batch on the device (m = 1): the
// single-design control flow of SolveDIP (MMA.cpp:639-676) per design, one
// thread each: epsilon levels from 1 by factors of 10 while above the
// tolerance, at most 100 Newton steps per level while err > 0.9 epsilon, err
// carried across levels; DualGrad / DualHess (MMA.cpp:730-868), the line
// search (MMA.cpp:870-888) and DualResidual (MMA.cpp:890-934) with m = 1.
// The fused dual pass between the kernels gives every design's constraint
// sum (B_RED) and Hessian sum (B_RED + 1) at its lambda (lamd).  No
// contraction of products into fused multiply-adds: the arithmetic is the
// host formulation's, operation by operation.
__device__ __forceinline__ void mma_yz(double *S, double a, double c) {  // XYZofLAMBDA (MMA.cpp:693-701)
#pragma clang fp contract(off)
  if (S[B_MLAM] < 0.0) S[B_MLAM] = 0.0;
  S[B_MY] = fmax(0.0, S[B_MLAM] - c);
  S[B_MZ] = fmax(0.0, 10.0 * (S[B_MLAM] * a - 1.0));
}
// gt[2 d] = gx of design d, gt[2 d + 1] = its tolerance 1e-9 sqrt(1 + n_d)
__global__ void k_bsc_mma_init(int nseg, double *sc, const double *gt, double a, double c, double *lamd) {
#pragma clang fp contract(off)
  FOR_DESIGN(nseg)
  S[B_MB] = S[B_RED] - gt[2 * d];
  S[B_MTOL] = gt[2 * d + 1];
  S[B_MLAM] = c / 2.0;
  S[B_MMU] = 1.0;
  S[B_MEPS] = 1.0;
  S[B_MERR] = 1.0;
  S[B_MLOOP] = 0.0;
  S[B_MSTEPS] = 0.0;
  S[B_MDONE] = 0.0;
  S[B_MSTEP] = 0.0;
  mma_yz(S, a, c);
  lamd[d] = S[B_MLAM];
}
// leave the epsilon levels whose inner loop has ended; decide who steps
// (flags[MF_STEP]), note a negative lambda among them (flags[MF_NEG]: the
// sums are then taken again at the clamped lambda before the step)
__global__ void k_bsc_mma_decide(int nseg, double *sc, double a, double c, double *lamd, double *flags) {
#pragma clang fp contract(off)
  const int d = threadIdx.x;
  bool step = false, neg = false;
  if (d < nseg) {
    double *S = sc + (size_t)d * UFT_BSC;
    S[B_MSTEP] = 0.0;
    if (S[B_MDONE] == 0.0) {
      while (!(S[B_MERR] > 0.9 * S[B_MEPS] && S[B_MLOOP] < 100.0)) {
        S[B_MEPS] *= 0.1;
        S[B_MLOOP] = 0.0;
        if (!(S[B_MEPS] > S[B_MTOL])) {
          S[B_MDONE] = 1.0;
          break;
        }
      }
      if (S[B_MDONE] == 0.0) {
        S[B_MSTEP] = 1.0;
        step = true;
        neg = S[B_MLAM] < 0.0;
        mma_yz(S, a, c);
        lamd[d] = S[B_MLAM];
      }
    }
  }
  step = __syncthreads_or(step);
  neg = __syncthreads_or(neg);
  if (d == 0) {
    flags[MF_STEP] = step ? 1.0 : 0.0;
    flags[MF_NEG] = neg ? 1.0 : 0.0;
  }
}
__global__ void k_bsc_mma_step(int nseg, double *sc, double a, double c, double *lamd) {
#pragma clang fp contract(off)
  FOR_DESIGN(nseg)
  if (S[B_MSTEP] == 0.0) return;
  S[B_MLOOP] += 1.0;
  S[B_MSTEPS] += 1.0;
  const double lam = S[B_MLAM], mu = S[B_MMU], ep = S[B_MEPS];
  double grad = S[B_RED] - S[B_MB] - a * S[B_MZ] - S[B_MY];
  grad = -1.0 * grad - ep / lam;
  double hs = S[B_RED + 1];
  if (lam > c) hs += -1.0;
  hs += -mu / lam;
  if (lam * a > 0.0) hs += -10.0 * a * a;
  double corr = 1e-4 * hs;
  if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
  hs += corr;
  const double s0 = grad / hs;
  const double s1 = -mu + ep / lam - s0 * mu / lam;
  double theta = 1.005;
  if (theta < -1.01 * s0 / lam) theta = -1.01 * s0 / lam;
  if (theta < -1.01 * s1 / mu) theta = -1.01 * s1 / mu;
  theta = 1.0 / theta;
  S[B_MLAM] += theta * s0;
  S[B_MMU] += theta * s1;
  mma_yz(S, a, c);
  lamd[d] = S[B_MLAM];
}
__global__ void k_bsc_mma_resid(int nseg, double *sc, double a) {
#pragma clang fp contract(off)
  FOR_DESIGN(nseg)
  if (S[B_MSTEP] == 0.0) return;
  const double r1 = S[B_RED] - S[B_MB] - a * S[B_MZ] - S[B_MY] + S[B_MMU];
  const double r2 = S[B_MMU] * S[B_MLAM] - S[B_MEPS];
  S[B_MERR] = fmax(fabs(r1), fabs(r2));
}
#undef FOR_DESIGN

// ---------------------------------------------------------------- coarsest level of a batch
// Each design's block of the block-diagonal coarsest operator, dense and
// row-major at dense + coff[d] (n_d = 6 (cu[d+1] - cu[d]) rows).
__global__ __launch_bounds__(TPB) void k_seg_bsr_to_dense(int64_t nrows, const int32_t *__restrict__ rp,
                                                          const int32_t *__restrict__ ci,
                                                          const double *__restrict__ val,
                                                          const int32_t *__restrict__ udes,
                                                          const int32_t *__restrict__ cu,
                                                          const int64_t *__restrict__ coff, double *__restrict__ D) {
  const int64_t row = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (row >= nrows * 6) return;
  const int64_t A = row / 6;
  const int i = (int)(row - A * 6);
  const int d = udes[A];
  const int64_t u0 = cu[d], n = 6 * (int64_t)(cu[d + 1] - cu[d]);
  double *Dr = D + coff[d] + ((A - u0) * 6 + i) * n;
  for (int64_t c = 0; c < n; ++c) Dr[c] = 0.0;
  for (int k = rp[A]; k < rp[A + 1]; ++k) {
    const int64_t B = ci[k] - u0;
    if (B < 0 || 6 * B >= n) continue;  // designs do not couple; never write outside the block
#pragma unroll
    for (int j = 0; j < 6; ++j) Dr[B * 6 + j] = val[(size_t)k * 36 + i * 6 + j];
  }
}

// In-place Gauss-Jordan inversion of each design's SPD block (no pivoting: an
// SPD matrix's pivots are positive), one workgroup per design; step k with
// p = a_kk:  a_kk <- 1/p, a_kj <- a_kj/p, a_ik <- -a_ik/p,
// a_ij <- a_ij - a_ik a_kj / p.  The result is symmetrised into Ainv.
constexpr int GJB_TPB = 1024;
__global__ __launch_bounds__(GJB_TPB) void k_seg_gj(const int32_t *__restrict__ cu, const int64_t *__restrict__ coff,
                                                    double *__restrict__ A, double *__restrict__ Ainv) {
  __shared__ double rowk[UFT_TAIL_MAX], colk[UFT_TAIL_MAX];
  const int d = blockIdx.x;
  const int n = 6 * (cu[d + 1] - cu[d]);
  double *M = A + coff[d];
  double *Mi = Ainv + coff[d];
  const int64_t nn = (int64_t)n * n;
  for (int k = 0; k < n; ++k) {
    for (int j = threadIdx.x; j < n; j += GJB_TPB) {
      rowk[j] = M[(int64_t)k * n + j];
      colk[j] = M[(int64_t)j * n + k];
    }
    __syncthreads();
    const double piv = 1.0 / rowk[k];
    for (int64_t t = threadIdx.x; t < nn; t += GJB_TPB) {
      const int i = (int)(t / n), j = (int)(t - (int64_t)i * n);
      double v;
      if (i == k)
        v = j == k ? piv : rowk[j] * piv;
      else
        v = j == k ? -colk[i] * piv : M[t] - colk[i] * rowk[j] * piv;
      M[t] = v;
    }
    __syncthreads();
  }
  for (int64_t t = threadIdx.x; t < nn; t += GJB_TPB) {
    const int i = (int)(t / n), j = (int)(t - (int64_t)i * n);
    Mi[t] = 0.5 * (M[t] + M[(int64_t)j * n + i]);
  }
}

inline unsigned gsz(SegG g) { return (unsigned)g.nb; }

}  // namespace

namespace uftk {
void seg_final(hipStream_t s, const double *part, const int32_t *dfirst, int nseg, int K, double *sc, int base,
               const double *gate) {
  switch (K) {
    case 1: k_seg_final<1><<<nseg, TPB, 0, s>>>(part, dfirst, sc, base, gate); break;
    case 2: k_seg_final<2><<<nseg, TPB, 0, s>>>(part, dfirst, sc, base, gate); break;
    default: k_seg_final<3><<<nseg, TPB, 0, s>>>(part, dfirst, sc, base, gate); break;
  }
}
void seg_final_max(hipStream_t s, const double *part, const int32_t *dfirst, int nseg, double *sc, int base) {
  k_seg_final_max<<<nseg, TPB, 0, s>>>(part, dfirst, sc, base);
}
void seg_dot(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, const double *a, const double *b,
             double *part, double *sc, int base) {
  k_seg_dot<<<gsz(g), TPB, 0, s>>>(g, n, a, b, part);
  seg_final(s, part, dfirst, nseg, 1, sc, base);
}
template <typename T>
static void seg_dots_t(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, int K, const T *const *a,
                       const T *const *b, double *part, double *sc, int base) {
  if (K == 2)
    k_seg_dots<2, T><<<gsz(g), TPB, 0, s>>>(g, n, a[0], b[0], a[1], b[1], nullptr, nullptr, part);
  else
    k_seg_dots<3, T><<<gsz(g), TPB, 0, s>>>(g, n, a[0], b[0], a[1], b[1], a[2], b[2], part);
  seg_final(s, part, dfirst, nseg, K, sc, base);
}
void seg_dots(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, int K, const double *const *a,
              const double *const *b, double *part, double *sc, int base) {
  seg_dots_t<double>(s, g, dfirst, nseg, n, K, a, b, part, sc, base);
}
void seg_dots(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, int K, const float *const *a,
              const float *const *b, double *part, double *sc, int base) {
  seg_dots_t<float>(s, g, dfirst, nseg, n, K, a, b, part, sc, base);
}
void bsc_kc1(hipStream_t s, int nseg, double *sc, int base) { k_bsc_kc1<<<(nseg + 63) / 64, 64, 0, s>>>(nseg, sc, base); }
void bsc_kc2(hipStream_t s, int nseg, double *sc, int base) { k_bsc_kc2<<<(nseg + 63) / 64, 64, 0, s>>>(nseg, sc, base); }
template <typename T>
static void seg_lincomb_first_t(hipStream_t s, SegG g, int64_t n, const double *sc, int ic, const T *b, const T *v,
                                T *kr, T *r, const T *dinv, double it, T *d, T *kd) {
  k_seg_lincomb_first<T><<<gsz(g), TPB, 0, s>>>(g, n, sc, ic, b, v, kr, r, dinv, it, d, kd);
}
void seg_lincomb_first(hipStream_t s, SegG g, int64_t n, const double *sc, int ic, const double *b, const double *v,
                       double *kr, double *r, const double *dinv, double it, double *d, double *kd) {
  seg_lincomb_first_t<double>(s, g, n, sc, ic, b, v, kr, r, dinv, it, d, kd);
}
void seg_lincomb_first(hipStream_t s, SegG g, int64_t n, const double *sc, int ic, const float *b, const float *v,
                       float *kr, float *r, const float *dinv, double it, float *d, float *kd) {
  seg_lincomb_first_t<float>(s, g, n, sc, ic, b, v, kr, r, dinv, it, d, kd);
}
template <typename T>
static void seg_prolong_kc_t(hipStream_t s, SegG g, int bs, int64_t n, const int32_t *parent, const T *P,
                             const double *sc, int i5, const T *xa, const T *xb, T *x) {
  if (bs == 3)
    k_seg_prolong_kc<3, T><<<gsz(g), TPB, 0, s>>>(g, n, parent, P, sc, i5, xa, xb, x);
  else
    k_seg_prolong_kc<6, T><<<gsz(g), TPB, 0, s>>>(g, n, parent, P, sc, i5, xa, xb, x);
}
void seg_prolong_kc(hipStream_t s, SegG g, int bs, int64_t n, const int32_t *parent, const double *P, const double *sc,
                    int i5, const double *xa, const double *xb, double *x) {
  seg_prolong_kc_t<double>(s, g, bs, n, parent, P, sc, i5, xa, xb, x);
}
void seg_prolong_kc(hipStream_t s, SegG g, int bs, int64_t n, const int32_t *parent, const float *P, const double *sc,
                    int i5, const float *xa, const float *xb, float *x) {
  seg_prolong_kc_t<float>(s, g, bs, n, parent, P, sc, i5, xa, xb, x);
}
void seg_dot_pr(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, const double *r, const double *z,
                const double *zo, double *part, double *sc, int base) {
  k_seg_dot_pr<<<gsz(g), TPB, 0, s>>>(g, n, r, z, zo, part);
  seg_final(s, part, dfirst, nseg, 2, sc, base);
}
void seg_cg_update(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, double *sc, int ia, double *x,
                   const double *p, double *r, const double *q, double *part, int base) {
  k_seg_cg_update<<<gsz(g), TPB, 0, s>>>(g, n, sc, ia, x, p, r, q, part);
  seg_final(s, part, dfirst, nseg, 1, sc, base);
}
void seg_xpby(hipStream_t s, SegG g, int64_t n, const double *z, const double *sc, int ib, double *p) {
  k_seg_xpby<<<gsz(g), TPB, 0, s>>>(g, n, z, sc, ib, p);
}
void seg_fill(hipStream_t s, SegG g, int64_t n, const double *vals, double *x) {
  k_seg_fill<<<gsz(g), TPB, 0, s>>>(g, n, vals, x);
}
void seg_scale_dinv(hipStream_t s, SegG g, int64_t n, const double *fac, double *dinv, float *dinvf) {
  k_seg_scale_dinv<<<gsz(g), TPB, 0, s>>>(g, n, fac, dinv, dinvf);
}
void seg_fcg_init(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, const double *b, const double *q,
                  const double *dinv, double *r, double *z, double *part, double *sc, int base) {
  k_seg_fcg_init<<<gsz(g), TPB, 0, s>>>(g, n, b, q, dinv, r, z, part);
  seg_final(s, part, dfirst, nseg, 3, sc, base);
}
void seg_fcg_update(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, double *sc, int ia, double *x,
                    const double *p, double *r, const double *q, const double *dinv, double *z, double *part,
                    int base) {
  k_seg_fcg_update<<<gsz(g), TPB, 0, s>>>(g, n, sc, ia, x, p, r, q, dinv, z, part);
  seg_final(s, part, dfirst, nseg, 2, sc, base);
}
void seg_fcg_p(hipStream_t s, SegG g, int64_t n, const double *z, const double *sc, int ic, int first, double *p) {
  k_seg_fcg_p<<<gsz(g), TPB, 0, s>>>(g, n, z, sc, ic, first, p);
}
void bsc_el_init(hipStream_t s, int nseg, double *sc, double rtol) { k_bsc_el_init<<<1, 64, 0, s>>>(nseg, sc, rtol); }
void bsc_el_alpha(hipStream_t s, int nseg, double *sc) { k_bsc_el_alpha<<<1, 64, 0, s>>>(nseg, sc); }
void bsc_el_check(hipStream_t s, int nseg, double *sc, double *flags) {
  k_bsc_el_check<<<1, 64, 0, s>>>(nseg, sc, flags);
}
void bsc_el_beta(hipStream_t s, int nseg, double *sc) { k_bsc_el_beta<<<1, 64, 0, s>>>(nseg, sc); }
void bsc_el_check_lo(hipStream_t s, int nseg, double *sc, double *flags) {
  k_bsc_el_check_lo<<<1, 64, 0, s>>>(nseg, sc, flags);
}
void bsc_el_true(hipStream_t s, int nseg, double *sc, double *flags) {
  k_bsc_el_true<<<1, 64, 0, s>>>(nseg, sc, flags);
}
void bsc_mma_init(hipStream_t s, int nseg, double *sc, const double *gt, double a, double c, double *lamd) {
  k_bsc_mma_init<<<1, 64, 0, s>>>(nseg, sc, gt, a, c, lamd);
}
void bsc_mma_decide(hipStream_t s, int nseg, double *sc, double a, double c, double *lamd, double *flags) {
  k_bsc_mma_decide<<<1, 64, 0, s>>>(nseg, sc, a, c, lamd, flags);
}
void bsc_mma_step(hipStream_t s, int nseg, double *sc, double a, double c, double *lamd) {
  k_bsc_mma_step<<<1, 64, 0, s>>>(nseg, sc, a, c, lamd);
}
void bsc_mma_resid(hipStream_t s, int nseg, double *sc, double a) { k_bsc_mma_resid<<<1, 64, 0, s>>>(nseg, sc, a); }
void seg_true_res(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, double *sc, const double *F,
                  const double *q, double *r, double *part, int base, const double *gate) {
  k_seg_true_res<<<gsz(g), TPB, 0, s>>>(g, n, sc, F, q, r, part, gate);
  seg_final(s, part, dfirst, nseg, 1, sc, base, gate);
}
void seg_dot_lo(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, const double *a, const float *b,
                double *part, double *sc, int base) {
  k_seg_dot_lo<<<gsz(g), TPB, 0, s>>>(g, n, a, b, part);
  seg_final(s, part, dfirst, nseg, 1, sc, base);
}
void seg_cg_update_lo(hipStream_t s, SegG g, const int32_t *dfirst, int nseg, int64_t n, double *sc, int ia, double *x,
                      const double *p, double *r, const float *q, double *part, int base) {
  k_seg_cg_update_lo<<<gsz(g), TPB, 0, s>>>(g, n, sc, ia, x, p, r, q, part);
  seg_final(s, part, dfirst, nseg, 1, sc, base);
}
void bsc_f_init(hipStream_t s, int nseg, double *sc, double rtol) { k_bsc_f_init<<<1, 64, 0, s>>>(nseg, sc, rtol); }
void bsc_f_alpha(hipStream_t s, int nseg, double *sc) { k_bsc_f_alpha<<<1, 64, 0, s>>>(nseg, sc); }
void bsc_f_after(hipStream_t s, int nseg, double *sc) { k_bsc_f_after<<<1, 64, 0, s>>>(nseg, sc); }
void bsc_lanczos_alpha(hipStream_t s, int nseg, double *sc, double *log, int step) {
  k_bsc_l_alpha<<<1, 64, 0, s>>>(nseg, sc, log, step);
}
void bsc_lanczos_beta(hipStream_t s, int nseg, double *sc, double *log, int step) {
  k_bsc_l_beta<<<1, 64, 0, s>>>(nseg, sc, log, step);
}
void seg_bsr_to_dense(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const double *val,
                      const int32_t *udes, const int32_t *cu, const int64_t *coff, double *dense) {
  k_seg_bsr_to_dense<<<(unsigned)((nb_rows * 6 + TPB - 1) / TPB), TPB, 0, s>>>(nb_rows, rp, ci, val, udes, cu, coff,
                                                                              dense);
}
void seg_dense_inverse(hipStream_t s, int nseg, const int32_t *cu, const int64_t *coff, double *A, double *Ainv) {
  k_seg_gj<<<nseg, GJB_TPB, 0, s>>>(cu, coff, A, Ainv);
}
}  // namespace uftk
