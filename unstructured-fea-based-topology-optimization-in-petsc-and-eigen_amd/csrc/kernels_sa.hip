// gfx950 kernels of the smoothed-aggregation prolongator of level 0.
//
// PCGAMG (SolveFem.cpp:129-136) smooths its tentative prolongator once with
// the current operator: P = (I - omega D^-1 A) P_tent, omega = 4 / (3
// lambda_max(D^-1 A)), and forms the coarse operator P^T A P; the reference
// rebuilds it at every design iteration (SolveFem.cpp:260-262).  Here, at
// every hierarchy refresh:
//   k_sa_prolongator  W = A P_tent per (node, aggregate) entry of P, then the
//                     smoothed entry (rows of fixed DOFs zero), stored for the
//                     products below and in the cycle's two layouts
//   k_sa_ap           V = A P per (node, aggregate) entry of A P
//   k_sa_galerkin     A_1[I][J] = sum over the support nodes i of I of
//                     P_i[I]^T V_i[J], one wavefront per upper block, lanes
//                     over its contribution pairs, a fixed butterfly combine;
//                     the lower block is written transposed, diagonal blocks
//                     symmetrised
// Each runs over contribution lists built once at setup from the patterns
// (k_sa_scan_*), so the refresh does no searching.  A is the fine operator
// assembled on the node pattern (3x3 blocks, fixed columns zeroed; a
// fixed row of A is the identity row, so its rows of A P_tent, P and A P
// vanish because P_tent vanishes there).  Every sum runs in a fixed order:
// results do not depend on scheduling.
// The cycle's transfers between levels 0 and 1 use P through its supports
// (restriction, fused with level 1's first Chebyshev step) and its rows
// (prolongation with the K-cycle's combination of level 1).
#include <type_traits>

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "uft_internal.h"
#include "uft_reduce.h"

namespace {

// storage of the prolongator copies: FP64, FP32 or FP16 (UFT_MIXED_H)
__device__ __forceinline__ double ld(const double *p) { return *p; }
__device__ __forceinline__ double ld(const float *p) { return (double)*p; }
__device__ __forceinline__ double ld(const __half *p) { return (double)__half2float(*p); }
__device__ __forceinline__ float ldf(const float *p) { return *p; }
__device__ __forceinline__ float ldf(const __half *p) { return __half2float(*p); }
__device__ __forceinline__ double ldd(const double *p) { return *p; }
template <typename S> __device__ __forceinline__ S st(double v);
template <> __device__ __forceinline__ double st<double>(double v) { return v; }
template <> __device__ __forceinline__ float st<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ __half st<__half>(double v) { return __float2half((float)v); }

constexpr int TPB = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// slot q of node n on the sliced-ELL node pattern
__device__ __forceinline__ int64_t ell_slot(const int64_t *off, int n, int q) {
  return off[n / UFT_ELL_SLICE] + (int64_t)q * UFT_ELL_SLICE + n % UFT_ELL_SLICE;
}

// ---------------------------------------------------------------- setup: lists
// which 0, per prolongator entry t = (i, K): the slots of row i whose column
// lies in K, in slot order
// which 1, per entry t = (i, J) of A P: (slot of row i, support position of
// the column's prolongator entry naming J), in slot order
// which 2, per upper block (I, J): (support position q of I at node i,
// entry of A P of node i naming J), in support order
// count: a = per-row counts; fill: a, b = the lists at the rows' offsets ptr
__global__ __launch_bounds__(TPB) void k_sa_scan_w(int64_t np, const int32_t *__restrict__ pnode,
                                                   const int32_t *__restrict__ pagg, const int64_t *__restrict__ off,
                                                   const int32_t *__restrict__ width, const int32_t *__restrict__ col,
                                                   const int32_t *__restrict__ cptr, const int32_t *__restrict__ parent,
                                                   const int32_t *__restrict__ ptr, int32_t *__restrict__ a) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= np) return;
  const int i = pnode[t], K = pagg[t];
  const int w = width[i / UFT_ELL_SLICE];
  int c = 0;
  for (int q = 0; q < w; ++q) {
    const int64_t k = ell_slot(off, i, q);
    if (cptr[k + 1] == cptr[k] || parent[col[k]] != K) continue;
    if (ptr) a[ptr[t] + c] = (int32_t)k;
    ++c;
  }
  if (!ptr) a[t] = c;
}

__global__ __launch_bounds__(TPB) void k_sa_scan_v(int64_t nv, const int32_t *__restrict__ vnode,
                                                   const int32_t *__restrict__ vagg, const int64_t *__restrict__ off,
                                                   const int32_t *__restrict__ width, const int32_t *__restrict__ col,
                                                   const int32_t *__restrict__ cptr, const int32_t *__restrict__ pptr,
                                                   const int32_t *__restrict__ pagg, const int32_t *__restrict__ spos,
                                                   const int32_t *__restrict__ ptr, int32_t *__restrict__ a,
                                                   int32_t *__restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= nv) return;
  const int i = vnode[t], J = vagg[t];
  const int w = width[i / UFT_ELL_SLICE];
  int c = 0;
  for (int q = 0; q < w; ++q) {
    const int64_t k = ell_slot(off, i, q);
    if (cptr[k + 1] == cptr[k]) continue;
    const int j = col[k];
    for (int e = pptr[j]; e < pptr[j + 1]; ++e) {
      const int g = pagg[e];
      if (g < J) continue;
      if (g == J) {
        if (ptr) {
          a[ptr[t] + c] = (int32_t)k;
          b[ptr[t] + c] = spos[e];
        }
        ++c;
      }
      break;
    }
  }
  if (!ptr) a[t] = c;
}

__global__ __launch_bounds__(TPB) void k_sa_scan_g(int64_t nub, const int32_t *__restrict__ ublk,
                                                   const int32_t *__restrict__ brow, const int32_t *__restrict__ bci,
                                                   const int32_t *__restrict__ sptr, const int32_t *__restrict__ snode,
                                                   const int32_t *__restrict__ sidx, const int32_t *__restrict__ vptr,
                                                   const int32_t *__restrict__ vagg, const int32_t *__restrict__ ptr,
                                                   int32_t *__restrict__ a, int32_t *__restrict__ b) {
  const int64_t u = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (u >= nub) return;
  const int bk = ublk[u];
  const int I = brow[bk], J = bci[bk];
  int c = 0;
  for (int q = sptr[I]; q < sptr[I + 1]; ++q) {
    const int i = snode[q];
    int lo = vptr[i], hi = vptr[i + 1];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (vagg[mid] < J) lo = mid + 1;
      else hi = mid;
    }
    if (lo >= vptr[i + 1] || vagg[lo] != J) continue;
    if (ptr) {
      a[ptr[u] + c] = q;  // support position
      b[ptr[u] + c] = lo;
    }
    ++c;
  }
  if (!ptr) a[u] = c;
}

// ---------------------------------------------------------------- refresh
// The refresh kernels read and write in the cycle's precision T (FP64 in the
// FP64 form, FP32 in the mixed forms, whose coarse operators are FP32 or
// lower) and accumulate in FP64.  Their blocks are handed out XCD by XCD in
// contiguous ranges (block b runs on XCD b mod 8), so neighbouring rows,
// which share their operands, share one L2.
using uftr::xcd_block;

// The FP16 node-sliced copy (UFT_MIXED_H) is stored in three planes over
// the entry slots s (ne of them): values 0-7 at halves 8 s, values 8-15 at
// 8 (ne + s), values 16-17 at 16 ne + 2 s, so a lane reads its 3x6 entry
// with two 16-byte loads and one 4-byte load, each coalesced over the
// wavefront's 64 consecutive slots.
__device__ __forceinline__ int64_t pe_half_index(int64_t ne, int64_t s, int v) {
  return v < 8 ? 8 * s + v : v < 16 ? 8 * (ne + s) + (v - 8) : 16 * ne + 2 * s + (v - 16);
}

// two FP16 values packed in a 32-bit word
__device__ __forceinline__ float2 unpack_h2(uint32_t w) {
  __half2 h;
  __builtin_memcpy(&h, &w, 4);
  return __half22float2(h);
}

// acc (3x6) += A (3x3, row-major) * B (3x6, row-major)
template <typename T, typename S>
__device__ __forceinline__ void mac33x36t(double (&acc)[18], const T *__restrict__ A, const S *__restrict__ B) {
  double a[9], b[18];
#pragma unroll
  for (int t = 0; t < 9; ++t) a[t] = (double)A[t];
#pragma unroll
  for (int t = 0; t < 18; ++t) b[t] = ld(B + t);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int m = 0; m < 6; ++m) acc[6 * c + m] += a[3 * c] * b[m] + a[3 * c + 1] * b[6 + m] + a[3 * c + 2] * b[12 + m];
}

// one thread per prolongator entry t = (node i, aggregate K):
// W = sum over its slots of A_slot P_tent(col); P[t] = [K == parent(i)]
// P_tent(i) - omega D^-1 W, stored as the node-sliced copy and the
// support-ordered copy
template <typename T, typename S>
__global__ __launch_bounds__(TPB) void k_sa_prolongator(int64_t np, const int32_t *__restrict__ pnode,
                                                        const int32_t *__restrict__ pagg,
                                                        const int32_t *__restrict__ pptr,
                                                        const int32_t *__restrict__ wptr,
                                                        const int32_t *__restrict__ wslot,
                                                        const int32_t *__restrict__ col, const T *__restrict__ a,
                                                        const int32_t *__restrict__ parent, const T *__restrict__ Pt,
                                                        const double *__restrict__ dinv,
                                                        const uint8_t *__restrict__ fixm, double omega,
                                                        const int64_t *__restrict__ poff, int64_t ne_ell,
                                                        const int32_t *__restrict__ spos, S *__restrict__ Pe,
                                                        S *__restrict__ Ps) {
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
  if (t >= np) return;
  const int i = pnode[t];
  const int K = pagg[t];
  double acc[18];
#pragma unroll
  for (int m = 0; m < 18; ++m) acc[m] = 0.0;
  for (int q = wptr[t]; q < wptr[t + 1]; ++q) {
    const int64_t k = wslot[q];
    mac33x36t<T, T>(acc, a + 9 * k, Pt + 18 * (size_t)col[k]);
  }
  const bool own = parent[i] == K;
  const unsigned fm = fixm[i];
  const int q = (int)(t - pptr[i]);
  const int64_t slot = poff[i / UFT_ELL_SLICE] + (int64_t)q * UFT_ELL_SLICE + i % UFT_ELL_SLICE;
  S *pe = Pe + (slot - i % UFT_ELL_SLICE) * 18 + i % UFT_ELL_SLICE;
  S *ps = Ps + (size_t)spos[t] * 18;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double di = dinv[3 * (size_t)i + c];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const double pt = own ? (double)Pt[18 * (size_t)i + 6 * c + m] : 0.0;
      const S v = st<S>(((fm >> c) & 1u) ? 0.0 : pt - omega * di * acc[6 * c + m]);
      if constexpr (std::is_same<S, __half>::value)
        Pe[pe_half_index(ne_ell, slot, 6 * c + m)] = v;
      else
        pe[(6 * c + m) * UFT_ELL_SLICE] = v;
      ps[6 * c + m] = v;
    }
  }
}

// one thread per entry t of A P: the sum over its (slot, support position)
// pairs of A_slot P_position (rows of fixed DOFs zero)
template <typename T, typename S>
__global__ __launch_bounds__(TPB) void k_sa_ap(int64_t nv, const int32_t *__restrict__ vnode,
                                               const int32_t *__restrict__ vlptr, const int32_t *__restrict__ vlslot,
                                               const int32_t *__restrict__ vlpos, const T *__restrict__ a,
                                               const S *__restrict__ Ps, const uint8_t *__restrict__ fixm,
                                               T *__restrict__ V) {
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
  if (t >= nv) return;
  double acc[18];
#pragma unroll
  for (int m = 0; m < 18; ++m) acc[m] = 0.0;
  for (int q = vlptr[t]; q < vlptr[t + 1]; ++q)
    mac33x36t<T, S>(acc, a + 9 * (size_t)vlslot[q], Ps + 18 * (size_t)vlpos[q]);
  const unsigned fm = fixm[vnode[t]];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int m = 0; m < 6; ++m) V[18 * t + 6 * c + m] = (T)(((fm >> c) & 1u) ? 0.0 : acc[6 * c + m]);
}

// one wavefront per upper block b = (I, J), I <= J, of the level-1 pattern;
// its pairs (support position of I, entry of A P) run in support order
template <typename T, typename S>
__global__ __launch_bounds__(TPB) void k_sa_galerkin(int64_t nub, const int32_t *__restrict__ ublk,
                                                     const int32_t *__restrict__ brow, const int32_t *__restrict__ bci,
                                                     const int32_t *__restrict__ btr, const int32_t *__restrict__ gptr,
                                                     const int32_t *__restrict__ gp, const int32_t *__restrict__ gv,
                                                     const S *__restrict__ Ps, const T *__restrict__ V,
                                                     double *__restrict__ val) {
  const int64_t wv = (xcd_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (wv >= nub) return;  // uniform per wavefront
  const int b = ublk[wv];
  double acc[36];
#pragma unroll
  for (int e = 0; e < 36; ++e) acc[e] = 0.0;
  for (int q = gptr[wv] + lane; q < gptr[wv + 1]; q += 64) {
    const S *p = Ps + 18 * (size_t)gp[q];
    const T *v = V + 18 * (size_t)gv[q];
    double pr[18], vr[18];
#pragma unroll
    for (int m = 0; m < 18; ++m) {
      pr[m] = ld(p + m);
      vr[m] = (double)v[m];
    }
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int bb = 0; bb < 6; ++bb)
        acc[6 * a + bb] += pr[a] * vr[bb] + pr[6 + a] * vr[6 + bb] + pr[12 + a] * vr[12 + bb];
  }
#pragma unroll
  for (int e = 0; e < 36; ++e) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[e] += __shfl_xor(acc[e], o, 64);
  }
  if (lane != 0) return;
  const int I = brow[b], J = bci[b];
  double *vb = val + 36 * (size_t)b;
  if (I == J) {
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int bb = 0; bb < 6; ++bb) vb[6 * a + bb] = 0.5 * (acc[6 * a + bb] + acc[6 * bb + a]);
    return;
  }
  double *vt = val + 36 * (size_t)btr[b];
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int bb = 0; bb < 6; ++bb) {
      vb[6 * a + bb] = acc[6 * a + bb];
      vt[6 * bb + a] = acc[6 * a + bb];
    }
}

// ---------------------------------------------------------------- cycle transfers
// G lanes per coarse unit: bc_I = sum over the support positions of I of
// P^T r at the position's node (P in support order, each position's 3x6 block
// contiguous), then (dinv given) level 1's first Chebyshev step from a zero
// guess: rc = bc, dc = it D^-1 bc, xc = dc
template <int G, typename T, typename S>
__global__ __launch_bounds__(TPB) void k_sa_restrict_first(int64_t nc, const int32_t *__restrict__ sptr,
                                                           const int32_t *__restrict__ snode, const S *__restrict__ Ps,
                                                           const T *__restrict__ r, T *__restrict__ bc,
                                                           const T *__restrict__ dinv, double it, T *__restrict__ rc,
                                                           T *__restrict__ dc, T *__restrict__ xc) {
  static_assert(G >= 1 && G <= 64 && (64 % G) == 0, "G lanes must tile a wavefront");
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int64_t A = t / G;
  const int lane = (int)(t % G);
  const bool on = A < nc;
  const int k0 = on ? sptr[A] : 0, npair = on ? (sptr[A + 1] - k0) * 3 : 0;
  T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int q = lane; q < npair; q += G) {
    const int kk = q / 3, c = q - kk * 3;
    const int i = snode[k0 + kk];
    const S *Pe = Ps + (size_t)(k0 + kk) * 18 + c * 6;
    const T ri = r[(size_t)i * 3 + c];
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[j] += (T)ld(Pe + j) * ri;
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc[j] += __shfl_xor(acc[j], o, 64);
  }
  if (on && lane == 0) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int64_t i = A * 6 + j;
      bc[i] = acc[j];
      if (dinv) {
        const T d = (T)it * dinv[i] * acc[j];
        rc[i] = acc[j];
        dc[i] = d;
        xc[i] = d;
      }
    }
  }
}

// one thread per node u: x_u += sum over its entry slots of P (ca xa + cb
// xb)_{aggregate}; ca, cb from sc[ia], sc[ib] (index < 0: 1.0); xb may be
// null.  P in the node-sliced layout: one load instruction reads a value of
// 64 consecutive nodes.
template <typename T, typename S>
__global__ __launch_bounds__(TPB) void k_sa_prolong_comb(int64_t nu, const int64_t *__restrict__ poff,
                                                         const int32_t *__restrict__ pw,
                                                         const int32_t *__restrict__ eagg, const S *__restrict__ Pe,
                                                         const double *__restrict__ sc, int ia, int ib,
                                                         const T *__restrict__ xa, const T *__restrict__ xb,
                                                         T *__restrict__ x, const int32_t *__restrict__ ndes) {
  const int64_t u = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (u >= nu) return;
  const double *scu = ndes ? sc + (size_t)ndes[u] * UFT_BSC : sc;  // a batch: the node's design's scalars
  const T ca = ia < 0 ? T(1) : (T)scu[ia];
  const T cb = ib < 0 ? T(1) : (T)scu[ib];
  const int64_t w = u / UFT_ELL_SLICE, lane = u % UFT_ELL_SLICE;
  const int nq = pw[w];
  T acc[3] = {T(0), T(0), T(0)};
  for (int q = 0; q < nq; ++q) {
    const int64_t s = poff[w] + (int64_t)q * UFT_ELL_SLICE + lane;
    const int64_t A = eagg[s];
    if (A < 0) break;  // padding follows the node's entries
    T c[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) c[j] = xb ? ca * xa[A * 6 + j] + cb * xb[A * 6 + j] : xa[A * 6 + j];
    const S *Pq = Pe + poff[w] * 18 + (int64_t)q * 18 * UFT_ELL_SLICE + lane;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) acc[i] += (T)ld(Pq + (i * 6 + j) * UFT_ELL_SLICE) * c[j];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) x[u * 3 + i] += acc[i];
}

// FP16 copies (UFT_MIXED_H): G lanes per coarse unit, one support node per
// lane and step (its whole 3x6 entry: nine 4-byte words in support order and
// the node's three residual entries), then the same butterfly and first
// Chebyshev step as k_sa_restrict_first
template <int G>
__global__ __launch_bounds__(TPB) void k_sa_restrict_first_h(int64_t nc, const int32_t *__restrict__ sptr,
                                                             const int32_t *__restrict__ snode,
                                                             const uint32_t *__restrict__ Ps,
                                                             const float *__restrict__ r, float *__restrict__ bc,
                                                             const float *__restrict__ dinv, double it,
                                                             float *__restrict__ rc, float *__restrict__ dc,
                                                             float *__restrict__ xc) {
  static_assert(G >= 1 && G <= 64 && (64 % G) == 0, "G lanes must tile a wavefront");
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int64_t A = t / G;
  const int lane = (int)(t % G);
  const bool on = A < nc;
  const int k1 = on ? sptr[A + 1] : 0;
  float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int k = (on ? sptr[A] : 0) + lane; k < k1; k += G) {
    const int64_t i = snode[k];
    const uint32_t *pw = Ps + (int64_t)k * 9;
    float p[18];
#pragma unroll
    for (int h = 0; h < 9; ++h) {
      const float2 v = unpack_h2(pw[h]);
      p[2 * h] = v.x;
      p[2 * h + 1] = v.y;
    }
    const float r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[j] += p[j] * r0 + p[6 + j] * r1 + p[12 + j] * r2;
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc[j] += __shfl_xor(acc[j], o, 64);
  }
  if (on && lane == 0) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int64_t i = A * 6 + j;
      bc[i] = acc[j];
      if (dinv) {
        const float d = (float)it * dinv[i] * acc[j];
        rc[i] = acc[j];
        dc[i] = d;
        xc[i] = d;
      }
    }
  }
}

// x_u += sum of P (ca xa + cb xb)_{aggregate} (as k_sa_prolong_comb) from
// the FP16 node-sliced copy's three planes (pe_half_index).  The node's
// entries are taken TG at a time with every load of the group (aggregate
// ids, P planes, the aggregates' six coarse entries) issued before the
// products; padding slots (aggregate -1) read aggregate 0 with zero
// weights, so no branch separates the loads.  TWO: a second coarse vector xb.
template <bool TWO>
__global__ __launch_bounds__(TPB) void k_sa_prolong_comb_h(int64_t nu, int64_t ne, const int64_t *__restrict__ poff,
                                                           const int32_t *__restrict__ pw,
                                                           const int32_t *__restrict__ eagg,
                                                           const uint4 *__restrict__ P01,
                                                           const uint32_t *__restrict__ P2,
                                                           const double *__restrict__ sc, int ia, int ib,
                                                           const float *__restrict__ xa, const float *__restrict__ xb,
                                                           float *__restrict__ x, const int32_t *__restrict__ ndes) {
  constexpr int TG = 4;
  const int64_t u = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (u >= nu) return;
  const double *scu = ndes ? sc + (size_t)ndes[u] * UFT_BSC : sc;  // a batch: the node's design's scalars
  const float ca = ia < 0 ? 1.f : (float)scu[ia];
  const float cb = ib < 0 ? 1.f : (float)scu[ib];
  const int64_t w = u / UFT_ELL_SLICE;
  const int nq = pw[w];
  const int64_t s0 = poff[w] + u % UFT_ELL_SLICE;
  float a0 = 0.f, a1 = 0.f, a2 = 0.f;
  for (int q0 = 0; q0 < nq; q0 += TG) {
    int A[TG];
    uint32_t wds[TG][9];
    float c[TG][6];
#pragma unroll
    for (int g = 0; g < TG; ++g) {
      const int64_t s = s0 + (int64_t)min(q0 + g, nq - 1) * UFT_ELL_SLICE;
      A[g] = eagg[s];
      const uint4 lo = P01[s], hi = P01[ne + s];
      wds[g][0] = lo.x; wds[g][1] = lo.y; wds[g][2] = lo.z; wds[g][3] = lo.w;
      wds[g][4] = hi.x; wds[g][5] = hi.y; wds[g][6] = hi.z; wds[g][7] = hi.w;
      wds[g][8] = P2[s];
    }
#pragma unroll
    for (int g = 0; g < TG; ++g) {
      const int64_t Ag = A[g] >= 0 ? A[g] : 0;
#pragma unroll
      for (int j = 0; j < 6; ++j) c[g][j] = TWO ? ca * xa[6 * Ag + j] + cb * xb[6 * Ag + j] : xa[6 * Ag + j];
    }
#pragma unroll
    for (int g = 0; g < TG; ++g) {
      const bool ok = q0 + g < nq && A[g] >= 0;
      float p[18];
#pragma unroll
      for (int h = 0; h < 9; ++h) {
        const float2 v = unpack_h2(wds[g][h]);
        p[2 * h] = ok ? v.x : 0.f;
        p[2 * h + 1] = ok ? v.y : 0.f;
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        a0 += p[j] * c[g][j];
        a1 += p[6 + j] * c[g][j];
        a2 += p[12 + j] * c[g][j];
      }
    }
  }
  x[u * 3] += a0;
  x[u * 3 + 1] += a1;
  x[u * 3 + 2] += a2;
}

// ---- the same transfers for NR interleaved right-hand sides (the
// multi-load-case form, vectors [row][NR]): every P entry is read once for
// all right-hand sides.
// P entry e (3x6) of the support-ordered copy, in FP32 registers
template <typename S>
__device__ __forceinline__ void ld_entry_s(const S *Ps, int64_t e, float (&p)[18]) {
  if constexpr (std::is_same<S, __half>::value) {
    const uint32_t *pw = reinterpret_cast<const uint32_t *>(Ps) + e * 9;
#pragma unroll
    for (int h = 0; h < 9; ++h) {
      const float2 v = unpack_h2(pw[h]);
      p[2 * h] = v.x;
      p[2 * h + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int v = 0; v < 18; ++v) p[v] = (float)Ps[e * 18 + v];
  }
}
// G lanes per coarse unit, one support node per lane and step, all NR
// right-hand sides per lane; then the butterfly over the G lanes and level
// 1's first Chebyshev step per right-hand side
template <int G, typename T, typename S, int NR>
__global__ __launch_bounds__(TPB) void k_mr_sa_restrict_first(int64_t nc, const int32_t *__restrict__ sptr,
                                                              const int32_t *__restrict__ snode,
                                                              const S *__restrict__ Ps, const T *__restrict__ r,
                                                              T *__restrict__ bc, const T *__restrict__ dinv,
                                                              double it, T *__restrict__ rc, T *__restrict__ dc,
                                                              T *__restrict__ xc) {
  static_assert(G >= 1 && G <= 64 && (64 % G) == 0, "G lanes must tile a wavefront");
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int64_t A = t / G;
  const int lane = (int)(t % G);
  const bool on = A < nc;
  const int k1 = on ? sptr[A + 1] : 0;
  T acc[6][NR];
#pragma unroll
  for (int j = 0; j < 6; ++j)
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[j][q] = T(0);
  for (int k = (on ? sptr[A] : 0) + lane; k < k1; k += G) {
    const int64_t i = snode[k];
    float p[18];
    ld_entry_s<S>(Ps, k, p);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T *ri = r + (3 * i + c) * NR;
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        const T rv = ri[q];
#pragma unroll
        for (int j = 0; j < 6; ++j) acc[j][q] += (T)p[6 * c + j] * rv;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 6; ++j)
#pragma unroll
    for (int q = 0; q < NR; ++q) {
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) acc[j][q] += __shfl_xor(acc[j][q], o, 64);
    }
  if (on && lane == 0) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int64_t i = A * 6 + j;
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        const int64_t ii = i * NR + q;
        bc[ii] = acc[j][q];
        if (dinv) {
          const T d = (T)it * dinv[i] * acc[j][q];
          rc[ii] = acc[j][q];
          dc[ii] = d;
          xc[ii] = d;
        }
      }
    }
  }
}
// c = ca_q xa + cb_q xb on a coarse vector of NR interleaved right-hand sides
// (ca_q = sc[8 q + ia], index < 0: 1), ahead of the prolongation, so that it
// gathers one coarse vector per entry instead of two
template <typename T, int NR>
__global__ __launch_bounds__(TPB) void k_mr_lincomb2(int64_t n, const double *__restrict__ sc, int ia, int ib,
                                                     const T *__restrict__ xa, const T *__restrict__ xb,
                                                     T *__restrict__ c) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= n * NR) return;
  const int q = (int)(t % NR);
  const T ca = ia < 0 ? T(1) : (T)sc[8 * q + ia];
  const T cb = ib < 0 ? T(1) : (T)sc[8 * q + ib];
  c[t] = ca * xa[t] + cb * xb[t];
}
// one thread per node u (all NR right-hand sides): x_u += ca_q sum over its
// entry slots of P xa_{aggregate}, ca_q = sc[8 q + ia] (index < 0: 1; a
// second vector is combined into xa beforehand); P from the node-sliced copy
// (FP16: its three planes).  The entries are
// taken TG at a time with all their loads (P values, the aggregates' 6 NR
// coarse entries in 16-byte loads) issued before the products; entries past
// the node's last one (aggregate < 0, or past the slice width) read the last
// valid slot again with zero weights, so no branch separates the loads.
template <typename T, int NR>
__device__ __forceinline__ void ld_coarse6(const T *__restrict__ p, T (&o)[6][NR]) {
  if constexpr (std::is_same<T, float>::value && NR % 4 == 0) {
#pragma unroll
    for (int j = 0; j < 6; ++j)
#pragma unroll
      for (int q = 0; q < NR / 4; ++q) {
        const float4 t = reinterpret_cast<const float4 *>(p + j * NR)[q];
        o[j][4 * q] = t.x; o[j][4 * q + 1] = t.y; o[j][4 * q + 2] = t.z; o[j][4 * q + 3] = t.w;
      }
  } else {
#pragma unroll
    for (int j = 0; j < 6; ++j)
#pragma unroll
      for (int q = 0; q < NR; ++q) o[j][q] = p[j * NR + q];
  }
}
template <typename T, typename S, int NR>
__global__ __launch_bounds__(TPB) void k_mr_sa_prolong_comb(int64_t nu, int64_t ne, const int64_t *__restrict__ poff,
                                                            const int32_t *__restrict__ pw,
                                                            const int32_t *__restrict__ eagg,
                                                            const S *__restrict__ Pe, const double *__restrict__ sc,
                                                            int ia, const T *__restrict__ xa, T *__restrict__ x) {
  constexpr int TG = 2;
  const int64_t u = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (u >= nu) return;
  T ca[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) ca[q] = ia < 0 ? T(1) : (T)sc[8 * q + ia];
  const int64_t w = u / UFT_ELL_SLICE, lane = u % UFT_ELL_SLICE;
  const int nq = pw[w];
  T acc[3][NR];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[i][q] = T(0);
  int last = 0;  // the node's last valid entry slot seen so far (entries come first, padding after)
  for (int q0 = 0; q0 < nq; q0 += TG) {
    int64_t sl[TG], A[TG];
    bool ok[TG];
#pragma unroll
    for (int g = 0; g < TG; ++g) {
      const int qq = min(q0 + g, nq - 1);
      sl[g] = poff[w] + (int64_t)qq * UFT_ELL_SLICE + lane;
      A[g] = eagg[sl[g]];
      ok[g] = q0 + g < nq && A[g] >= 0;
      if (ok[g]) last = qq;
    }
    float p[TG][18];
    T xv[TG][6][NR];
#pragma unroll
    for (int g = 0; g < TG; ++g) {
      const int64_t s1 = ok[g] ? sl[g] : poff[w] + (int64_t)last * UFT_ELL_SLICE + lane;
      const int64_t a1 = ok[g] ? A[g] : 0;
      if constexpr (std::is_same<S, __half>::value) {
        const uint4 lo = reinterpret_cast<const uint4 *>(Pe)[s1], hi = reinterpret_cast<const uint4 *>(Pe)[ne + s1];
        const uint32_t tl = reinterpret_cast<const uint32_t *>(Pe)[8 * ne + s1];
        const uint32_t wds[9] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w, tl};
#pragma unroll
        for (int h = 0; h < 9; ++h) {
          const float2 v = unpack_h2(wds[h]);
          p[g][2 * h] = v.x;
          p[g][2 * h + 1] = v.y;
        }
      } else {
        const int qq = ok[g] ? min(q0 + g, nq - 1) : last;
        const S *Pq = Pe + poff[w] * 18 + (int64_t)qq * 18 * UFT_ELL_SLICE + lane;
#pragma unroll
        for (int v = 0; v < 18; ++v) p[g][v] = (float)Pq[v * UFT_ELL_SLICE];
      }
      ld_coarse6<T, NR>(xa + a1 * 6 * NR, xv[g]);
    }
#pragma unroll
    for (int g = 0; g < TG; ++g) {
#pragma unroll
      for (int v = 0; v < 18; ++v) p[g][v] = ok[g] ? p[g][v] : 0.f;
#pragma unroll
      for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int q = 0; q < NR; ++q)
#pragma unroll
          for (int i = 0; i < 3; ++i) acc[i][q] += (T)p[g][i * 6 + j] * xv[g][j][q];
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int q = 0; q < NR; ++q) x[(u * 3 + i) * NR + q] += ca[q] * acc[i][q];
}

}  // namespace

namespace uftk {

void sa_lists(hipStream_t s, const uft_ctx *C, int which, const int32_t *ptr, int32_t *a, int32_t *b) {
  const AmgLevel &L0 = C->amg.L[0], &L1 = C->amg.L[1];
  const auto &F = C->fell;
  if (which == 0)
    k_sa_scan_w<<<nblk(L0.sa_np), TPB, 0, s>>>(L0.sa_np, L0.d_sa_pnode, L0.d_sa_pagg, F.d_off, F.d_width, F.d_col,
                                               F.d_cptr, L0.d_parent, ptr, a);
  else if (which == 1)
    k_sa_scan_v<<<nblk(L0.sa_nv), TPB, 0, s>>>(L0.sa_nv, L0.d_sa_vnode, L0.d_sa_vagg, F.d_off, F.d_width, F.d_col,
                                               F.d_cptr, L0.d_sa_pptr, L0.d_sa_pagg, L0.d_sa_spos, ptr, a, b);
  else
    k_sa_scan_g<<<nblk(L0.sa_nub), TPB, 0, s>>>(L0.sa_nub, L0.d_sa_ublk, L1.d_brow, L1.d_ci, L0.d_sa_sptr,
                                                L0.d_sa_snode, L0.d_sa_sidx, L0.d_sa_vptr, L0.d_sa_vagg, ptr, a, b);
}

template <typename T, typename S>
static void sa_refresh_t(hipStream_t s, const uft_ctx *C, const T *a, const T *Pt, S *Pe, S *Ps, T *V) {
  const AmgLevel &L0 = C->amg.L[0], &L1 = C->amg.L[1];
  k_sa_prolongator<T, S><<<nblk(L0.sa_np), TPB, 0, s>>>(L0.sa_np, L0.d_sa_pnode, L0.d_sa_pagg, L0.d_sa_pptr,
                                                        L0.d_sa_wptr, L0.d_sa_wslot, C->fell.d_col, a, L0.d_parent, Pt,
                                                        L0.d_dinv, C->d_fixm, L0.sa_omega, L0.d_sa_poff, L0.sa_ne_ell,
                                                        L0.d_sa_spos, Pe, Ps);
  k_sa_ap<T, S><<<nblk(L0.sa_nv), TPB, 0, s>>>(L0.sa_nv, L0.d_sa_vnode, L0.d_sa_vlptr, L0.d_sa_vlslot, L0.d_sa_vlent,
                                               a, Ps, C->d_fixm, V);
  k_sa_galerkin<T, S><<<nblk(L0.sa_nub * 64), TPB, 0, s>>>(L0.sa_nub, L0.d_sa_ublk, L1.d_brow, L1.d_ci, L0.d_gtr,
                                                           L0.d_sa_gptr, L0.d_sa_gp, L0.d_sa_gv, Ps, V, L1.d_val);
}

void sa_refresh(hipStream_t s, uft_ctx *C) {
  AmgLevel &L0 = C->amg.L[0];
  L0.sa_omega_used = L0.sa_omega;
  if (L0.d_sa_Peh)
    sa_refresh_t<float, __half>(s, C, C->d_a32, L0.d_Pf, (__half *)L0.d_sa_Peh, (__half *)L0.d_sa_Psh, L0.d_sa_Vf);
  else if (L0.d_sa_Pef)
    sa_refresh_t<float, float>(s, C, C->d_a32, L0.d_Pf, L0.d_sa_Pef, L0.d_sa_Psf, L0.d_sa_Vf);
  else
    sa_refresh_t<double, double>(s, C, C->d_a64, L0.d_P, L0.d_sa_Pe, L0.d_sa_Ps, L0.d_sa_V);
}

void sa_restrict_first(hipStream_t s, const AmgLevel &L0, int64_t nc, const double *r, double *bc, const double *dinv,
                       double it, double *rc, double *dc, double *xc) {
  k_sa_restrict_first<16, double, double><<<nblk(nc * 16), TPB, 0, s>>>(nc, L0.d_sa_sptr, L0.d_sa_snode, L0.d_sa_Ps, r,
                                                                        bc, dinv, it, rc, dc, xc);
}
void sa_restrict_first(hipStream_t s, const AmgLevel &L0, int64_t nc, const float *r, float *bc, const float *dinv,
                       double it, float *rc, float *dc, float *xc) {
  if (L0.d_sa_Psh)
    k_sa_restrict_first_h<16><<<nblk(nc * 16), TPB, 0, s>>>(nc, L0.d_sa_sptr, L0.d_sa_snode,
                                                            (const uint32_t *)L0.d_sa_Psh, r, bc, dinv, it, rc, dc, xc);
  else
    k_sa_restrict_first<16, float, float><<<nblk(nc * 16), TPB, 0, s>>>(nc, L0.d_sa_sptr, L0.d_sa_snode, L0.d_sa_Psf, r,
                                                                        bc, dinv, it, rc, dc, xc);
}
void sa_prolong_comb(hipStream_t s, const AmgLevel &L0, const double *sc, int ia, int ib, const double *xa,
                     const double *xb, double *x, const int32_t *ndes) {
  k_sa_prolong_comb<double, double><<<nblk(L0.nunit), TPB, 0, s>>>(L0.nunit, L0.d_sa_poff, L0.d_sa_pw, L0.d_sa_eagg,
                                                                   L0.d_sa_Pe, sc, ia, ib, xa, xb, x, ndes);
}
#define UFT_SA_NR(nr, CALL)                        \
  switch (nr) {                                    \
    case 2: { constexpr int NR = 2; CALL; } break; \
    case 4: { constexpr int NR = 4; CALL; } break; \
    default: { constexpr int NR = 8; CALL; } break; \
  }
void mr_sa_restrict_first(hipStream_t s, int nr, const AmgLevel &L0, int64_t nc, const double *r, double *bc,
                          const double *dinv, double it, double *rc, double *dc, double *xc) {
  UFT_SA_NR(nr, (k_mr_sa_restrict_first<16, double, double, NR><<<nblk(nc * 16), TPB, 0, s>>>(
                    nc, L0.d_sa_sptr, L0.d_sa_snode, L0.d_sa_Ps, r, bc, dinv, it, rc, dc, xc)));
}
void mr_sa_restrict_first(hipStream_t s, int nr, const AmgLevel &L0, int64_t nc, const float *r, float *bc,
                          const float *dinv, double it, float *rc, float *dc, float *xc) {
  if (L0.d_sa_Psh) {
    UFT_SA_NR(nr, (k_mr_sa_restrict_first<16, float, __half, NR><<<nblk(nc * 16), TPB, 0, s>>>(
                      nc, L0.d_sa_sptr, L0.d_sa_snode, (const __half *)L0.d_sa_Psh, r, bc, dinv, it, rc, dc, xc)));
  } else {
    UFT_SA_NR(nr, (k_mr_sa_restrict_first<16, float, float, NR><<<nblk(nc * 16), TPB, 0, s>>>(
                      nc, L0.d_sa_sptr, L0.d_sa_snode, L0.d_sa_Psf, r, bc, dinv, it, rc, dc, xc)));
  }
}
void mr_sa_prolong_comb(hipStream_t s, int nr, const AmgLevel &L0, int64_t ncu, const double *sc, int ia, int ib,
                        const double *xa, const double *xb, double *x, double *tmp) {
  if (xb) {  // combine first (tmp: a free coarse vector of nr right-hand sides)
    const int64_t nc = 6 * ncu;
    UFT_SA_NR(nr, (k_mr_lincomb2<double, NR><<<nblk(nc * NR), TPB, 0, s>>>(nc, sc, ia, ib, xa, xb, tmp)));
    xa = tmp;
    xb = nullptr;
    ia = ib = -1;
  }
  UFT_SA_NR(nr, (k_mr_sa_prolong_comb<double, double, NR><<<nblk(L0.nunit), TPB, 0, s>>>(
                    L0.nunit, L0.sa_ne_ell, L0.d_sa_poff, L0.d_sa_pw, L0.d_sa_eagg, L0.d_sa_Pe, sc, ia, xa, x)));
}
void mr_sa_prolong_comb(hipStream_t s, int nr, const AmgLevel &L0, int64_t ncu, const double *sc, int ia, int ib,
                        const float *xa, const float *xb, float *x, float *tmp) {
  if (xb) {
    const int64_t nc = 6 * ncu;
    UFT_SA_NR(nr, (k_mr_lincomb2<float, NR><<<nblk(nc * NR), TPB, 0, s>>>(nc, sc, ia, ib, xa, xb, tmp)));
    xa = tmp;
    xb = nullptr;
    ia = ib = -1;
  }
  if (L0.d_sa_Peh) {
    UFT_SA_NR(nr, (k_mr_sa_prolong_comb<float, __half, NR><<<nblk(L0.nunit), TPB, 0, s>>>(
                      L0.nunit, L0.sa_ne_ell, L0.d_sa_poff, L0.d_sa_pw, L0.d_sa_eagg, (const __half *)L0.d_sa_Peh, sc,
                      ia, xa, x)));
  } else {
    UFT_SA_NR(nr, (k_mr_sa_prolong_comb<float, float, NR><<<nblk(L0.nunit), TPB, 0, s>>>(
                      L0.nunit, L0.sa_ne_ell, L0.d_sa_poff, L0.d_sa_pw, L0.d_sa_eagg, L0.d_sa_Pef, sc, ia, xa, x)));
  }
}
#undef UFT_SA_NR

void sa_prolong_comb(hipStream_t s, const AmgLevel &L0, const double *sc, int ia, int ib, const float *xa,
                     const float *xb, float *x, const int32_t *ndes) {
  if (L0.d_sa_Peh) {
    if (xb)
      k_sa_prolong_comb_h<true><<<nblk(L0.nunit), TPB, 0, s>>>(
          L0.nunit, L0.sa_ne_ell, L0.d_sa_poff, L0.d_sa_pw, L0.d_sa_eagg, (const uint4 *)L0.d_sa_Peh,
          (const uint32_t *)L0.d_sa_Peh + 8 * L0.sa_ne_ell, sc, ia, ib, xa, xb, x, ndes);
    else
      k_sa_prolong_comb_h<false><<<nblk(L0.nunit), TPB, 0, s>>>(
          L0.nunit, L0.sa_ne_ell, L0.d_sa_poff, L0.d_sa_pw, L0.d_sa_eagg, (const uint4 *)L0.d_sa_Peh,
          (const uint32_t *)L0.d_sa_Peh + 8 * L0.sa_ne_ell, sc, ia, ib, xa, xb, x, ndes);
  } else {
    k_sa_prolong_comb<float, float><<<nblk(L0.nunit), TPB, 0, s>>>(L0.nunit, L0.d_sa_poff, L0.d_sa_pw, L0.d_sa_eagg,
                                                                   L0.d_sa_Pef, sc, ia, ib, xa, xb, x, ndes);
  }
}

}  // namespace uftk
