// Deterministic block reductions for the gfx950 kernels.
//
// Every block writes its K partial sums in a fixed intra-block order (64-lane
// shuffle tree, then the four waves in order); a one-block final kernel
// (uftk::final_sum) then sums the partials in a fixed order.  Results do not
// depend on scheduling, so reductions are bit-reproducible run to run.
//
// A single-launch variant (last block finalises after a device-scope fence and
// a ticket atomic) was measured slower on MI355X: the device-scope release per
// block has to make every block's partials visible across the eight XCD L2
// caches, which costs more than the second launch it saves.
#pragma once
#include <hip/hip_runtime.h>

#include "uft_internal.h"

namespace uftr {

constexpr int RTPB = 256;

// Workgroup b of a launch runs on XCD b mod 8.  xcd_block maps it to a
// logical block so that each XCD takes one contiguous range of the nb
// logical blocks (a bijection): kernels whose neighbouring rows share
// gathered operands then keep those operands in one XCD's L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t x = b % 8, q = nb / 8, r = nb % 8;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + b / 8;
}

// Row of this thread under a segment grid (blocks of RTPB threads): the
// plain grid (row = block * RTPB + thread, design 0, valid below n) or the
// block's own row range and design.
__device__ __forceinline__ bool seg_row(const SegG &g, int64_t n, int64_t &i, int &d) {
  if (g.span) {
    const int64_t lo = g.span[2 * blockIdx.x], hi = g.span[2 * blockIdx.x + 1];
    i = lo + threadIdx.x;
    d = g.des[blockIdx.x];
    return i < hi;
  }
  i = (int64_t)blockIdx.x * RTPB + threadIdx.x;
  d = 0;
  return i < n;
}
inline unsigned seg_grid(const SegG &g, int64_t n) {
  return g.span ? (unsigned)g.nb : (unsigned)((n + RTPB - 1) / RTPB);
}

template <int K>
__device__ __forceinline__ void block_partials(double (&v)[K], double *part) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  }
  __shared__ double sm[K][RTPB / 64];
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

// One block (RTPB threads) sums the K columns of nb partial rows in a fixed
// order: strided per-thread sums, a 64-lane shuffle tree, then the four waves
// in order.  Every thread receives the K results.
template <int K>
__device__ __forceinline__ void block_final(const double *__restrict__ part, int nb, double (&res)[K]) {
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (int i = threadIdx.x; i < nb; i += RTPB) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += part[(size_t)i * K + k];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  }
  __shared__ double sm[K][RTPB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sm[k][w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) res[k] = (sm[k][0] + sm[k][1]) + (sm[k][2] + sm[k][3]);
}

}  // namespace uftr
