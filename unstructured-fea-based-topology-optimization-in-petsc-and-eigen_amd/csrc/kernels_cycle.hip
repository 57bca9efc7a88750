// gfx950 kernels of the fused multigrid cycle.  Every operator apply of the
// cycle (the matrix-free fine operator's node phase, and the 6x6 block-sparse
// coarse operators) finishes with a per-row epilogue that performs the vector
// update the cycle needs next (Chebyshev step, residual, first smoothing step,
// K-cycle inner products), so a smoothing step is one launch instead of two.
// Restriction also performs the first Chebyshev step of the next level,
// prolongation forms the K-cycle's linear combination on the fly, and the
// last restriction and the dense coarsest solve share one launch.  Row-local epilogues never write the vector that the
// same launch multiplies, and every reduction keeps a fixed order.
#include <hip/hip_fp16.h>

#include "uft_elem.h"
#include "uft_internal.h"
#include "uft_reduce.h"

namespace {
using namespace uftd;

constexpr int TPB = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// Per-row epilogue of an operator apply, t = (A u)_i, the mode (RowOp) fixed
// at compile time: the operands that do not depend on the product (EpiPre,
// u_i included) are loaded when the kernel starts, so their latency overlaps
// the product's gathers instead of following them (the coarse levels'
// launches are latency bound).
template <typename T, int MODE>
struct EpiPre {
  T ui = T(0), a = T(0), b = T(0), c = T(0);
  __device__ __forceinline__ void load(const RowOp<T> &op, const T *u, int64_t i) {
    ui = u[i];
    if constexpr (MODE == ROW_STEP) {
      a = op.r[i];
      b = op.dinv[i];
      c = op.x[i];
    } else if constexpr (MODE == ROW_RES) {
      a = op.base[i];
    } else if constexpr (MODE == ROW_RESFIRST || MODE == ROW_RESFIRSTX) {
      a = op.base[i];
      b = op.dinv[i];
    } else if constexpr (MODE == ROW_DOTS2) {
      a = op.w1[i];
    } else if constexpr (MODE == ROW_DOTS3) {
      a = op.w1[i];
      b = op.w2[i];
    }
  }
};
template <typename T, int MODE>
__device__ __forceinline__ void row_epi_m(const RowOp<T> &op, int64_t i, T t, const EpiPre<T, MODE> &p,
                                          double (&dv)[3]) {
  const T ui = p.ui;
  if constexpr (MODE == ROW_STORE) {
    op.y[i] = t;
  } else if constexpr (MODE == ROW_STEP) {
    const T ri = p.a - t;
    op.r[i] = ri;
    const T dn = (T)op.c1 * ui + (T)op.c2 * p.b * ri;
    op.dout[i] = dn;
    op.x[i] = p.c + (op.addprev ? dn + ui : dn);
  } else if constexpr (MODE == ROW_RES) {
    op.r[i] = p.a - t;
  } else if constexpr (MODE == ROW_RESFIRST) {
    const T ri = p.a - t;
    op.r[i] = ri;
    op.dout[i] = (T)op.it * p.b * ri;
  } else if constexpr (MODE == ROW_RESFIRSTX) {
    const T ri = p.a - t;
    op.r[i] = ri;
    op.kr[i] = ri;
    const T dn = (T)op.it * p.b * ri;
    op.dout[i] = dn;
    op.kd[i] = dn;
  } else if constexpr (MODE == ROW_DOTS2) {
    op.y[i] = t;
    dv[0] += (double)ui * (double)t;
    dv[1] += (double)ui * (double)p.a;
  } else if constexpr (MODE == ROW_DOTS3) {
    dv[0] += (double)ui * (double)p.a;
    dv[1] += (double)ui * (double)t;
    dv[2] += (double)ui * (double)p.b;
  } else {  // ROW_STOREDOT
    op.y[i] = t;
    dv[0] += (double)ui * (double)t;
  }
}
template <typename T, int MODE>
__device__ __forceinline__ void epi_partials_m(const RowOp<T> &op, double (&dv)[3]) {
  if constexpr (MODE == ROW_DOTS2) {
    double v2[2] = {dv[0], dv[1]};
    uftr::block_partials<2>(v2, op.part);
  } else if constexpr (MODE == ROW_DOTS3) {
    uftr::block_partials<3>(dv, op.part);
  } else if constexpr (MODE == ROW_STOREDOT) {
    double v1[1] = {dv[0]};
    uftr::block_partials<1>(v1, op.part);
  }
}

// ---------------------------------------------------------------- 6x6 BSR
// LPR lanes per scalar row, fixed shuffle-tree combine: 8 on large levels,
// 32 on small ones so that the launch still spreads over the chip.
inline int lpr_for(int64_t nb_rows) { return nb_rows * 6 < 65536 ? 32 : 8; }
template <typename T, int LPR, int MODE>
__global__ __launch_bounds__(TPB) void k_bsr6_op(int64_t nrows, const int32_t *__restrict__ rp,
                                                 const int32_t *__restrict__ ci, const T *__restrict__ val,
                                                 const T *__restrict__ u, RowOp<T> op) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int64_t row = t / LPR;
  const int l = (int)(t - row * LPR);
  EpiPre<T, MODE> pre;
  if (l == 0 && row < nrows * 6) pre.load(op, u, row);
  T acc = T(0);
  if (row < nrows * 6) {
    const int64_t A = row / 6;
    const int i = (int)(row - A * 6);
    for (int k = rp[A] + l; k < rp[A + 1]; k += LPR) {
      const int64_t B = ci[k];
      const T *v = val + (size_t)k * 36 + i * 6;
      const T *xb = u + B * 6;
#pragma unroll
      for (int j = 0; j < 6; ++j) acc += v[j] * xb[j];
    }
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, LPR);
  double dv[3] = {0.0, 0.0, 0.0};
  if (l == 0 && row < nrows * 6) row_epi_m<T, MODE>(op, row, acc, pre, dv);
  epi_partials_m<T, MODE>(op, dv);
}

// FP16 blocks of the symmetrically scaled operator D^-1/2 A D^-1/2 (every
// entry in [-1, 1]; dh = D^1/2 per scalar row, FP32):
// t_i = dh_i sum_j Ah_ij (dh_j u_j), FP32 arithmetic.  LPR lanes per block
// row; a lane reads whole 72-byte blocks (nine 8-byte loads) and the six
// scaled entries of u once per block, accumulating all six rows; a fixed
// butterfly combines the lanes and six lanes finish the six rows.
inline int lpr_h(int64_t nb_rows) { return nb_rows < 4096 ? 32 : 16; }
template <int LPR, int MODE>
__global__ __launch_bounds__(TPB) void k_bsr6_op_h(int64_t nbrows, const int32_t *__restrict__ rp,
                                                   const int32_t *__restrict__ ci, const uint2 *__restrict__ valh,
                                                   const float *__restrict__ dh, const float *__restrict__ u,
                                                   RowOp<float> op) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int64_t A = t / LPR;
  const int l = (int)(t - A * LPR);
  EpiPre<float, MODE> pre;
  float dhr = 0.f;
  if (l < 6 && A < nbrows) {
    pre.load(op, u, A * 6 + l);
    dhr = dh[A * 6 + l];
  }
  float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (A < nbrows) {
    for (int k = rp[A] + l; k < rp[A + 1]; k += LPR) {
      const int64_t B = ci[k];
      const float2 *ub = reinterpret_cast<const float2 *>(u + B * 6);
      const float2 *db = reinterpret_cast<const float2 *>(dh + B * 6);
      float xs[6];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float2 a = ub[j], b = db[j];
        xs[2 * j] = a.x * b.x;
        xs[2 * j + 1] = a.y * b.y;
      }
      const uint2 *vb = valh + (size_t)k * 9;
#pragma unroll
      for (int w = 0; w < 9; ++w) {
        const uint2 h = vb[w];
        const unsigned hw[2] = {h.x, h.y};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const float2 f = __half22float2(*reinterpret_cast<const __half2 *>(&hw[q]));
          const int e = 4 * w + 2 * q;  // entries e, e+1 of the row-major block (same row: 6 is even)
          acc[e / 6] += f.x * xs[e % 6] + f.y * xs[e % 6 + 1];
        }
      }
    }
  }
  // butterfly: every lane of the group ends with the six sums (each lane in
  // its own fixed order); lane i < 6 finishes row 6A + i
#pragma unroll
  for (int i = 0; i < 6; ++i) {
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) acc[i] += __shfl_xor(acc[i], off, LPR);
  }
  double dv[3] = {0.0, 0.0, 0.0};
  if (l < 6 && A < nbrows) {
    float ti = acc[0];
#pragma unroll
    for (int i = 1; i < 6; ++i)
      if (l == i) ti = acc[i];
    row_epi_m<float, MODE>(op, A * 6 + l, ti * dhr, pre, dv);
  }
  epi_partials_m<float, MODE>(op, dv);
}

// ---------------------------------------------------------------- fine node phase
// y_n = sum of the element-local results of node n (fixed incidence order);
// fixed DOFs have the unit diagonal row (y = u).
template <int NPE, typename T, int MODE>
__global__ __launch_bounds__(TPB) void k_node_gather_op(int nn, int ne, const int32_t *__restrict__ iptr,
                                                        const int32_t *__restrict__ inc, const T *__restrict__ Ye,
                                                        const uint8_t *__restrict__ fixm, const T *__restrict__ u,
                                                        RowOp<T> op) {
  const int n = blockIdx.x * TPB + threadIdx.x;
  double dv[3] = {0.0, 0.0, 0.0};
  if (n < nn) {
    EpiPre<T, MODE> pre[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) pre[c].load(op, u, 3 * (int64_t)n + c);
    T acc[3] = {T(0), T(0), T(0)};
    const int k0 = iptr[n], k1 = iptr[n + 1];
    for (int k = k0; k < k1; ++k) {
      const int v = inc[k];
      const int e = v / NPE, l = v - e * NPE;
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[c] += Ye[(size_t)(c * NPE + l) * ne + e];
    }
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      row_epi_m<T, MODE>(op, 3 * (int64_t)n + c, ((m >> c) & 1u) ? pre[c].ui : acc[c], pre[c], dv);
  }
  epi_partials_m<T, MODE>(op, dv);
}

// ---------------------------------------------------------------- assembled fine operator
// Sliced ELL of 3x3 node blocks (FP16, each scalar row divided by its
// diagonal dg_i): one node per lane and one slice per wavefront, so the slot
// loop is uniform and every slot load of a wavefront reads consecutive
// memory.  t_i = dg_i sum_s Ah_(n,s) u_col(n,s) in FP32; fixed DOFs have the
// unit row (t = u), their columns are zero in the blocks.  Bytes per node:
// slots x (16 + 2 + 2) (16-bit column offsets; 4 when some neighbour is more
// than 32767 nodes away) plus the row scales and the epilogue's vectors.
__device__ __forceinline__ float2 h2f2(unsigned w) { return __half22float2(*reinterpret_cast<const __half2 *>(&w)); }
template <bool NT, int MODE, typename CI>
__global__ __launch_bounds__(TPB) void k_fine_ell_op(int nn, const int64_t *__restrict__ off,
                                                     const int32_t *__restrict__ width, const CI *__restrict__ col,
                                                     const uint4 *__restrict__ v8, const uint16_t *__restrict__ v1,
                                                     const float *__restrict__ dg, const uint8_t *__restrict__ fixm,
                                                     const float *__restrict__ u, RowOp<float> op) {
  const int n = blockIdx.x * TPB + threadIdx.x;
  double dv[3] = {0.0, 0.0, 0.0};
  if (n < nn) {
    EpiPre<float, MODE> pre[3];
    float dgr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pre[c].load(op, u, 3 * (int64_t)n + c);
      dgr[c] = dg[3 * (int64_t)n + c];
    }
    const int w = n / UFT_ELL_SLICE;
    const int64_t base = off[w] + n % UFT_ELL_SLICE;
    const int W = width[w];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll 3
    for (int s = 0; s < W; ++s) {
      const int64_t k = base + (int64_t)s * UFT_ELL_SLICE;
      const int64_t j = std::is_same<CI, int16_t>::value ? n + (int64_t)col[k] : (int64_t)col[k];
      unsigned h[4];
      if constexpr (NT) {
        typedef unsigned v4u __attribute__((ext_vector_type(4)));
        const v4u v = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(&v8[k]));
        h[0] = v.x; h[1] = v.y; h[2] = v.z; h[3] = v.w;
      } else {
        const uint4 v = v8[k];
        h[0] = v.x; h[1] = v.y; h[2] = v.z; h[3] = v.w;
      }
      const float e8 = __half2float(__ushort_as_half(v1[k]));
      const float x0 = u[3 * j], x1 = u[3 * j + 1], x2 = u[3 * j + 2];
      const float2 e01 = h2f2(h[0]), e23 = h2f2(h[1]), e45 = h2f2(h[2]), e67 = h2f2(h[3]);
      a0 += e01.x * x0 + e01.y * x1 + e23.x * x2;
      a1 += e23.y * x0 + e45.x * x1 + e45.y * x2;
      a2 += e67.x * x0 + e67.y * x1 + e8 * x2;
    }
    const float acc[3] = {a0, a1, a2};
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      row_epi_m<float, MODE>(op, 3 * (int64_t)n + c, ((m >> c) & 1u) ? pre[c].ui : dgr[c] * acc[c], pre[c], dv);
  }
  epi_partials_m<float, MODE>(op, dv);
}

// The same smoother operator with int8 blocks: per slot eight entries in a
// uint2 and the block's FP32 scale with entry 8 in its low byte (12 bytes
// instead of 18): t_i = dg_i sum_s sc_s (q_s u_col(n,s))_i.  The int8 grid
// of one block spans +-127 of its largest entry, so an entry's rounding is
// at most 1/254 of the block's largest (FP16: 2^-11 of the entry itself).
__device__ __forceinline__ float sb8(unsigned w, int sh) { return (float)((int)(w << (24 - sh)) >> 24); }
template <int MODE, typename CI>
__global__ __launch_bounds__(TPB) void k_fine_ell_op_q(int nn, const int64_t *__restrict__ off,
                                                       const int32_t *__restrict__ width, const CI *__restrict__ col,
                                                       const uint2 *__restrict__ q8, const uint32_t *__restrict__ qs,
                                                       const float *__restrict__ dg, const uint8_t *__restrict__ fixm,
                                                       const float *__restrict__ u, RowOp<float> op) {
  const int n = blockIdx.x * TPB + threadIdx.x;
  double dv[3] = {0.0, 0.0, 0.0};
  if (n < nn) {
    EpiPre<float, MODE> pre[3];
    float dgr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pre[c].load(op, u, 3 * (int64_t)n + c);
      dgr[c] = dg[3 * (int64_t)n + c];
    }
    const int w = n / UFT_ELL_SLICE;
    const int64_t base = off[w] + n % UFT_ELL_SLICE;
    const int W = width[w];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll 3
    for (int s = 0; s < W; ++s) {
      const int64_t k = base + (int64_t)s * UFT_ELL_SLICE;
      const int64_t j = std::is_same<CI, int16_t>::value ? n + (int64_t)col[k] : (int64_t)col[k];
      typedef unsigned v2u __attribute__((ext_vector_type(2)));
      const v2u v = __builtin_nontemporal_load(reinterpret_cast<const v2u *>(&q8[k]));
      const unsigned xs = __builtin_nontemporal_load(&qs[k]);
      const float sc = __uint_as_float(xs & 0xFFFFFF00u);
      const float x0 = u[3 * j], x1 = u[3 * j + 1], x2 = u[3 * j + 2];
      const float r0 = sb8(v.x, 0) * x0 + sb8(v.x, 8) * x1 + sb8(v.x, 16) * x2;
      const float r1 = sb8(v.x, 24) * x0 + sb8(v.y, 0) * x1 + sb8(v.y, 8) * x2;
      const float r2 = sb8(v.y, 16) * x0 + sb8(v.y, 24) * x1 + sb8(xs, 0) * x2;
      a0 += sc * r0;
      a1 += sc * r1;
      a2 += sc * r2;
    }
    const float acc[3] = {a0, a1, a2};
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      row_epi_m<float, MODE>(op, 3 * (int64_t)n + c, ((m >> c) & 1u) ? pre[c].ui : dgr[c] * acc[c], pre[c], dv);
  }
  epi_partials_m<float, MODE>(op, dv);
}

// The same pattern with FP32 entries (no row scale): 40 bytes per slot.
template <int MODE, typename CI>
__global__ __launch_bounds__(TPB) void k_fine_ell_op32(int nn, const int64_t *__restrict__ off,
                                                       const int32_t *__restrict__ width,
                                                       const CI *__restrict__ col, const float4 *__restrict__ w8,
                                                       const float *__restrict__ w1, const uint8_t *__restrict__ fixm,
                                                       const float *__restrict__ u, RowOp<float> op) {
  const int n = blockIdx.x * TPB + threadIdx.x;
  double dv[3] = {0.0, 0.0, 0.0};
  if (n < nn) {
    EpiPre<float, MODE> pre[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) pre[c].load(op, u, 3 * (int64_t)n + c);
    const int w = n / UFT_ELL_SLICE;
    const int64_t base = off[w] + n % UFT_ELL_SLICE;
    const int W = width[w];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll 3
    for (int s = 0; s < W; ++s) {
      const int64_t k = base + (int64_t)s * UFT_ELL_SLICE;
      const int64_t j = std::is_same<CI, int16_t>::value ? n + (int64_t)col[k] : (int64_t)col[k];
      typedef float v4f __attribute__((ext_vector_type(4)));
      const v4f e0 = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&w8[2 * k]));
      const v4f e1 = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&w8[2 * k + 1]));
      const float e8 = __builtin_nontemporal_load(&w1[k]);
      const float x0 = u[3 * j], x1 = u[3 * j + 1], x2 = u[3 * j + 2];
      a0 += e0.x * x0 + e0.y * x1 + e0.z * x2;
      a1 += e0.w * x0 + e1.x * x1 + e1.y * x2;
      a2 += e1.z * x0 + e1.w * x1 + e8 * x2;
    }
    const float acc[3] = {a0, a1, a2};
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      row_epi_m<float, MODE>(op, 3 * (int64_t)n + c, ((m >> c) & 1u) ? pre[c].ui : acc[c], pre[c], dv);
  }
  epi_partials_m<float, MODE>(op, dv);
}

// The FP64 form's operator on the same pattern: the FP64 blocks assembled at
// the refresh from the FP64 element matrices (72 bytes per slot: entries
// 2q, 2q+1 in plane q at e8[q ns + slot], entry 8 in e1, so every load of a
// wavefront reads consecutive memory), q = K u in FP64 with the fused row
// epilogues; one launch per apply instead of the element phase, its
// element-local results and the node phase.
template <int MODE, typename CI>
__global__ __launch_bounds__(TPB) void k_fine_ell_op64(int nn, int64_t ns, const int64_t *__restrict__ off,
                                                       const int32_t *__restrict__ width,
                                                       const CI *__restrict__ col, const double2 *__restrict__ e8,
                                                       const double *__restrict__ e1, const uint8_t *__restrict__ fixm,
                                                       const double *__restrict__ u, RowOp<double> op) {
  const int n = blockIdx.x * TPB + threadIdx.x;
  double dv[3] = {0.0, 0.0, 0.0};
  if (n < nn) {
    EpiPre<double, MODE> pre[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) pre[c].load(op, u, 3 * (int64_t)n + c);
    const int w = n / UFT_ELL_SLICE;
    const int64_t base = off[w] + n % UFT_ELL_SLICE;
    const int W = width[w];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll 3
    for (int s = 0; s < W; ++s) {
      const int64_t k = base + (int64_t)s * UFT_ELL_SLICE;
      const int64_t j = std::is_same<CI, int16_t>::value ? n + (int64_t)col[k] : (int64_t)col[k];
      typedef double v2d __attribute__((ext_vector_type(2)));
      const v2d b0 = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(&e8[k]));
      const v2d b1 = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(&e8[ns + k]));
      const v2d b2 = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(&e8[2 * ns + k]));
      const v2d b3 = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(&e8[3 * ns + k]));
      const double b8 = __builtin_nontemporal_load(&e1[k]);
      const double x0 = u[3 * j], x1 = u[3 * j + 1], x2 = u[3 * j + 2];
      a0 += b0.x * x0 + b0.y * x1 + b1.x * x2;
      a1 += b1.y * x0 + b2.x * x1 + b2.y * x2;
      a2 += b3.x * x0 + b3.y * x1 + b8 * x2;
    }
    const double acc[3] = {a0, a1, a2};
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      row_epi_m<double, MODE>(op, 3 * (int64_t)n + c, ((m >> c) & 1u) ? pre[c].ui : acc[c], pre[c], dv);
  }
  epi_partials_m<double, MODE>(op, dv);
}

// Assembly at a hierarchy refresh: lane (n, slot s) sums E(x_e) KE_e[a][b]
// (3x3) over the slot's element contributions (incidence order, FP64),
// zeroes fixed columns, divides row c by the diagonal 1 / dinv_3n+c and
// stores FP16, or int8 with a block scale when q8 is given (and, with w8, the
// unscaled entries in FP32).  The element blocks are read from their FP32 copy and the
// moduli E(x_e) from the refresh's per-element array.
template <int NPE, typename KT>
__global__ __launch_bounds__(TPB) void k_fine_ell_fill(int nn, int ne, const int64_t *__restrict__ off,
                                                       const int32_t *__restrict__ width,
                                                       const int32_t *__restrict__ col,
                                                       const int32_t *__restrict__ cptr,
                                                       const int32_t *__restrict__ clist,
                                                       const KT *__restrict__ KEf,
                                                       const double *__restrict__ Ee,
                                                       const uint8_t *__restrict__ fixm,
                                                       const double *__restrict__ dinv, uint4 *__restrict__ v8,
                                                       uint16_t *__restrict__ v1, uint2 *__restrict__ q8,
                                                       uint32_t *__restrict__ qs, float *__restrict__ dg,
                                                       float4 *__restrict__ w8, float *__restrict__ w1,
                                                       double *__restrict__ a64, float *__restrict__ a32,
                                                       const int32_t *__restrict__ kmap, uint4 *__restrict__ tv8,
                                                       uint16_t *__restrict__ tv1, float4 *__restrict__ tw8,
                                                       float *__restrict__ tw1, double2 *__restrict__ e8,
                                                       double *__restrict__ e1, int64_t ns) {
  constexpr int ND = 3 * NPE;
  const int n = blockIdx.x * TPB + threadIdx.x;
  const int s = blockIdx.y;
  if (n >= nn) return;
  const int w = n / UFT_ELL_SLICE;
  if (s >= width[w]) return;
  const int64_t k = off[w] + (int64_t)s * UFT_ELL_SLICE + n % UFT_ELL_SLICE;
  double b[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) b[t] = 0.0;
  for (int q = cptr[k]; q < cptr[k + 1]; ++q) {
    const int code = clist[q];
    const int lj = code % NPE, v = code / NPE;
    const int e = v / NPE, li = v - e * NPE;
    const double E = Ee[e];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int p = pk_index(c * NPE + li, d * NPE + lj, ND);
        const KT pr = KEf[(size_t)(p >> 1) * ne + e];
        b[3 * c + d] += E * (double)((p & 1) ? pr.y : pr.x);
      }
  }
  const unsigned mj = fixm[col[k]];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if ((mj >> d) & 1u) b[3 * c + d] = 0.0;
  if (a64) {
#pragma unroll
    for (int t = 0; t < 9; ++t) a64[9 * k + t] = b[t];
  }
  if (a32) {
#pragma unroll
    for (int t = 0; t < 9; ++t) a32[9 * k + t] = (float)b[t];
  }
  if (e8) {
#pragma unroll
    for (int q = 0; q < 4; ++q) e8[q * ns + k] = make_double2(b[2 * q], b[2 * q + 1]);
    e1[k] = b[8];
  }
  if (!q8 && !v8 && !w8) return;  // the smoothed prolongator's blocks only
  // the node-tile copies (several load cases) at the slot's tiled position
  const int64_t k2 = kmap ? (int64_t)kmap[k] : -1;
  if (w8) {
    const float4 f0 = make_float4((float)b[0], (float)b[1], (float)b[2], (float)b[3]);
    const float4 f1 = make_float4((float)b[4], (float)b[5], (float)b[6], (float)b[7]);
    w8[2 * k] = f0;
    w8[2 * k + 1] = f1;
    w1[k] = (float)b[8];
    if (k2 >= 0) {
      tw8[2 * k2] = f0;
      tw8[2 * k2 + 1] = f1;
      tw1[k2] = (float)b[8];
    }
  }
  if (!q8 && !v8) return;  // UFT_MIXED: the FP32 blocks only
  double a[9];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double di = dinv[3 * (size_t)n + c];
    const double dd = (di > 0.0 && di < 1e300) ? 1.0 / di : 1.0;
    if (s == 0) dg[3 * (size_t)n + c] = (float)dd;
#pragma unroll
    for (int d = 0; d < 3; ++d) a[3 * c + d] = b[3 * c + d] / dd;
  }
  if (q8) {
    // int8 with one scale per block: sc = max|a| / 127 in FP32 with its low
    // mantissa byte cleared (it carries entry 8), q = rint(a / sc)
    double mx = 0.0;
#pragma unroll
    for (int t = 0; t < 9; ++t) mx = fmax(mx, fabs(a[t]));
    const unsigned sbits = __float_as_uint((float)(mx / 127.0)) & 0xFFFFFF00u;
    const double sc = (double)__uint_as_float(sbits);
    unsigned qb[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const double q = sc > 0.0 ? fmin(fmax(rint(a[t] / sc), -127.0), 127.0) : 0.0;
      qb[t] = (unsigned)(int)q & 0xFFu;
    }
    q8[k] = make_uint2(qb[0] | (qb[1] << 8) | (qb[2] << 16) | (qb[3] << 24),
                       qb[4] | (qb[5] << 8) | (qb[6] << 16) | (qb[7] << 24));
    qs[k] = (sc > 0.0 ? sbits : 0u) | qb[8];
    return;
  }
  unsigned short hb[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) hb[t] = __half_as_ushort(__float2half_rn((float)a[t]));
  const uint4 hv = make_uint4((unsigned)hb[0] | ((unsigned)hb[1] << 16), (unsigned)hb[2] | ((unsigned)hb[3] << 16),
                              (unsigned)hb[4] | ((unsigned)hb[5] << 16), (unsigned)hb[6] | ((unsigned)hb[7] << 16));
  v8[k] = hv;
  v1[k] = hb[8];
  if (k2 >= 0) {
    tv8[k2] = hv;
    tv1[k2] = hb[8];
  }
}

// The filter's Helmholtz operator on the same node pattern (FP64 scalar
// entries, assembled once at setup: KF does not depend on the density):
// y_n = sum_s H_(n,s) x_col(n,s); with part, the block partials of x.y; nothing
// once *stop is set (the filter CG's device-side convergence flag)
template <int NPE>
__global__ __launch_bounds__(TPB) void k_helm_ell_fill(int nn, int ne, const int64_t *__restrict__ off,
                                                       const int32_t *__restrict__ width,
                                                       const int32_t *__restrict__ cptr,
                                                       const int32_t *__restrict__ clist,
                                                       const double2 *__restrict__ KF, double *__restrict__ hv) {
  const int n = blockIdx.x * TPB + threadIdx.x;
  const int s = blockIdx.y;
  if (n >= nn) return;
  const int w = n / UFT_ELL_SLICE;
  if (s >= width[w]) return;
  const int64_t k = off[w] + (int64_t)s * UFT_ELL_SLICE + n % UFT_ELL_SLICE;
  double h = 0.0;
  for (int q = cptr[k]; q < cptr[k + 1]; ++q) {
    const int code = clist[q];
    const int lj = code % NPE, v = code / NPE;
    const int e = v / NPE, li = v - e * NPE;
    const int p = pk_index(li, lj, NPE);
    const double2 pr = KF[(size_t)(p >> 1) * ne + e];
    h += (p & 1) ? pr.y : pr.x;
  }
  hv[k] = h;
}
template <typename CI>
__global__ __launch_bounds__(TPB) void k_helm_ell(int nn, const int64_t *__restrict__ off,
                                                  const int32_t *__restrict__ width, const CI *__restrict__ col,
                                                  const double *__restrict__ hv, const double *__restrict__ x,
                                                  double *__restrict__ y, double *__restrict__ part,
                                                  const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // the same for every block
  const int n = blockIdx.x * TPB + threadIdx.x;
  double ds[1] = {0.0};
  if (n < nn) {
    const int w = n / UFT_ELL_SLICE;
    const int64_t base = off[w] + n % UFT_ELL_SLICE;
    const int W = width[w];
    double acc = 0.0;
#pragma unroll 3
    for (int s = 0; s < W; ++s) {
      const int64_t k = base + (int64_t)s * UFT_ELL_SLICE;
      const int64_t j = std::is_same<CI, int16_t>::value ? n + (int64_t)col[k] : (int64_t)col[k];
      acc += __builtin_nontemporal_load(&hv[k]) * x[j];
    }
    y[n] = acc;
    ds[0] = acc * x[n];
  }
  if (part) uftr::block_partials<1>(ds, part);
}

// ---------------------------------------------------------------- transfers
// G lanes per coarse unit A (64 / G units per wavefront): bc_A = sum_children
// P_u^T r_u (each lane a fixed stride of the unit's (child, row) pairs, then
// a fixed butterfly within the G lanes); with dinv given, also the next
// level's first Chebyshev step from a zero guess: rc = bc, dc = it D^-1 bc,
// xc = dc.
template <int BS, int G, typename T>
__global__ __launch_bounds__(TPB) void k_restrict_first(int64_t nc, const int32_t *__restrict__ cptr,
                                                        const int32_t *__restrict__ child, const T *__restrict__ P,
                                                        const T *__restrict__ r, T *__restrict__ bc,
                                                        const T *__restrict__ dinv, double it, T *__restrict__ rc,
                                                        T *__restrict__ dc, T *__restrict__ xc) {
  static_assert(G >= 1 && G <= 64 && (64 % G) == 0, "G lanes must tile a wavefront");
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int64_t A = t / G;
  const int lane = (int)(t % G);
  const bool on = A < nc;
  const int k0 = on ? cptr[A] : 0, npair = on ? (cptr[A + 1] - k0) * BS : 0;
  T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int q = lane; q < npair; q += G) {
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
    for (int off = G / 2; off > 0; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
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

// x_u += P_u (ca xa + cb xb)_{parent(u)}: ca, cb from the device scalars
// sc[ia], sc[ib] (index < 0: 1.0); xb may be null.
template <int BS, typename T>
__global__ __launch_bounds__(TPB) void k_prolong_comb(int64_t nu, const int32_t *__restrict__ parent,
                                                      const T *__restrict__ P, const double *__restrict__ sc, int ia,
                                                      int ib, const T *__restrict__ xa, const T *__restrict__ xb,
                                                      T *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= nu * BS) return;
  const int64_t u = t / BS;
  const int i = (int)(t - u * BS);
  const int64_t A = parent[u];
  if (A < 0) return;
  const T ca = ia < 0 ? T(1) : (T)sc[ia];
  const T cb = ib < 0 ? T(1) : (T)sc[ib];
  const T *Pu = P + (size_t)u * BS * 6 + i * 6;
  T acc = T(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const T c = xb ? ca * xa[A * 6 + j] + cb * xb[A * 6 + j] : xa[A * 6 + j];
    acc += Pu[j] * c;
  }
  x[t] += acc;
}

// The same for 3 rows per unit (the fine level) with one thread per unit:
// the parent's combination is formed once for the unit's three rows, whose
// 3x6 block is one contiguous 18-entry run; each row's sum in the order of
// k_prolong_comb.
template <typename T>
__global__ __launch_bounds__(TPB) void k_prolong_comb_node(int64_t nu, const int32_t *__restrict__ parent,
                                                           const T *__restrict__ P, const double *__restrict__ sc,
                                                           int ia, int ib, const T *__restrict__ xa,
                                                           const T *__restrict__ xb, T *__restrict__ x) {
  const int64_t u = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (u >= nu) return;
  const int64_t A = parent[u];
  if (A < 0) return;
  const T ca = ia < 0 ? T(1) : (T)sc[ia];
  const T cb = ib < 0 ? T(1) : (T)sc[ib];
  T c[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) c[j] = xb ? ca * xa[A * 6 + j] + cb * xb[A * 6 + j] : xa[A * 6 + j];
  const T *Pu = P + (size_t)u * 18;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < 6; ++j) acc += Pu[i * 6 + j] * c[j];
    x[u * 3 + i] += acc;
  }
}

// K-cycle second start: kr = b + sc[7] v ; r = kr ; d = it D^-1 kr ; kd = d
template <typename T>
__global__ __launch_bounds__(TPB) void k_lincomb_first(int64_t n, const double *__restrict__ sc,
                                                       const T *__restrict__ b, const T *__restrict__ v,
                                                       T *__restrict__ kr, T *__restrict__ r, const T *__restrict__ dinv,
                                                       double it, T *__restrict__ d, T *__restrict__ kd) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const T c = (T)sc[7];
  const T ri = b[i] + c * v[i];
  kr[i] = ri;
  r[i] = ri;
  const T di = (T)it * dinv[i] * ri;
  d[i] = di;
  kd[i] = di;
}

// K-cycle coefficients of the second inner step from the sums of d.v, d.Ad,
// d.r and the first step's rho1 = sc[0], alpha1 = sc[1] (the formulas of
// k_final_kc<3>).
__device__ __forceinline__ void kc2_coef(const double *sc, const double (&res)[3], double &c5, double &c6) {
  const double rho1 = sc[0], a1 = sc[1], g = res[0], be = res[1], a2 = res[2];
  const double t1 = rho1 != 0.0 ? a1 / rho1 : 0.0;
  const double rho2 = rho1 != 0.0 ? be - g * g / rho1 : 0.0;
  if (rho2 > 0.0) {
    c5 = t1 - g * a2 / (rho1 * rho2);
    c6 = a2 / rho2;
  } else {
    c5 = t1;
    c6 = 0.0;
  }
}

// k_final_kc<2> and k_lincomb_first in one launch: every block sums the nb
// partial rows of c.v, c.b itself (the same fixed order in each block, so
// all blocks agree) and block 0 stores sc[0], sc[1], sc[7].
template <typename T>
__global__ __launch_bounds__(TPB) void k_lincomb_first_kc(int64_t n, const double *__restrict__ part, int nb,
                                                          double *__restrict__ sc, const T *__restrict__ b,
                                                          const T *__restrict__ v, T *__restrict__ kr,
                                                          T *__restrict__ r, const T *__restrict__ dinv, double it,
                                                          T *__restrict__ d, T *__restrict__ kd) {
  double res[2];
  uftr::block_final<2>(part, nb, res);
  const double c7 = res[0] != 0.0 ? -res[1] / res[0] : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[0] = res[0];
    sc[1] = res[1];
    sc[7] = c7;
  }
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const T c = (T)c7;
  const T ri = b[i] + c * v[i];
  kr[i] = ri;
  r[i] = ri;
  const T di = (T)it * dinv[i] * ri;
  d[i] = di;
  kd[i] = di;
}

// k_final_kc<3> and the K-cycle prolongation x_u += P_u (c5 kc + c6 kd) in one
// launch (every block sums the partials; block 0 stores sc[2..6]).
template <int BS, typename T>
__global__ __launch_bounds__(TPB) void k_prolong_kc(int64_t nu, const int32_t *__restrict__ parent,
                                                    const T *__restrict__ P, const double *__restrict__ part, int nb,
                                                    double *__restrict__ sc, const T *__restrict__ xa,
                                                    const T *__restrict__ xb, T *__restrict__ x) {
  double res[3];
  uftr::block_final<3>(part, nb, res);
  double c5, c6;
  kc2_coef(sc, res, c5, c6);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[2] = res[0];
    sc[3] = res[1];
    sc[4] = res[2];
    sc[5] = c5;
    sc[6] = c6;
  }
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= nu * BS) return;
  const int64_t u = t / BS;
  const int i = (int)(t - u * BS);
  const int64_t A = parent[u];
  if (A < 0) return;
  const T ca = (T)c5, cb = (T)c6;
  const T *Pu = P + (size_t)u * BS * 6 + i * 6;
  T acc = T(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const T c = ca * xa[A * 6 + j] + cb * xb[A * 6 + j];
    acc += Pu[j] * c;
  }
  x[t] += acc;
}

// fixed-order sum of K partial columns by one block, then the K-cycle
// coefficients (Notay & Vassilevski 2008, Alg. 2.1 with two inner steps):
//  which 1: sc[0] = rho1 = c.v, sc[1] = alpha1 = c.b, sc[7] = -alpha1/rho1
//  which 2: sc[2] = gamma = d.v, sc[3] = beta = d.Ad, sc[4] = alpha2 = d.r,
//           sc[5], sc[6] = coefficients of c and d in the result
template <int K>
__global__ __launch_bounds__(TPB) void k_final_kc(const double *__restrict__ part, int nb, double *__restrict__ sc) {
  double res[K];
  uftr::block_final<K>(part, nb, res);
  if (threadIdx.x != 0) return;
  if constexpr (K == 2) {
    sc[0] = res[0];
    sc[1] = res[1];
    sc[7] = res[0] != 0.0 ? -res[1] / res[0] : 0.0;
  } else {
    double c5, c6;
    kc2_coef(sc, res, c5, c6);
    sc[2] = res[0];
    sc[3] = res[1];
    sc[4] = res[2];
    sc[5] = c5;
    sc[6] = c6;
  }
}

// The last restriction and the coarsest solve in one workgroup (upper level
// with 6 unknowns per unit): bc = P^T r with one wavefront per coarsest unit
// (fixed order), then xc = Ainv bc in FP64 with the exactly symmetric inverse
// read by columns (loads coalesced over rows; four column groups added in a
// fixed order).  The prolongation x += P xc follows as an ordinary launch.
constexpr int TAIL_TPB = 1024;
constexpr int TAIL_CH = 1024;  // children staged in LDS per pass
constexpr int TAIL_PU = 4;     // prolongation rows per thread in flight
// xc = Ainv P^T r on the coarsest level, one workgroup (a design batch: one
// per design, units [u0, u0 + nu) with its inverse block at Ainv + coff).
// Restriction: one thread per child computes P_u^T r_u, staged in LDS; one
// thread per coarse unknown then adds its unit's children in child order
// (chunks of TAIL_CH children, in order).  Dense product: four groups of
// threads take every fourth column, each with four independent
// accumulators; fixed summation order throughout.  With xup (one design), the
// same workgroup then prolongs: xup_u += P_u xc_parent(u) for the nu_up units
// of the level above (the arithmetic of k_prolong_comb), saving a launch.
// A single workgroup is latency-bound, so the dense product loads 16 inverse
// entries per thread before its first multiply and the prolongation loads
// TAIL_PU rows per thread at once; neither changes an operation's order.
template <typename T, typename TA>
__global__ __launch_bounds__(TAIL_TPB) void k_tail_solve(int nc_units, const int32_t *__restrict__ cptr,
                                                         const int32_t *__restrict__ child, const T *__restrict__ P,
                                                         const TA *__restrict__ Ainv, const T *__restrict__ r,
                                                         T *__restrict__ xc, const int32_t *__restrict__ cu,
                                                         const int64_t *__restrict__ coff, int64_t nu_up,
                                                         const int32_t *__restrict__ parent, T *__restrict__ xup) {
  __shared__ double bc[UFT_TAIL_MAX];
  __shared__ double cv[TAIL_CH * 6];
  __shared__ double part4[4][TAIL_TPB / 4];
  __shared__ T xs[UFT_TAIL_MAX];
  int u0 = 0;
  if (cu) {
    u0 = cu[blockIdx.x];
    nc_units = cu[blockIdx.x + 1] - u0;
    Ainv += coff[blockIdx.x];
    xc += (size_t)u0 * 6;
  }
  const int n = 6 * nc_units;
  for (int t = threadIdx.x; t < n; t += TAIL_TPB) bc[t] = 0.0;
  const int kb = cptr[u0], ke = cptr[u0 + nc_units];
  for (int c0 = kb; c0 < ke; c0 += TAIL_CH) {
    const int c1 = min(c0 + TAIL_CH, ke);
    __syncthreads();  // bc cleared / the previous chunk consumed
    for (int k = c0 + (int)threadIdx.x; k < c1; k += TAIL_TPB) {
      const int64_t u = child[k];
      const T *Pu = P + (size_t)u * 36;
      const T *ru = r + u * 6;
      double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const double ri = (double)ru[i];
#pragma unroll
        for (int j = 0; j < 6; ++j) v[j] += (double)Pu[i * 6 + j] * ri;
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) cv[(k - c0) * 6 + j] = v[j];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += TAIL_TPB) {
      const int A = t / 6, j = t - A * 6;
      const int a0 = max(cptr[u0 + A], c0), a1 = min(cptr[u0 + A + 1], c1);
      double s = bc[t];
      for (int k = a0; k < a1; ++k) s += cv[(k - c0) * 6 + j];
      bc[t] = s;
    }
  }
  __syncthreads();
  const int g = threadIdx.x / (TAIL_TPB / 4), row0 = threadIdx.x % (TAIL_TPB / 4);
  for (int rb = 0; rb < n; rb += TAIL_TPB / 4) {
    const int row = rb + row0;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (row < n) {
      int c = g;
      for (; c + 60 < n; c += 64) {  // columns c, c + 4, ..., c + 60
        TA buf[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) buf[q] = Ainv[(size_t)(c + 4 * q) * n + row];
#pragma unroll
        for (int q = 0; q < 16; ++q) a[q & 3] += (double)buf[q] * bc[c + 4 * q];
      }
      for (; c + 12 < n; c += 16) {
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] += (double)Ainv[(size_t)(c + 4 * q) * n + row] * bc[c + 4 * q];
      }
      for (; c < n; c += 4) a[0] += (double)Ainv[(size_t)c * n + row] * bc[c];
    }
    part4[g][row0] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    if (g == 0 && row < n) {
      const T v = (T)((part4[0][row0] + part4[1][row0]) + (part4[2][row0] + part4[3][row0]));
      xc[row] = v;
      xs[row] = v;
    }
    __syncthreads();
  }
  if (!xup) return;
  const int64_t nt = nu_up * 6;
  for (int64_t t0 = threadIdx.x; t0 < nt; t0 += (int64_t)TAIL_TPB * TAIL_PU) {
    int A[TAIL_PU];
    T pv[TAIL_PU][6], xo[TAIL_PU];
#pragma unroll
    for (int q = 0; q < TAIL_PU; ++q) {
      const int64_t t = t0 + (int64_t)q * TAIL_TPB;
      A[q] = -1;
      if (t < nt) {
        const int64_t u = t / 6;
        const int i = (int)(t - u * 6);
        A[q] = parent[u];
        const T *Pu = P + (size_t)u * 36 + i * 6;
#pragma unroll
        for (int j = 0; j < 6; ++j) pv[q][j] = Pu[j];
        xo[q] = xup[t];
      }
    }
#pragma unroll
    for (int q = 0; q < TAIL_PU; ++q) {
      if (A[q] < 0) continue;
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < 6; ++j) acc += pv[q][j] * xs[A[q] * 6 + j];
      xup[t0 + (int64_t)q * TAIL_TPB] = xo[q] + acc;
    }
  }
}

}  // namespace

namespace uftk {

// one instantiation per epilogue mode
#define UFT_MODE_SWITCH(mode, CALL)                                   \
  switch (mode) {                                                     \
    case ROW_STORE: { constexpr int MODE = ROW_STORE; CALL; } break;  \
    case ROW_STEP: { constexpr int MODE = ROW_STEP; CALL; } break;    \
    case ROW_RES: { constexpr int MODE = ROW_RES; CALL; } break;      \
    case ROW_RESFIRST: { constexpr int MODE = ROW_RESFIRST; CALL; } break; \
    case ROW_RESFIRSTX: { constexpr int MODE = ROW_RESFIRSTX; CALL; } break; \
    case ROW_DOTS2: { constexpr int MODE = ROW_DOTS2; CALL; } break;  \
    case ROW_DOTS3: { constexpr int MODE = ROW_DOTS3; CALL; } break;  \
    default: { constexpr int MODE = ROW_STOREDOT; CALL; } break;      \
  }

template <typename T>
static void bsr6_op_t(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const T *val, const T *u,
                      const RowOp<T> &op) {
  if (lpr_for(nb_rows) == 32)
    UFT_MODE_SWITCH(op.mode, (k_bsr6_op<T, 32, MODE><<<nblk(nb_rows * 6 * 32), TPB, 0, s>>>(nb_rows, rp, ci, val, u, op)))
  else
    UFT_MODE_SWITCH(op.mode, (k_bsr6_op<T, 8, MODE><<<nblk(nb_rows * 6 * 8), TPB, 0, s>>>(nb_rows, rp, ci, val, u, op)))
}
void bsr6_op(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const double *val, const double *u,
             const RowOp<double> &op) {
  bsr6_op_t<double>(s, nb_rows, rp, ci, val, u, op);
}
void bsr6_op(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const float *val, const float *u,
             const RowOp<float> &op) {
  bsr6_op_t<float>(s, nb_rows, rp, ci, val, u, op);
}
int bsr6_op_blocks(int64_t nb_rows) { return (int)nblk(nb_rows * 6 * lpr_for(nb_rows)); }
void bsr6_op_h(hipStream_t s, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const uint16_t *valh,
               const float *dh, const float *u, const RowOp<float> &op) {
  const uint2 *vh = reinterpret_cast<const uint2 *>(valh);
  if (lpr_h(nb_rows) == 32)
    UFT_MODE_SWITCH(op.mode, (k_bsr6_op_h<32, MODE><<<nblk(nb_rows * 32), TPB, 0, s>>>(nb_rows, rp, ci, vh, dh, u, op)))
  else
    UFT_MODE_SWITCH(op.mode, (k_bsr6_op_h<16, MODE><<<nblk(nb_rows * 16), TPB, 0, s>>>(nb_rows, rp, ci, vh, dh, u, op)))
}
int bsr6_op_h_blocks(int64_t nb_rows) { return (int)nblk(nb_rows * lpr_h(nb_rows)); }

template <typename T>
static void gather_op_t(hipStream_t s, const uft_ctx *C, const T *u, const T *Ye, const RowOp<T> &op) {
  const int nn = (int)C->nn, ne = (int)C->ne;
  if (C->npe == 8)
    UFT_MODE_SWITCH(op.mode, (k_node_gather_op<8, T, MODE><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, Ye,
                                                                                    C->d_fixm, u, op)))
  else
    UFT_MODE_SWITCH(op.mode, (k_node_gather_op<4, T, MODE><<<nblk(nn), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, Ye,
                                                                                    C->d_fixm, u, op)))
}
void gather_op(hipStream_t s, const uft_ctx *C, const double *u, const double *Ye, const RowOp<double> &op) {
  gather_op_t<double>(s, C, u, Ye, op);
}
void gather_op(hipStream_t s, const uft_ctx *C, const float *u, const float *Ye, const RowOp<float> &op) {
  gather_op_t<float>(s, C, u, Ye, op);
}
int gather_op_blocks(int64_t nn) { return (int)nblk(nn); }

// the mixed forms' smoother and outer blocks (from the FP32 element blocks)
// and, when a64 / a32 is given, the smoothed prolongator's unscaled blocks
// in FP64 / FP32 too; without the mixed forms' blocks only the latter, from
// the FP64 element blocks (FP64 form) or the FP32 ones
void fine_ell_fill(hipStream_t s, const uft_ctx *C, const double *Ee, const double *dinv, double *a64, float *a32) {
  const int nn = (int)C->nn, ne = (int)C->ne;
  const auto &F = C->fell;
  const dim3 g(nblk(nn), (unsigned)F.wmax);
  const bool blocks = F.d_q8 || F.d_v8 || F.d_w8;
  if (!blocks && !a64 && !a32) return;
  if (blocks || !C->d_KE) {
    const float2 *KEf = (const float2 *)C->d_KEf;
    if (C->npe == 8)
      k_fine_ell_fill<8, float2><<<g, TPB, 0, s>>>(nn, ne, F.d_off, F.d_width, F.d_col, F.d_cptr, F.d_clist, KEf, Ee,
                                                   C->d_fixm, dinv, F.d_v8, F.d_v1, F.d_q8, F.d_qs, F.d_dg, F.d_w8,
                                                   F.d_w1, a64, a32, F.d_kmap, F.d_tv8, F.d_tv1, F.d_tw8, F.d_tw1,
                                                   nullptr, nullptr, F.nslot);
    else
      k_fine_ell_fill<4, float2><<<g, TPB, 0, s>>>(nn, ne, F.d_off, F.d_width, F.d_col, F.d_cptr, F.d_clist, KEf, Ee,
                                                   C->d_fixm, dinv, F.d_v8, F.d_v1, F.d_q8, F.d_qs, F.d_dg, F.d_w8,
                                                   F.d_w1, a64, a32, F.d_kmap, F.d_tv8, F.d_tv1, F.d_tw8, F.d_tw1,
                                                   nullptr, nullptr, F.nslot);
    return;
  }
  const double2 *KE = (const double2 *)C->d_KE;
  if (C->npe == 8)
    k_fine_ell_fill<8, double2><<<g, TPB, 0, s>>>(nn, ne, F.d_off, F.d_width, F.d_col, F.d_cptr, F.d_clist, KE, Ee,
                                                  C->d_fixm, dinv, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                                  nullptr, a64, a32, nullptr, nullptr, nullptr, nullptr, nullptr,
                                                  a64 ? F.d_e8 : nullptr, F.d_e1, F.nslot);
  else
    k_fine_ell_fill<4, double2><<<g, TPB, 0, s>>>(nn, ne, F.d_off, F.d_width, F.d_col, F.d_cptr, F.d_clist, KE, Ee,
                                                  C->d_fixm, dinv, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                                  nullptr, a64, a32, nullptr, nullptr, nullptr, nullptr, nullptr,
                                                  a64 ? F.d_e8 : nullptr, F.d_e1, F.nslot);
}
void fine_ell_op64(hipStream_t s, const uft_ctx *C, const double *u, const RowOp<double> &op) {
  const auto &F = C->fell;
  if (F.d_cd)
    UFT_MODE_SWITCH(op.mode, (k_fine_ell_op64<MODE, int16_t><<<nblk(C->nn), TPB, 0, s>>>(
                                  (int)C->nn, F.nslot, F.d_off, F.d_width, F.d_cd, F.d_e8, F.d_e1, C->d_fixm, u,
                                  op)))
  else
    UFT_MODE_SWITCH(op.mode, (k_fine_ell_op64<MODE, int32_t><<<nblk(C->nn), TPB, 0, s>>>(
                                  (int)C->nn, F.nslot, F.d_off, F.d_width, F.d_col, F.d_e8, F.d_e1, C->d_fixm, u,
                                  op)))
}
void fine_ell_op32(hipStream_t s, const uft_ctx *C, const float *u, const RowOp<float> &op) {
  const auto &F = C->fell;
  if (F.d_cd)
    UFT_MODE_SWITCH(op.mode, (k_fine_ell_op32<MODE, int16_t><<<nblk(C->nn), TPB, 0, s>>>(
                                  (int)C->nn, F.d_off, F.d_width, F.d_cd, F.d_w8, F.d_w1, C->d_fixm, u, op)))
  else
    UFT_MODE_SWITCH(op.mode, (k_fine_ell_op32<MODE, int32_t><<<nblk(C->nn), TPB, 0, s>>>(
                                  (int)C->nn, F.d_off, F.d_width, F.d_col, F.d_w8, F.d_w1, C->d_fixm, u, op)))
}
void helm_ell_fill(hipStream_t s, const uft_ctx *C) {
  const int nn = (int)C->nn, ne = (int)C->ne;
  const auto &F = C->fell;
  const dim3 g(nblk(nn), (unsigned)F.wmax);
  const double2 *KF = (const double2 *)C->d_KF;
  if (C->npe == 8)
    k_helm_ell_fill<8><<<g, TPB, 0, s>>>(nn, ne, F.d_off, F.d_width, F.d_cptr, F.d_clist, KF, F.d_hv);
  else
    k_helm_ell_fill<4><<<g, TPB, 0, s>>>(nn, ne, F.d_off, F.d_width, F.d_cptr, F.d_clist, KF, F.d_hv);
}
void helm_ell_apply(hipStream_t s, const uft_ctx *C, const double *x, double *y, double *dot_out) {
  const auto &F = C->fell;
  const int nn = (int)C->nn;
  double *part = dot_out ? C->d_part : nullptr;
  if (F.d_cd)
    k_helm_ell<int16_t><<<nblk(nn), TPB, 0, s>>>(nn, F.d_off, F.d_width, F.d_cd, F.d_hv, x, y, part, nullptr);
  else
    k_helm_ell<int32_t><<<nblk(nn), TPB, 0, s>>>(nn, F.d_off, F.d_width, F.d_col, F.d_hv, x, y, part, nullptr);
  if (dot_out) final_sum(s, part, (int)nblk(nn), 1, dot_out);
}
void helm_ell_cg(hipStream_t s, const uft_ctx *C, const double *z, const double *sc, int ib, int ibo, int first,
                 const double *stop, double *p, double *y, double *dot_out) {
  const auto &F = C->fell;
  const int nn = (int)C->nn;
  fcg_p_ctl(s, nn, z, sc, ib, ibo, first, stop, p);
  if (F.d_cd)
    k_helm_ell<int16_t><<<nblk(nn), TPB, 0, s>>>(nn, F.d_off, F.d_width, F.d_cd, F.d_hv, p, y, C->d_part, stop);
  else
    k_helm_ell<int32_t><<<nblk(nn), TPB, 0, s>>>(nn, F.d_off, F.d_width, F.d_col, F.d_hv, p, y, C->d_part, stop);
  final_sum(s, C->d_part, (int)nblk(nn), 1, dot_out);
}
void fine_ell_op(hipStream_t s, const uft_ctx *C, const float *u, const RowOp<float> &op) {
  const auto &F = C->fell;
  if (F.d_q8) {
    if (F.d_cd)
      UFT_MODE_SWITCH(op.mode, (k_fine_ell_op_q<MODE, int16_t><<<nblk(C->nn), TPB, 0, s>>>(
                                    (int)C->nn, F.d_off, F.d_width, F.d_cd, F.d_q8, F.d_qs, F.d_dg, C->d_fixm, u, op)))
    else
      UFT_MODE_SWITCH(op.mode, (k_fine_ell_op_q<MODE, int32_t><<<nblk(C->nn), TPB, 0, s>>>(
                                    (int)C->nn, F.d_off, F.d_width, F.d_col, F.d_q8, F.d_qs, F.d_dg, C->d_fixm, u, op)))
    return;
  }
  if (F.d_cd)
    UFT_MODE_SWITCH(op.mode, (k_fine_ell_op<true, MODE, int16_t><<<nblk(C->nn), TPB, 0, s>>>(
                                  (int)C->nn, F.d_off, F.d_width, F.d_cd, F.d_v8, F.d_v1, F.d_dg, C->d_fixm, u, op)))
  else
    UFT_MODE_SWITCH(op.mode, (k_fine_ell_op<true, MODE, int32_t><<<nblk(C->nn), TPB, 0, s>>>(
                                  (int)C->nn, F.d_off, F.d_width, F.d_col, F.d_v8, F.d_v1, F.d_dg, C->d_fixm, u, op)))
}

template <typename T>
static void restrict_first_t(hipStream_t s, int bs, int64_t nc, const int32_t *cptr, const int32_t *child,
                             const T *P, const T *r, T *bc, const T *dinv, double it, T *rc, T *dc, T *xc) {
  if (bs == 3)
    k_restrict_first<3, 16, T><<<nblk(nc * 16), TPB, 0, s>>>(nc, cptr, child, P, r, bc, dinv, it, rc, dc, xc);
  else
    k_restrict_first<6, 64, T><<<nblk(nc * 64), TPB, 0, s>>>(nc, cptr, child, P, r, bc, dinv, it, rc, dc, xc);
}
void restrict_first(hipStream_t s, int bs, int64_t nc, const int32_t *cptr, const int32_t *child, const double *P,
                    const double *r, double *bc, const double *dinv, double it, double *rc, double *dc, double *xc) {
  restrict_first_t<double>(s, bs, nc, cptr, child, P, r, bc, dinv, it, rc, dc, xc);
}
void restrict_first(hipStream_t s, int bs, int64_t nc, const int32_t *cptr, const int32_t *child, const float *P,
                    const float *r, float *bc, const float *dinv, double it, float *rc, float *dc, float *xc) {
  restrict_first_t<float>(s, bs, nc, cptr, child, P, r, bc, dinv, it, rc, dc, xc);
}

template <typename T>
static void prolong_comb_t(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const T *P, const double *sc,
                           int ia, int ib, const T *xa, const T *xb, T *x) {
  if (bs == 3)
    k_prolong_comb_node<T><<<nblk(nu), TPB, 0, s>>>(nu, parent, P, sc, ia, ib, xa, xb, x);
  else
    k_prolong_comb<6, T><<<nblk(nu * 6), TPB, 0, s>>>(nu, parent, P, sc, ia, ib, xa, xb, x);
}
void prolong_comb(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const double *P, const double *sc, int ia,
                  int ib, const double *xa, const double *xb, double *x) {
  prolong_comb_t<double>(s, bs, nu, parent, P, sc, ia, ib, xa, xb, x);
}
void prolong_comb(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const float *P, const double *sc, int ia,
                  int ib, const float *xa, const float *xb, float *x) {
  prolong_comb_t<float>(s, bs, nu, parent, P, sc, ia, ib, xa, xb, x);
}

void lincomb_first(hipStream_t s, int64_t n, const double *sc, const double *b, const double *v, double *kr, double *r,
                   const double *dinv, double it, double *d, double *kd) {
  k_lincomb_first<double><<<nblk(n), TPB, 0, s>>>(n, sc, b, v, kr, r, dinv, it, d, kd);
}
void lincomb_first(hipStream_t s, int64_t n, const double *sc, const float *b, const float *v, float *kr, float *r,
                   const float *dinv, double it, float *d, float *kd) {
  k_lincomb_first<float><<<nblk(n), TPB, 0, s>>>(n, sc, b, v, kr, r, dinv, it, d, kd);
}

void final_kc1(hipStream_t s, const double *part, int nb, double *sc) { k_final_kc<2><<<1, TPB, 0, s>>>(part, nb, sc); }
template <typename T>
static void lincomb_first_kc_t(hipStream_t s, int64_t n, const double *part, int nb, double *sc, const T *b,
                               const T *v, T *kr, T *r, const T *dinv, double it, T *d, T *kd) {
  k_lincomb_first_kc<T><<<nblk(n), TPB, 0, s>>>(n, part, nb, sc, b, v, kr, r, dinv, it, d, kd);
}
void lincomb_first_kc(hipStream_t s, int64_t n, const double *part, int nb, double *sc, const double *b,
                      const double *v, double *kr, double *r, const double *dinv, double it, double *d, double *kd) {
  lincomb_first_kc_t<double>(s, n, part, nb, sc, b, v, kr, r, dinv, it, d, kd);
}
void lincomb_first_kc(hipStream_t s, int64_t n, const double *part, int nb, double *sc, const float *b,
                      const float *v, float *kr, float *r, const float *dinv, double it, float *d, float *kd) {
  lincomb_first_kc_t<float>(s, n, part, nb, sc, b, v, kr, r, dinv, it, d, kd);
}
template <typename T>
static void prolong_kc_t(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const T *P, const double *part,
                         int nb, double *sc, const T *xa, const T *xb, T *x) {
  if (bs == 3)
    k_prolong_kc<3, T><<<nblk(nu * 3), TPB, 0, s>>>(nu, parent, P, part, nb, sc, xa, xb, x);
  else
    k_prolong_kc<6, T><<<nblk(nu * 6), TPB, 0, s>>>(nu, parent, P, part, nb, sc, xa, xb, x);
}
void prolong_kc(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const double *P, const double *part, int nb,
                double *sc, const double *xa, const double *xb, double *x) {
  prolong_kc_t<double>(s, bs, nu, parent, P, part, nb, sc, xa, xb, x);
}
void prolong_kc(hipStream_t s, int bs, int64_t nu, const int32_t *parent, const float *P, const double *part, int nb,
                double *sc, const float *xa, const float *xb, float *x) {
  prolong_kc_t<float>(s, bs, nu, parent, P, part, nb, sc, xa, xb, x);
}
void final_kc2(hipStream_t s, const double *part, int nb, double *sc) { k_final_kc<3><<<1, TPB, 0, s>>>(part, nb, sc); }

void tail_solve(hipStream_t s, int nc_units, const int32_t *cptr, const int32_t *child, const double *P,
                const double *Ainv, const double *r, double *xc, int nseg, const int32_t *cu, const int64_t *coff) {
  k_tail_solve<double, double><<<nseg, TAIL_TPB, 0, s>>>(nc_units, cptr, child, P, Ainv, r, xc, cu, coff, 0, nullptr,
                                                          nullptr);
}
void tail_solve(hipStream_t s, int nc_units, const int32_t *cptr, const int32_t *child, const float *P,
                const double *Ainv, const float *r, float *xc, int nseg, const int32_t *cu, const int64_t *coff) {
  k_tail_solve<float, double><<<nseg, TAIL_TPB, 0, s>>>(nc_units, cptr, child, P, Ainv, r, xc, cu, coff, 0, nullptr,
                                                         nullptr);
}
void tail_solve_prolong(hipStream_t s, int nc_units, const int32_t *cptr, const int32_t *child, const double *P,
                        const double *Ainv, const double *r, double *xc, int64_t nu_up, const int32_t *parent,
                        double *x) {
  k_tail_solve<double, double><<<1, TAIL_TPB, 0, s>>>(nc_units, cptr, child, P, Ainv, r, xc, nullptr, nullptr, nu_up,
                                                      parent, x);
}
void tail_solve_prolong(hipStream_t s, int nc_units, const int32_t *cptr, const int32_t *child, const float *P,
                        const float *Ainv, const float *r, float *xc, int64_t nu_up, const int32_t *parent,
                        float *x) {
  k_tail_solve<float, float><<<1, TAIL_TPB, 0, s>>>(nc_units, cptr, child, P, Ainv, r, xc, nullptr, nullptr, nu_up,
                                                    parent, x);
}

}  // namespace uftk
