// gfx950 kernels of the multi-load-case form: NR right-hand sides (NR = 2,
// 4 or 8; unused ones are zero) solved together.  Every vector is stored
// interleaved, [unknown][NR], so one read of an operator entry (element
// block, 6x6 coarse block, transfer block) serves all NR vectors, and the NR
// values of one unknown sit in one 16-64 byte segment.  Per-vector scalars
// (Krylov and K-cycle coefficients) live in device arrays indexed by the
// right-hand side.
//
// The fine element products of the mixed form run on the matrix cores:
// Y_e (nd x NR) = KE_e (nd x nd) U_e (nd x NR) as v_mfma_f32_16x16x32_bf16
// tiles (BF16 operands, FP32 accumulation), one element per wavefront step.
// Element-local results are stored node-row major per element,
// [ne][npe][3][NR], so the node phase reads each incidence's 3 x NR values
// as one contiguous segment.
#include <hip/hip_fp16.h>

#include "uft_elem.h"
#include "uft_internal.h"
#include "uft_reduce.h"

namespace {
using namespace uftd;

constexpr int TPB = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// Per-RHS block partials: v[k] of lane (.., r = lane % NR) is summed over the
// lanes with the same r (fixed shuffle tree, then the four waves in order)
// into part[block][k * NR + r].
template <int K, int NR>
__device__ __forceinline__ void mr_block_partials(double (&v)[K], double *part, int64_t blk = -1) {
  if (blk < 0) blk = blockIdx.x;
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int off = 32; off >= NR; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  }
  __shared__ double sm[K][NR][TPB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < NR) {
#pragma unroll
    for (int k = 0; k < K; ++k) sm[k][lane][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < K * NR) {
    const int k = threadIdx.x / NR, r = threadIdx.x % NR;
    part[(size_t)blk * K * NR + threadIdx.x] = (sm[k][r][0] + sm[k][r][1]) + (sm[k][r][2] + sm[k][r][3]);
  }
}

// per-row epilogue of a fused operator apply (the single-RHS RowOp modes);
// scalar coefficients that differ per right-hand side come from sc[r * 8 + .]
template <typename T>
struct MrOp {
  int mode;
  int addprev;
  T *y, *r, *dout, *x, *kr, *kd;
  const T *base, *dinv, *w1, *w2;
  double c1, c2, it;
  double *part;
};
enum { MR_STORE = 0, MR_STEP, MR_RES, MR_RESFIRST, MR_RESFIRSTX, MR_DOTS2, MR_DOTS3, MR_STOREDOT };

// element row i (dinv index), interleaved index ii = i * NR + rr
template <typename T>
__device__ __forceinline__ void mr_epi(const MrOp<T> &op, int64_t i, int64_t ii, T t, T ui, double (&dv)[3]) {
  switch (op.mode) {
    case MR_STORE:
      op.y[ii] = t;
      break;
    case MR_STEP: {
      const T ri = op.r[ii] - t;
      op.r[ii] = ri;
      const T dn = (T)op.c1 * ui + (T)op.c2 * op.dinv[i] * ri;
      op.dout[ii] = dn;
      op.x[ii] += op.addprev ? dn + ui : dn;
      break;
    }
    case MR_RES:
      op.r[ii] = op.base[ii] - t;
      break;
    case MR_RESFIRST: {
      const T ri = op.base[ii] - t;
      op.r[ii] = ri;
      op.dout[ii] = (T)op.it * op.dinv[i] * ri;
      break;
    }
    case MR_RESFIRSTX: {
      const T ri = op.base[ii] - t;
      op.r[ii] = ri;
      op.kr[ii] = ri;
      const T dn = (T)op.it * op.dinv[i] * ri;
      op.dout[ii] = dn;
      op.kd[ii] = dn;
      break;
    }
    case MR_DOTS2:
      op.y[ii] = t;
      dv[0] += (double)ui * (double)t;
      dv[1] += (double)ui * (double)op.w1[ii];
      break;
    case MR_DOTS3:
      dv[0] += (double)ui * (double)op.w1[ii];
      dv[1] += (double)ui * (double)t;
      dv[2] += (double)ui * (double)op.w2[ii];
      break;
    case MR_STOREDOT:
      op.y[ii] = t;
      dv[0] += (double)ui * (double)t;
      break;
    default:
      break;
  }
}
template <typename T, int NR>
__device__ __forceinline__ void mr_epi_partials(const MrOp<T> &op, double (&dv)[3], int64_t blk = -1) {
  if (op.mode == MR_DOTS2) {
    double v[2] = {dv[0], dv[1]};
    mr_block_partials<2, NR>(v, op.part, blk);
  } else if (op.mode == MR_DOTS3) {
    mr_block_partials<3, NR>(dv, op.part, blk);
  } else if (op.mode == MR_STOREDOT) {
    double v[1] = {dv[0]};
    mr_block_partials<1, NR>(v, op.part, blk);
  }
}

// ---------------------------------------------------------------- elementwise
template <typename T, int NR>
__global__ __launch_bounds__(TPB) void k_mr_cheb_first(int64_t n, const T *__restrict__ b, const T *__restrict__ dinv,
                                                        double it, T *__restrict__ r, T *__restrict__ d,
                                                        T *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= n * NR) return;
  const T ri = b[t];
  r[t] = ri;
  const T di = (T)it * dinv[t / NR] * ri;
  d[t] = di;
  x[t] = di;
}

// ---------------------------------------------------------------- element phase, vector form
// lane (e, rr), rr = t % NR: the NR lanes of an element read the same packed
// KE pair (one coalesced request) and their own vector.
template <int NPE, typename T, typename T2, int NR>
__global__ __launch_bounds__(TPB) void k_mr_elem(int ne, const int32_t *__restrict__ elem, const T2 *__restrict__ KE,
                                                 const double *__restrict__ xphys, SimpCoef sc,
                                                 const uint8_t *__restrict__ fixm, const T *__restrict__ u,
                                                 T *__restrict__ Ye) {
  constexpr int ND = 3 * NPE, NP = ND * (ND + 1) / 2, NQ = NP / 2;
  const int64_t t = uftr::xcd_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
  const int64_t e = t / NR;
  const int rr = (int)(t % NR);
  if (e >= ne) return;
  T ue[ND], ye[ND];
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    const int n = elem[(size_t)a * ne + e];
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T v = u[(3 * (size_t)n + c) * NR + rr];
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
  for (int k = 0; k < ND; ++k) Ye[(((size_t)e * NPE + k % NPE) * 3 + k / NPE) * NR + rr] = E * ye[k];
}

// ---------------------------------------------------------------- element phase, matrix cores
// Y_e (nd x 8) = E(x_e) KE_e U_e for the 8 right-hand sides on the matrix
// cores (E(x_e) from the per-element moduli of the last hierarchy refresh,
// which used the same density)
// cores; per element two (hex8) or one (tet4) v_mfma_f32_16x16x32_bf16 tiles
// (BF16 operands, FP32 accumulation).  Operand layout (lane l):
// A[l % 16][8 (l / 16) + j], B[8 (l / 16) + j][l % 16], D[4 (l / 16) + i][l % 16];
// A rows are element DOFs (16 per tile), K is the element DOF (padded to 32),
// the columns are the right-hand sides (8 of 16 used).  The element blocks
// are stored for this product as full BF16 rows ([ne][nd][ndp], ndp = nd
// rounded up to 8, zero padded) so each lane's A fragment is one 16-byte
// load; U_e of MF_EL elements is staged in LDS as BF16 (fixed DOFs zero,
// one 96-byte node row per load) and read as one 16-byte B fragment.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int MF_WAVES = 4;  // wavefronts per workgroup
constexpr int MF_EL = 16;    // elements per wavefront batch

template <int NPE>
__global__ __launch_bounds__(64 * MF_WAVES) void k_mr_elem_mfma(int ne, const int32_t *__restrict__ elem,
                                                               const uint4 *__restrict__ KEr,
                                                               const double *__restrict__ Ee,
                                                               const uint8_t *__restrict__ fixm,
                                                               const float *__restrict__ u, float *__restrict__ Ye) {
  constexpr int ND = 3 * NPE, NDP = (ND + 7) / 8 * 8, CH = NDP / 8, TILES = (ND + 15) / 16;
  constexpr int NR = 8;
  __shared__ __attribute__((aligned(16))) __bf16 ub[MF_WAVES][MF_EL][NR][32];  // U_e^T, K padded to 32
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t e0 = ((int64_t)blockIdx.x * MF_WAVES + w) * MF_EL;
  {
    const int el = lane % MF_EL, rs = lane / MF_EL;
    const int64_t e = e0 + el;
    const bool ok = e < ne;
    for (int a = rs; a < NPE; a += 64 / MF_EL) {
      float4 v[6];
      unsigned m = 7u;
      if (ok) {
        const int n = elem[(size_t)a * ne + e];
        m = fixm[n];
        const float4 *row = reinterpret_cast<const float4 *>(u + 3 * (size_t)n * NR);
#pragma unroll
        for (int h = 0; h < 6; ++h) v[h] = row[h];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const bool z = !ok || ((m >> c) & 1u);
        const float4 lo = v[2 * c], hi = v[2 * c + 1];
        const float vals[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
        for (int r = 0; r < NR; ++r) ub[w][el][r][c * NPE + a] = (__bf16)(z ? 0.f : vals[r]);
      }
    }
    for (int t = lane; t < MF_EL * NR * (32 - ND); t += 64) {
      const int el2 = t / (NR * (32 - ND)), rem = t % (NR * (32 - ND));
      ub[w][el2][rem / (32 - ND)][ND + rem % (32 - ND)] = (__bf16)0.f;
    }
  }
  __syncthreads();
  const int col = lane % 16, kq = lane / 16, kb = 8 * kq;
  uint4 an[TILES];
#pragma unroll
  for (int tl = 0; tl < TILES; ++tl) an[tl] = make_uint4(0u, 0u, 0u, 0u);
  if (e0 < ne) {
#pragma unroll
    for (int tl = 0; tl < TILES; ++tl) {
      const int row = 16 * tl + col;
      if (row < ND && kb < ND) an[tl] = KEr[((size_t)e0 * ND + row) * CH + kq];
    }
  }
  for (int el = 0; el < MF_EL; ++el) {
    const int64_t e = e0 + el;
    if (e >= ne) break;  // uniform across the wavefront
    uint4 ac[TILES];
#pragma unroll
    for (int tl = 0; tl < TILES; ++tl) ac[tl] = an[tl];
    if (el + 1 < MF_EL && e + 1 < ne) {  // the next element's blocks in flight
#pragma unroll
      for (int tl = 0; tl < TILES; ++tl) {
        const int row = 16 * tl + col;
        if (row < ND && kb < ND) an[tl] = KEr[((size_t)(e + 1) * ND + row) * CH + kq];
      }
    }
    bf16x8 bv;
    if (col < NR) {
      bv = *reinterpret_cast<const bf16x8 *>(&ub[w][el][col][kb]);
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) bv[j] = (__bf16)0.f;
    }
    const float E = (float)Ee[e];
    float *Yel = Ye + (size_t)e * ND * NR;
#pragma unroll
    for (int tl = 0; tl < TILES; ++tl) {
      const bf16x8 av = __builtin_bit_cast(bf16x8, ac[tl]);
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
      if (col < NR) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * tl + 4 * kq + i;
          if (row < ND) Yel[((row % NPE) * 3 + row / NPE) * NR + col] = E * acc[i];
        }
      }
    }
  }
}

// FP64 packed KE -> full BF16 rows [ne][nd][ndp] (round to nearest even,
// zero padded): one thread per (element, row)
template <int NPE>
__global__ __launch_bounds__(TPB) void k_ke_to_bf16_rows(int ne, const double2 *__restrict__ KE,
                                                         __bf16 *__restrict__ out) {
  constexpr int ND = 3 * NPE, NDP = (ND + 7) / 8 * 8;
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= (int64_t)ne * ND) return;
  const int64_t e = t / ND;
  const int row = (int)(t - e * ND);
  __bf16 *o = out + t * NDP;
  for (int c = 0; c < NDP; ++c) {
    float v = 0.f;
    if (c < ND) {
      const int p = pk_index(row, c, ND);
      const double2 pr = KE[(size_t)(p >> 1) * ne + e];
      v = (float)((p & 1) ? pr.y : pr.x);
    }
    o[c] = (__bf16)v;
  }
}

// ---------------------------------------------------------------- node phase
template <int NPE, typename T, int NR>
__global__ __launch_bounds__(TPB) void k_mr_gather_op(int nn, int ne, const int32_t *__restrict__ iptr,
                                                      const int32_t *__restrict__ inc, const T *__restrict__ Ye,
                                                      const uint8_t *__restrict__ fixm, const T *__restrict__ u,
                                                      MrOp<T> op) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  const int64_t n = t / NR;
  const int rr = (int)(t % NR);
  double dv[3] = {0.0, 0.0, 0.0};
  if (n < nn) {
    T acc[3] = {T(0), T(0), T(0)};
    const int k0 = iptr[n], k1 = iptr[n + 1];
    for (int k = k0; k < k1; ++k) {
      const int v = inc[k];
      const int e = v / NPE, l = v - e * NPE;
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[c] += Ye[(((size_t)e * NPE + l) * 3 + c) * NR + rr];
    }
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = 3 * n + c, ii = i * NR + rr;
      const T ui = u[ii];
      mr_epi(op, i, ii, ((m >> c) & 1u) ? ui : acc[c], ui, dv);
    }
  }
  mr_epi_partials<T, NR>(op, dv);
}

// ---------------------------------------------------------------- assembled fine operator
// The sliced-ELL 3x3 node blocks (one node pattern, see the single right-hand
// side kernels) applied to NR interleaved right-hand sides: lane (n, rr), the
// NR lanes of a node read the same block (one request) and their own vector
// entries, so each block streams once for all right-hand sides.  H: FP16
// blocks of D^-1 A (rows scaled by dg), else FP32 blocks.
template <int NR, bool H, typename CI>
__global__ __launch_bounds__(TPB) void k_mr_ell_op(int nn, const int64_t *__restrict__ off,
                                                   const int32_t *__restrict__ width, const CI *__restrict__ col,
                                                   const uint4 *__restrict__ v8, const uint16_t *__restrict__ v1,
                                                   const float *__restrict__ dg, const float4 *__restrict__ w8,
                                                   const float *__restrict__ w1, const uint8_t *__restrict__ fixm,
                                                   const float *__restrict__ u, MrOp<float> op) {
  const int64_t blk = uftr::xcd_block(blockIdx.x, gridDim.x);
  const int64_t t = blk * TPB + threadIdx.x;
  const int64_t n = t / NR;
  const int rr = (int)(t % NR);
  double dv[3] = {0.0, 0.0, 0.0};
  if (n < nn) {
    const int64_t w = n / UFT_ELL_SLICE;
    const int64_t base = off[w] + n % UFT_ELL_SLICE;
    const int W = width[w];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll 4
    for (int s = 0; s < W; ++s) {
      const int64_t k = base + (int64_t)s * UFT_ELL_SLICE;
      const int64_t j = std::is_same<CI, int16_t>::value ? n + (int64_t)col[k] : (int64_t)col[k];
      float e[9];
      if constexpr (H) {
        const uint4 h = v8[k];
        const unsigned hw[4] = {h.x, h.y, h.z, h.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float2 f = __half22float2(*reinterpret_cast<const __half2 *>(&hw[q]));
          e[2 * q] = f.x;
          e[2 * q + 1] = f.y;
        }
        e[8] = __half2float(__ushort_as_half(v1[k]));
      } else {
        const float4 a = w8[2 * k], b = w8[2 * k + 1];
        e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
        e[4] = b.x; e[5] = b.y; e[6] = b.z; e[7] = b.w;
        e[8] = w1[k];
      }
      const float x0 = u[(3 * j) * NR + rr], x1 = u[(3 * j + 1) * NR + rr], x2 = u[(3 * j + 2) * NR + rr];
      a0 += e[0] * x0 + e[1] * x1 + e[2] * x2;
      a1 += e[3] * x0 + e[4] * x1 + e[5] * x2;
      a2 += e[6] * x0 + e[7] * x1 + e[8] * x2;
    }
    const float acc[3] = {a0, a1, a2};
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = 3 * n + c, ii = i * NR + rr;
      const float ui = u[ii];
      const float tv = H ? dg[i] * acc[c] : acc[c];
      mr_epi(op, i, ii, ((m >> c) & 1u) ? ui : tv, ui, dv);
    }
  }
  mr_epi_partials<float, NR>(op, dv, blk);
}

// The FP64 form's load-case operator on the assembled FP64 node blocks (the
// refresh's, in four planes of 16-byte pairs plus the ninth entry): lane
// (n, rr) as above, the NR lanes of a node share each block load, FP64
// throughout; replaces the element phase and its element-local results.
template <int NR, typename CI>
__global__ __launch_bounds__(TPB) void k_mr_ell_op64(int nn, int64_t ns, const int64_t *__restrict__ off,
                                                     const int32_t *__restrict__ width, const CI *__restrict__ col,
                                                     const double2 *__restrict__ e8, const double *__restrict__ e1,
                                                     const uint8_t *__restrict__ fixm, const double *__restrict__ u,
                                                     MrOp<double> op) {
  const int64_t blk = uftr::xcd_block(blockIdx.x, gridDim.x);
  const int64_t t = blk * TPB + threadIdx.x;
  const int64_t n = t / NR;
  const int rr = (int)(t % NR);
  double dv[3] = {0.0, 0.0, 0.0};
  if (n < nn) {
    const int64_t w = n / UFT_ELL_SLICE;
    const int64_t base = off[w] + n % UFT_ELL_SLICE;
    const int W = width[w];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll 2
    for (int s = 0; s < W; ++s) {
      const int64_t k = base + (int64_t)s * UFT_ELL_SLICE;
      const int64_t j = std::is_same<CI, int16_t>::value ? n + (int64_t)col[k] : (int64_t)col[k];
      const double2 b0 = e8[k], b1 = e8[ns + k], b2 = e8[2 * ns + k], b3 = e8[3 * ns + k];
      const double b8 = e1[k];
      const double x0 = u[(3 * j) * NR + rr], x1 = u[(3 * j + 1) * NR + rr], x2 = u[(3 * j + 2) * NR + rr];
      a0 += b0.x * x0 + b0.y * x1 + b1.x * x2;
      a1 += b1.y * x0 + b2.x * x1 + b2.y * x2;
      a2 += b3.x * x0 + b3.y * x1 + b8 * x2;
    }
    const double acc[3] = {a0, a1, a2};
    const unsigned m = fixm[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = 3 * n + c, ii = i * NR + rr;
      const double ui = u[ii];
      mr_epi(op, i, ii, ((m >> c) & 1u) ? ui : acc[c], ui, dv);
    }
  }
  mr_epi_partials<double, NR>(op, dv, blk);
}

// The same product with one lane per node: the lane holds the node's NR
// right-hand sides, so a slot's block is loaded once per node (not once per
// right-hand side), the neighbour's 3 NR vector entries are one contiguous
// run read with 16-byte (NR >= 4) or 8-byte loads, and the wavefront's 64
// consecutive nodes read 64 consecutive slots: NR times fewer memory
// instructions than a lane per (node, right-hand side).  MODE is the rows'
// epilogue (MR_STORE ... MR_RESFIRSTX, MR_STOREDOT).  Measured on c1 with 8
// load cases: the FP32 outer product (MR_STOREDOT) 167 against 207 us, the
// FP16 residual 149 against 188 us, the Chebyshev step 204 against 190 us
// (its epilogue touches three vectors with 32-byte runs per lane), so
// mr_ell_op keeps the lane per (node, right-hand side) for that step.  Two
// lanes per node (four right-hand sides each) ran slower in every mode.
template <int NR>
__device__ __forceinline__ void ld_nr(const float *__restrict__ p, float (&o)[NR]) {
  if constexpr (NR % 4 == 0) {
#pragma unroll
    for (int q = 0; q < NR / 4; ++q) {
      const float4 t = reinterpret_cast<const float4 *>(p)[q];
      o[4 * q] = t.x; o[4 * q + 1] = t.y; o[4 * q + 2] = t.z; o[4 * q + 3] = t.w;
    }
  } else if constexpr (NR % 2 == 0) {
#pragma unroll
    for (int q = 0; q < NR / 2; ++q) {
      const float2 t = reinterpret_cast<const float2 *>(p)[q];
      o[2 * q] = t.x; o[2 * q + 1] = t.y;
    }
  } else {
    o[0] = p[0];
  }
}
template <int NR>
__device__ __forceinline__ void st_nr(float *__restrict__ p, const float (&o)[NR]) {
  if constexpr (NR % 4 == 0) {
#pragma unroll
    for (int q = 0; q < NR / 4; ++q)
      reinterpret_cast<float4 *>(p)[q] = make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
  } else if constexpr (NR % 2 == 0) {
#pragma unroll
    for (int q = 0; q < NR / 2; ++q) reinterpret_cast<float2 *>(p)[q] = make_float2(o[2 * q], o[2 * q + 1]);
  } else {
    p[0] = o[0];
  }
}

// The rows of node n once its products a[c][r] = (A u)_(3n+c, r) are
// formed: fixed DOFs take t = u, H scales by the row's diagonal, then MODE's
// epilogue on the node's 3 NR entries (16- or 8-byte runs); MR_STOREDOT adds
// u.t per right-hand side into dv.  load() reads every operand of the three
// rows (it does not depend on the products, so a kernel may issue it before
// its slot loop); finish() forms and stores the rows.
template <int NR, bool H, int MODE>
struct NodeRows {
  static constexpr bool kA = MODE == MR_STEP || MODE == MR_RES || MODE == MR_RESFIRST || MODE == MR_RESFIRSTX;
  static constexpr bool kB = MODE == MR_STEP;
  static constexpr bool kD = MODE == MR_STEP || MODE == MR_RESFIRST || MODE == MR_RESFIRSTX;
  float ui[3][NR], va[3][kA ? NR : 1], vb[3][kB ? NR : 1], sc[3], di[3];
  unsigned m;
  __device__ __forceinline__ void load(const MrOp<float> &op, int64_t n, const float *__restrict__ u,
                                       const float *__restrict__ dg, const uint8_t *__restrict__ fixm) {
    m = fixm[n];
    const int64_t i0 = 3 * n, ii0 = i0 * NR;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ld_nr<NR>(u + ii0 + c * NR, ui[c]);
      sc[c] = H ? dg[i0 + c] : 1.f;
      if constexpr (MODE == MR_STEP) {
        ld_nr<NR>(op.r + ii0 + c * NR, va[c]);
        ld_nr<NR>(op.x + ii0 + c * NR, vb[c]);
      } else if constexpr (kA) {
        ld_nr<NR>(op.base + ii0 + c * NR, va[c]);
      }
      if constexpr (kD) di[c] = op.dinv[i0 + c];
    }
  }
  __device__ __forceinline__ void finish(const MrOp<float> &op, int64_t n, const float (&a)[3][NR], double (&dv)[NR]) {
    const int64_t ii0 = 3 * n * NR;
    float tv[3][NR];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const bool fx = (m >> c) & 1u;
#pragma unroll
      for (int r = 0; r < NR; ++r) tv[c][r] = fx ? ui[c][r] : sc[c] * a[c][r];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t ii = ii0 + c * NR;
      if constexpr (MODE == MR_STORE) {
        st_nr<NR>(op.y + ii, tv[c]);
      } else if constexpr (MODE == MR_STEP) {
        float dn[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          va[c][r] -= tv[c][r];
          dn[r] = (float)op.c1 * ui[c][r] + (float)op.c2 * di[c] * va[c][r];
          vb[c][r] += op.addprev ? dn[r] + ui[c][r] : dn[r];
        }
        st_nr<NR>(op.r + ii, va[c]);
        st_nr<NR>(op.dout + ii, dn);
        st_nr<NR>(op.x + ii, vb[c]);
      } else if constexpr (MODE == MR_RES) {
#pragma unroll
        for (int r = 0; r < NR; ++r) va[c][r] -= tv[c][r];
        st_nr<NR>(op.r + ii, va[c]);
      } else if constexpr (MODE == MR_RESFIRST || MODE == MR_RESFIRSTX) {
        float dn[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          va[c][r] -= tv[c][r];
          dn[r] = (float)op.it * di[c] * va[c][r];
        }
        st_nr<NR>(op.r + ii, va[c]);
        st_nr<NR>(op.dout + ii, dn);
        if constexpr (MODE == MR_RESFIRSTX) {
          st_nr<NR>(op.kr + ii, va[c]);
          st_nr<NR>(op.kd + ii, dn);
        }
      } else {  // MR_STOREDOT
        st_nr<NR>(op.y + ii, tv[c]);
#pragma unroll
        for (int r = 0; r < NR; ++r) dv[r] += (double)ui[c][r] * (double)tv[c][r];
      }
    }
  }
};
// MR_STOREDOT's block partials, one lane per node: per right-hand side a
// fixed shuffle tree over the wavefront, then the four wavefronts in order,
// into part[blk][r]
template <int NR>
__device__ __forceinline__ void mr_node_partials(const MrOp<float> &op, double (&dv)[NR], int64_t blk) {
  __shared__ double sm[NR][TPB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double v = dv[r];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) sm[r][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < NR) {
    const int r = threadIdx.x;
    op.part[(size_t)blk * NR + r] = (sm[r][0] + sm[r][1]) + (sm[r][2] + sm[r][3]);
  }
}
// one slot's 3x3 block as loaded (FP16: eight entries in 16 bytes and the
// ninth; FP32: two float4 and the ninth), converted when used; the block
// stream is read once per apply, so its loads are non-temporal and leave
// L2 to the gathered vector entries
template <bool H>
struct RawBlk;
template <>
struct RawBlk<true> {
  uint4 v;
  uint16_t w;
  template <typename I>
  __device__ __forceinline__ void load(I k, const uint4 *__restrict__ v8, const uint16_t *__restrict__ v1,
                                       const float4 *__restrict__, const float *__restrict__) {
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    const v4u t = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(&v8[k]));
    v = make_uint4(t.x, t.y, t.z, t.w);
    w = __builtin_nontemporal_load(&v1[k]);
  }
  __device__ __forceinline__ void get(float (&e)[9]) const {
    const unsigned hw[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float2 f = __half22float2(*reinterpret_cast<const __half2 *>(&hw[q]));
      e[2 * q] = f.x;
      e[2 * q + 1] = f.y;
    }
    e[8] = __half2float(__ushort_as_half(w));
  }
};
template <>
struct RawBlk<false> {
  float4 a, b;
  float c;
  template <typename I>
  __device__ __forceinline__ void load(I k, const uint4 *__restrict__, const uint16_t *__restrict__,
                                       const float4 *__restrict__ w8, const float *__restrict__ w1) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f ta = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&w8[2 * k]));
    const v4f tb = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(&w8[2 * k + 1]));
    a = make_float4(ta.x, ta.y, ta.z, ta.w);
    b = make_float4(tb.x, tb.y, tb.z, tb.w);
    c = __builtin_nontemporal_load(&w1[k]);
  }
  __device__ __forceinline__ void get(float (&e)[9]) const {
    e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
    e[4] = b.x; e[5] = b.y; e[6] = b.z; e[7] = b.w;
    e[8] = c;
  }
};
// one slot's 3x3 block: FP16 (eight entries in 16 bytes and the ninth) or FP32
template <bool H>
__device__ __forceinline__ void mr_block(int64_t k, const uint4 *__restrict__ v8, const uint16_t *__restrict__ v1,
                                         const float4 *__restrict__ w8, const float *__restrict__ w1, float (&e)[9]) {
  if constexpr (H) {
    const uint4 hv = v8[k];
    const unsigned hw[4] = {hv.x, hv.y, hv.z, hv.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float2 f = __half22float2(*reinterpret_cast<const __half2 *>(&hw[q]));
      e[2 * q] = f.x;
      e[2 * q + 1] = f.y;
    }
    e[8] = __half2float(__ushort_as_half(v1[k]));
  } else {
    const float4 a4 = w8[2 * k], b4 = w8[2 * k + 1];
    e[0] = a4.x; e[1] = a4.y; e[2] = a4.z; e[3] = a4.w;
    e[4] = b4.x; e[5] = b4.y; e[6] = b4.z; e[7] = b4.w;
    e[8] = w1[k];
  }
}

template <int NR, bool H, typename CI, int MODE>
__global__ __launch_bounds__(TPB) void k_mr_ell_node(int nn, const int64_t *__restrict__ off,
                                                     const int32_t *__restrict__ width, const CI *__restrict__ col,
                                                     const uint4 *__restrict__ v8, const uint16_t *__restrict__ v1,
                                                     const float *__restrict__ dg, const float4 *__restrict__ w8,
                                                     const float *__restrict__ w1, const uint8_t *__restrict__ fixm,
                                                     const float *__restrict__ u, MrOp<float> op) {
  const int64_t blk = uftr::xcd_block(blockIdx.x, gridDim.x);
  const int64_t n = blk * TPB + threadIdx.x;
  double dv[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) dv[r] = 0.0;
  if (n < nn) {
    const int64_t w = n / UFT_ELL_SLICE;
    const int64_t base = off[w] + n % UFT_ELL_SLICE;
    const int W = width[w];
    float a[3][NR];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < NR; ++r) a[c][r] = 0.f;
#pragma unroll 2
    for (int s = 0; s < W; ++s) {
      const int64_t k = base + (int64_t)s * UFT_ELL_SLICE;
      const int64_t j = std::is_same<CI, int16_t>::value ? n + (int64_t)col[k] : (int64_t)col[k];
      float e[9];
      mr_block<H>(k, v8, v1, w8, w1, e);
      float x[3][NR];
#pragma unroll
      for (int c = 0; c < 3; ++c) ld_nr<NR>(u + (3 * j + c) * NR, x[c]);
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < NR; ++r) a[c][r] += e[3 * c] * x[0][r] + e[3 * c + 1] * x[1][r] + e[3 * c + 2] * x[2][r];
    }
    NodeRows<NR, H, MODE> rows;
    rows.load(op, n, u, dg, fixm);
    rows.finish(op, n, a, dv);
  }
  if constexpr (MODE == MR_STOREDOT) mr_node_partials<NR>(op, dv, blk);
}

// The same product on node tiles (8 right-hand sides): workgroup t takes tile
// t's UFT_TILE nodes, one per lane, in the tiles' position order.  It first
// copies the 3 x 8 vector entries of every node in the tile's halo (the
// columns of its rows, ascending) into LDS with 16-byte loads, six per node,
// so each neighbour's entries leave L2 once per tile instead of once per
// incident row; the slot loop then reads the blocks in tile order and the
// neighbour entries from LDS at the slot's 16-bit halo index.  LDS holds
// quarter q of halo node h at q HP + h (HP odd), so the lanes of one slot,
// whose halo indices mostly ascend, read distinct banks.
//
// LDS bounds the occupancy to two wavefronts per SIMD, so each wavefront
// keeps many loads in flight: the halo copy issues all its loads (node ids,
// then the quarters) before its first LDS store, the first group of TG slots
// (blocks and halo indices) is requested beside them, and the slot loop
// loads a whole group, with no branch between the loads, before its
// products.
constexpr int TILE_HP = UFT_TILE_HCAP + 1;
constexpr int TILE_QPT = (6 * UFT_TILE_HCAP + UFT_TILE - 1) / UFT_TILE;  // halo quarters per lane
template <bool H, int TG>
struct SlotGroup {
  RawBlk<H> rb[TG];
  int hh[TG];
  // slots s0 .. s0 + TG - 1 of the row; past its width W (uniform over the
  // wavefront) the last slot is read again and its block zeroed, so the
  // loads carry no branch and stay in flight together
  __device__ __forceinline__ void load(int s0, int W, unsigned base, const uint16_t *__restrict__ lc,
                                       const uint4 *__restrict__ v8, const uint16_t *__restrict__ v1,
                                       const float4 *__restrict__ w8, const float *__restrict__ w1) {
#pragma unroll
    for (int g = 0; g < TG; ++g) {
      const unsigned k = base + (unsigned)min(s0 + g, W - 1) * UFT_ELL_SLICE;
      hh[g] = __builtin_nontemporal_load(&lc[k]);
      rb[g].load(k, v8, v1, w8, w1);
    }
  }
  __device__ __forceinline__ void apply(int s0, int W, const float4 *xs, float (&a)[3][8]) const {
#pragma unroll
    for (int g = 0; g < TG; ++g) {
      float e[9];
      rb[g].get(e);
      const bool ok = s0 + g < W;
#pragma unroll
      for (int t = 0; t < 9; ++t) e[t] = ok ? e[t] : 0.f;
      float x[3][8];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const float4 v = xs[(2 * c + q) * TILE_HP + hh[g]];
          x[c][4 * q] = v.x; x[c][4 * q + 1] = v.y; x[c][4 * q + 2] = v.z; x[c][4 * q + 3] = v.w;
        }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 8; ++r)
          a[c][r] += e[3 * c] * x[0][r] + e[3 * c + 1] * x[1][r] + e[3 * c + 2] * x[2][r];
    }
  }
};
template <bool H, int MODE, int TG = (H ? 9 : 5)>
__global__ __launch_bounds__(UFT_TILE) void k_mr_tile_op(const int32_t *__restrict__ tnode,
                                                         const int32_t *__restrict__ hoff,
                                                         const int32_t *__restrict__ hnode,
                                                         const int64_t *__restrict__ toff,
                                                         const int32_t *__restrict__ twidth,
                                                         const uint16_t *__restrict__ lc,
                                                         const uint4 *__restrict__ v8, const uint16_t *__restrict__ v1,
                                                         const float *__restrict__ dg, const float4 *__restrict__ w8,
                                                         const float *__restrict__ w1, const uint8_t *__restrict__ fixm,
                                                         const float *__restrict__ u, MrOp<float> op) {
  constexpr int NR = 8;
  static_assert(UFT_TILE == TPB, "one lane per tile position");
  __shared__ float4 xs[6 * TILE_HP];
  const int64_t t = uftr::xcd_block(blockIdx.x, gridDim.x);
  const int64_t p = t * UFT_TILE + threadIdx.x;
  const int64_t n = tnode[p];
  const int64_t nv = n >= 0 ? n : 0;  // padding positions read node 0 and store nothing
  const int64_t w = p / UFT_ELL_SLICE;
  const unsigned base = (unsigned)(toff[w] + p % UFT_ELL_SLICE);  // tiled slots < 2^31 (set-up check)
  const int W = twidth[w];  // uniform over the wavefront (one slice); padding slots hold zero blocks
  // the halo copy: every lane TILE_QPT quarters (the surplus ones repeat
  // the last quarter), their node ids first, then all their loads; the first
  // slot group is requested after them, so waiting for the copy does not
  // wait for the group (loads complete in issue order)
  const int h0 = hoff[t], nq = 6 * (hoff[t + 1] - h0);
  const float4 *__restrict__ u4 = reinterpret_cast<const float4 *>(u);
  int hn[TILE_QPT], hq[TILE_QPT];
#pragma unroll
  for (int j = 0; j < TILE_QPT; ++j) {
    const int q = min((int)threadIdx.x + j * UFT_TILE, nq - 1);
    const int h = q / 6;
    hq[j] = q;
    hn[j] = hnode[h0 + h];
  }
  float4 hv[TILE_QPT];
#pragma unroll
  for (int j = 0; j < TILE_QPT; ++j) {
    const int h = hq[j] / 6, part = hq[j] - 6 * h;
    hv[j] = u4[(int64_t)hn[j] * 6 + part];
    hq[j] = part * TILE_HP + h;
  }
  SlotGroup<H, TG> ga;
  if (W > 0) ga.load(0, W, base, lc, v8, v1, w8, w1);
#pragma unroll
  for (int j = 0; j < TILE_QPT; ++j) xs[hq[j]] = hv[j];
  __syncthreads();
  double dv[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) dv[r] = 0.0;
  float a[3][NR];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < NR; ++r) a[c][r] = 0.f;
  for (int s0 = 0; s0 < W; s0 += TG) {
    if (s0 > 0) ga.load(s0, W, base, lc, v8, v1, w8, w1);
    ga.apply(s0, W, xs, a);
  }
  NodeRows<NR, H, MODE> rows;
  rows.load(op, nv, u, dg, fixm);
  if (n >= 0) rows.finish(op, n, a, dv);
  if constexpr (MODE == MR_STOREDOT) mr_node_partials<NR>(op, dv, t);
}

// ---------------------------------------------------------------- 6x6 BSR
// One wavefront per block row: lane (slot l, right-hand side rr), 64 / NR
// slots; a lane reads whole 6x6 blocks (the NR lanes of a slot share each
// load) and its right-hand side's six entries of u, accumulating all six
// rows; a fixed butterfly over the slots combines them, and the slot-0 lanes
// run the six rows' epilogues.
template <typename T, int NR>
__global__ __launch_bounds__(TPB) void k_mr_bsr_op(int64_t nbrows, const int32_t *__restrict__ rp,
                                                   const int32_t *__restrict__ ci, const T *__restrict__ val,
                                                   const T *__restrict__ u, MrOp<T> op) {
  constexpr int SLOTS = 64 / NR;
  typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type V16;
  constexpr int PER = 16 / (int)sizeof(T);
  const int64_t blk = uftr::xcd_block(blockIdx.x, gridDim.x);
  const int64_t t = blk * TPB + threadIdx.x;
  const int64_t A = t / 64;
  const int lane = (int)(t % 64), l = lane / NR, rr = lane % NR;
  T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  if (A < nbrows) {
    for (int k = rp[A] + l; k < rp[A + 1]; k += SLOTS) {
      const int64_t B = ci[k];
      const T *xb = u + (B * 6) * NR + rr;
      T x[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) x[j] = xb[j * NR];
      T v[36];  // the block in 16-byte loads
      const V16 *vp = reinterpret_cast<const V16 *>(val + (size_t)k * 36);
#pragma unroll
      for (int q = 0; q < 36 / PER; ++q) {
        const V16 w = vp[q];
        const T *ww = reinterpret_cast<const T *>(&w);
#pragma unroll
        for (int e = 0; e < PER; ++e) v[q * PER + e] = ww[e];
      }
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) acc[i] += v[i * 6 + j] * x[j];
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
#pragma unroll
    for (int off = 32; off >= NR; off >>= 1) acc[i] += __shfl_xor(acc[i], off, 64);
  }
  double dv[3] = {0.0, 0.0, 0.0};
  if (l == 0 && A < nbrows) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int64_t row = A * 6 + i;
      mr_epi(op, row, row * NR + rr, acc[i], u[row * NR + rr], dv);
    }
  }
  mr_epi_partials<T, NR>(op, dv, blk);
}

// The same product with the FP16 blocks of the int8-smoother form: the level
// stored as H = D^-1/2 A D^-1/2 in FP16 with dh = D^1/2 per row in FP32, so
// (A u)_i = dh_i sum_j H_ij dh_j u_j; 72 bytes per block instead of 144.
// (Taking two slot rounds per lane at once, with every load of both issued
// first, measured slower on c4a's level 1: 38 against 30 us, the extra
// registers cost more waves than the overlap gains.)
template <int NR>
__global__ __launch_bounds__(TPB) void k_mr_bsr_op_h(int64_t nbrows, const int32_t *__restrict__ rp,
                                                     const int32_t *__restrict__ ci, const uint2 *__restrict__ valh,
                                                     const float *__restrict__ dh, const float *__restrict__ u,
                                                     MrOp<float> op) {
  constexpr int SLOTS = 64 / NR;
  const int64_t blk = uftr::xcd_block(blockIdx.x, gridDim.x);
  const int64_t t = blk * TPB + threadIdx.x;
  const int64_t A = t / 64;
  const int lane = (int)(t % 64), l = lane / NR, rr = lane % NR;
  float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (A < nbrows) {
    for (int k = rp[A] + l; k < rp[A + 1]; k += SLOTS) {
      const int64_t B = ci[k];
      const float *xb = u + (B * 6) * NR + rr;
      const float2 *db = reinterpret_cast<const float2 *>(dh + B * 6);
      float x[6];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float2 d = db[j];
        x[2 * j] = xb[(2 * j) * NR] * d.x;
        x[2 * j + 1] = xb[(2 * j + 1) * NR] * d.y;
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
          acc[e / 6] += f.x * x[e % 6] + f.y * x[e % 6 + 1];
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
#pragma unroll
    for (int off = 32; off >= NR; off >>= 1) acc[i] += __shfl_xor(acc[i], off, 64);
  }
  double dv[3] = {0.0, 0.0, 0.0};
  if (l == 0 && A < nbrows) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int64_t row = A * 6 + i;
      mr_epi(op, row, row * NR + rr, acc[i] * dh[row], u[row * NR + rr], dv);
    }
  }
  mr_epi_partials<float, NR>(op, dv, blk);
}

// The same FP16 level with 8 right-hand sides, a block row per wavefront
// mapped as 32 slots x 2 halves of the cases: lane (l, h) takes blocks
// l, l + 32, ... of the row for cases 4h .. 4h + 3 (each neighbour's four
// entries of a row in one 16-byte load), so a row of a few dozen blocks
// takes one or two rounds instead of four or five dependent ones.  A fixed
// butterfly over the slots combines the sums; lanes 0 and 1 run the rows'
// epilogues and the per-case partials (part[blk][k * 8 + r], the layout of
// the other kernels).
template <int MODE>
__global__ __launch_bounds__(TPB) void k_mr_bsr_op_h8(int64_t nbrows, const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ ci, const uint2 *__restrict__ valh,
                                                      const float *__restrict__ dh, const float *__restrict__ u,
                                                      MrOp<float> op) {
  constexpr int NR = 8;
  const int64_t blk = uftr::xcd_block(blockIdx.x, gridDim.x);
  const int64_t t = blk * TPB + threadIdx.x;
  const int64_t A = t / 64;
  const int lane = (int)(t % 64), l = lane >> 1, h = lane & 1;
  float acc[6][4];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[i][q] = 0.f;
  if (A < nbrows) {
    for (int k = rp[A] + l; k < rp[A + 1]; k += 32) {
      const int64_t B = ci[k];
      const float2 *db = reinterpret_cast<const float2 *>(dh + B * 6);
      const float4 *xb = reinterpret_cast<const float4 *>(u + (B * 6) * NR + 4 * h);
      float x[6][4];
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const float4 v = xb[j * 2];  // unknown 6B + j, cases 4h .. 4h + 3
        x[j][0] = v.x; x[j][1] = v.y; x[j][2] = v.z; x[j][3] = v.w;
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float2 d = db[j];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          x[2 * j][q] *= d.x;
          x[2 * j + 1][q] *= d.y;
        }
      }
      const uint2 *vb = valh + (size_t)k * 9;
#pragma unroll
      for (int w = 0; w < 9; ++w) {
        const uint2 hv = vb[w];
        const unsigned hw[2] = {hv.x, hv.y};
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const float2 f = __half22float2(*reinterpret_cast<const __half2 *>(&hw[p]));
          const int e = 4 * w + 2 * p;  // entries e, e+1 of the row-major block (same row)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[e / 6][q] += f.x * x[e % 6][q] + f.y * x[e % 6 + 1][q];
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int off = 32; off >= 2; off >>= 1) acc[i][q] += __shfl_xor(acc[i][q], off, 64);
    }
  // every lane of a half now holds the row sums; lane (l < 24, h) runs the
  // epilogue of row 6A + l / 4, case 4h + l % 4
  double dv[3] = {0.0, 0.0, 0.0};
  if (l < 24 && A < nbrows) {
    const int i = l >> 2, q = l & 3;
    float ti = 0.f;
#pragma unroll
    for (int ii = 0; ii < 6; ++ii)
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) ti = (ii == i && qq == q) ? acc[ii][qq] : ti;
    const int64_t row = A * 6 + i, jj = row * NR + 4 * h + q;
    mr_epi(op, row, jj, ti * dh[row], u[jj], dv);
  }
  if (op.mode == MR_DOTS2 || op.mode == MR_DOTS3 || op.mode == MR_STOREDOT) {
    // per case r = 4h + q: the six rows' lanes (2 (4i + q) + h) of each
    // wavefront, then the four wavefronts, in a fixed order
    __shared__ double sp[3][TPB];
#pragma unroll
    for (int k = 0; k < 3; ++k) sp[k][threadIdx.x] = dv[k];
    __syncthreads();
    const int K = op.mode == MR_DOTS3 ? 3 : (op.mode == MR_DOTS2 ? 2 : 1);
    if (threadIdx.x < K * NR) {
      const int k = threadIdx.x / NR, r = threadIdx.x % NR, hh = r >> 2, q = r & 3;
      double tot = 0.0;
      for (int w = 0; w < TPB / 64; ++w) {
        double sw = 0.0;
        for (int i = 0; i < 6; ++i) sw += sp[k][64 * w + 2 * (4 * i + q) + hh];
        tot += sw;
      }
      op.part[(size_t)blk * K * NR + threadIdx.x] = tot;
    }
  }
}

// ---------------------------------------------------------------- transfers
// wavefront per coarse unit; lane = slot * NR + rr
template <int BS, typename T, int NR>
__global__ __launch_bounds__(TPB) void k_mr_restrict_first(int64_t nc, const int32_t *__restrict__ cptr,
                                                           const int32_t *__restrict__ child, const T *__restrict__ P,
                                                           const T *__restrict__ r, T *__restrict__ bc,
                                                           const T *__restrict__ dinv, double it, T *__restrict__ rc,
                                                           T *__restrict__ dc, T *__restrict__ xc) {
  constexpr int SL = 64 / NR;
  const int64_t A = ((int64_t)blockIdx.x * TPB + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63, slot = lane / NR, rr = lane % NR;
  if (A >= nc) return;  // whole wavefront
  const int k0 = cptr[A], npair = (cptr[A + 1] - k0) * BS;
  T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int q = slot; q < npair; q += SL) {
    const int kk = q / BS, i = q - kk * BS;
    const int64_t uu = child[k0 + kk];
    const T *Pu = P + (size_t)uu * BS * 6 + i * 6;
    const T ri = r[(uu * BS + i) * NR + rr];
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[j] += Pu[j] * ri;
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
#pragma unroll
    for (int off = 32; off >= NR; off >>= 1) acc[j] += __shfl_down(acc[j], off, 64);
  }
  if (slot == 0) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int64_t i = A * 6 + j, ii = i * NR + rr;
      bc[ii] = acc[j];
      if (dinv) {
        const T d = (T)it * dinv[i] * acc[j];
        rc[ii] = acc[j];
        dc[ii] = d;
        xc[ii] = d;
      }
    }
  }
}

// x_u += P_u (ca_r xa + cb_r xb)_{parent(u)}; ca_r = sc[r * 8 + ia] (ia < 0: 1)
template <int BS, typename T, int NR>
__global__ __launch_bounds__(TPB) void k_mr_prolong_comb(int64_t nu, const int32_t *__restrict__ parent,
                                                         const T *__restrict__ P, const double *__restrict__ sc, int ia,
                                                         int ib, const T *__restrict__ xa, const T *__restrict__ xb,
                                                         T *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= nu * BS * NR) return;
  const int64_t row = t / NR;
  const int rr = (int)(t % NR);
  const int64_t uu = row / BS;
  const int i = (int)(row - uu * BS);
  const int64_t A = parent[uu];
  if (A < 0) return;
  const T ca = ia < 0 ? T(1) : (T)sc[rr * 8 + ia];
  const T cb = ib < 0 ? T(1) : (T)sc[rr * 8 + ib];
  const T *Pu = P + (size_t)uu * BS * 6 + i * 6;
  T acc = T(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int64_t ii = (A * 6 + j) * NR + rr;
    const T c = xb ? ca * xa[ii] + cb * xb[ii] : xa[ii];
    acc += Pu[j] * c;
  }
  x[t] += acc;
}

template <typename T, int NR>
__global__ __launch_bounds__(TPB) void k_mr_lincomb_first(int64_t n, const double *__restrict__ sc,
                                                          const T *__restrict__ b, const T *__restrict__ v,
                                                          T *__restrict__ kr, T *__restrict__ r,
                                                          const T *__restrict__ dinv, double it, T *__restrict__ d,
                                                          T *__restrict__ kd) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= n * NR) return;
  const int rr = (int)(t % NR);
  const T ri = b[t] + (T)sc[rr * 8 + 7] * v[t];
  kr[t] = ri;
  r[t] = ri;
  const T di = (T)it * dinv[t / NR] * ri;
  d[t] = di;
  kd[t] = di;
}

// per-RHS K-cycle coefficients from the K * NR partial columns (same
// recurrences as the single right-hand-side K-cycle)
template <int K, int NR>
__global__ __launch_bounds__(TPB) void k_mr_final_kc(const double *__restrict__ part, int nb, double *__restrict__ sc) {
  double res[K * NR];
  uftr::block_final<K * NR>(part, nb, res);
  if (threadIdx.x >= NR) return;
  const int rr = threadIdx.x;
  double *s = sc + rr * 8;
  if constexpr (K == 2) {
    s[0] = res[rr];
    s[1] = res[NR + rr];
    s[7] = s[0] != 0.0 ? -s[1] / s[0] : 0.0;
  } else {
    const double rho1 = s[0], a1 = s[1], g = res[rr], be = res[NR + rr], a2 = res[2 * NR + rr];
    s[2] = g;
    s[3] = be;
    s[4] = a2;
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
}

// first stage of a long partial sum: block b sums rows [b chunk, (b+1) chunk)
// of the KN partial columns in the fixed block order into out[b]
template <int KN>
__global__ __launch_bounds__(TPB) void k_part_stage(const double *__restrict__ part, int nb, int chunk,
                                                    double *__restrict__ out) {
  const int r0 = blockIdx.x * chunk, r1 = min(nb, r0 + chunk);
  double res[KN];
  uftr::block_final<KN>(part + (size_t)r0 * KN, max(0, r1 - r0), res);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < KN; ++k) out[(size_t)blockIdx.x * KN + k] = res[k];
  }
}
constexpr int MR_STAGE_BLOCKS = 128;

// xc[row][r] = sum_c Ainv[row][c] bc[c][r] with the coarsest level's dense
// inverse in the cycle's precision T (row-major, n even): one wavefront per
// MR_DM_ROWS rows, lane l takes the column pairs (FP32) or columns (FP64)
// l, l + 64, ...; the interleaved right-hand-side entries of those columns
// are one contiguous run (16-byte loads) shared by the wavefront's rows, so
// the inverse streams once and bc is read once per MR_DM_ROWS rows; the
// lanes' partial sums are combined by a fixed butterfly.
constexpr int MR_DM_ROWS = 4;
template <typename T, int NR>
__global__ __launch_bounds__(TPB) void k_mr_dense_mv(int n, const T *__restrict__ A, const T *__restrict__ bc,
                                                     T *__restrict__ xc) {
  constexpr int VW = sizeof(T) == 4 ? 2 : 1;           // columns per lane and step
  constexpr int NV = VW * NR * (int)sizeof(T) / 16;    // 16-byte loads of bc per step
  typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type V16;
  constexpr int PER = 16 / (int)sizeof(T);
  const int64_t wave = ((int64_t)blockIdx.x * TPB + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t r0 = wave * MR_DM_ROWS;
  if (r0 >= n) return;  // whole wavefront
  T acc[MR_DM_ROWS][NR];
#pragma unroll
  for (int r = 0; r < MR_DM_ROWS; ++r)
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[r][q] = T(0);
#pragma unroll 4
  for (int64_t c = (int64_t)lane * VW; c < n; c += 64 * VW) {
    T xv[VW * NR];
    const V16 *xp = reinterpret_cast<const V16 *>(bc + c * NR);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const V16 t = xp[v];
      const T *tt = reinterpret_cast<const T *>(&t);
#pragma unroll
      for (int e = 0; e < PER; ++e) xv[v * PER + e] = tt[e];
    }
#pragma unroll
    for (int r = 0; r < MR_DM_ROWS; ++r) {
      if (r0 + r >= n) break;
      T av[VW];
      if constexpr (VW == 2) {
        const float2 a2 = *reinterpret_cast<const float2 *>(A + (r0 + r) * (int64_t)n + c);
        av[0] = a2.x;
        av[1] = a2.y;
      } else {
        av[0] = A[(r0 + r) * (int64_t)n + c];
      }
#pragma unroll
      for (int v = 0; v < VW; ++v)
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[r][q] += av[v] * xv[v * NR + q];
    }
  }
#pragma unroll
  for (int r = 0; r < MR_DM_ROWS; ++r)
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      T v = acc[r][q];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == r * NR + q && r0 + r < n) xc[(r0 + r) * NR + q] = v;
    }
}

// ---------------------------------------------------------------- outer Krylov (FP64)
// The outer-loop reductions run on a fixed grid (MR_RED_BLOCKS) with a
// grid-stride loop, so the final one-block sum reads few partials; the
// stride is a multiple of NR, so every thread keeps its right-hand side.
constexpr int MR_RED_BLOCKS = 1024;
inline unsigned mr_red_grid(int64_t n) { return (unsigned)std::min<int64_t>(MR_RED_BLOCKS, nblk(n)); }
template <int NR>
__global__ __launch_bounds__(TPB) void k_mr_dot(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                                                double *__restrict__ part) {
  double v[1] = {0.0};
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < n * NR; t += (int64_t)gridDim.x * TPB)
    v[0] += a[t] * b[t];
  mr_block_partials<1, NR>(v, part);
}
template <int NR>
__global__ __launch_bounds__(TPB) void k_mr_dot_pr(int64_t n, const double *__restrict__ r, const double *__restrict__ z,
                                                   const double *__restrict__ zo, double *__restrict__ part) {
  double v[2] = {0.0, 0.0};
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < n * NR; t += (int64_t)gridDim.x * TPB) {
    v[0] += r[t] * z[t];
    v[1] += r[t] * zo[t];
  }
  mr_block_partials<2, NR>(v, part);
}
// x += alpha_r p ; r -= alpha_r q ; partial ||r||^2 per RHS
template <int NR>
__global__ __launch_bounds__(TPB) void k_mr_cg_update(int64_t n, const double *__restrict__ alpha,
                                                      double *__restrict__ x, const double *__restrict__ p,
                                                      double *__restrict__ r, const double *__restrict__ q,
                                                      double *__restrict__ part) {
  double v[1] = {0.0};
  const double al = alpha[((int64_t)blockIdx.x * TPB + threadIdx.x) % NR];
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < n * NR; t += (int64_t)gridDim.x * TPB) {
    x[t] += al * p[t];
    const double ri = r[t] - al * q[t];
    r[t] = ri;
    v[0] += ri * ri;
  }
  mr_block_partials<1, NR>(v, part);
}
// mixed forms: the same update with the outer product q in FP32; the new
// residual's FP32 copy rf (the cycle's input) is written in the same pass
template <int NR>
__global__ __launch_bounds__(TPB) void k_mr_cg_update_lo(int64_t n, const double *__restrict__ alpha,
                                                         double *__restrict__ x, const double *__restrict__ p,
                                                         double *__restrict__ r, const float *__restrict__ q,
                                                         float *__restrict__ rf, double *__restrict__ part) {
  double v[1] = {0.0};
  const double al = alpha[((int64_t)blockIdx.x * TPB + threadIdx.x) % NR];
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < n * NR; t += (int64_t)gridDim.x * TPB) {
    x[t] += al * p[t];
    const double ri = r[t] - al * (double)q[t];
    r[t] = ri;
    rf[t] = (float)ri;
    v[0] += ri * ri;
  }
  mr_block_partials<1, NR>(v, part);
}
// reliable update of the right-hand sides in mask: r = res (the FP64
// residual F - K U) and its FP32 copy (rf may be null); partials of r.r for
// every right-hand side
template <int NR>
__global__ __launch_bounds__(TPB) void k_mr_take_res(int64_t n, unsigned mask, const double *__restrict__ res,
                                                     double *__restrict__ r, float *__restrict__ rf,
                                                     double *__restrict__ part) {
  double v[1] = {0.0};
  const int k = (int)(((int64_t)blockIdx.x * TPB + threadIdx.x) % NR);
  const bool take = (mask >> k) & 1u;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < n * NR; t += (int64_t)gridDim.x * TPB) {
    double ri;
    if (take) {
      ri = res[t];
      r[t] = ri;
      if (rf) rf[t] = (float)ri;
    } else {
      ri = r[t];
    }
    v[0] += ri * ri;
  }
  mr_block_partials<1, NR>(v, part);
}
// p = z + beta_r p and its FP32 copy (the mixed outer operator's input)
template <int NR>
__global__ __launch_bounds__(TPB) void k_mr_xpby_lo(int64_t n, const double *__restrict__ z,
                                                    const double *__restrict__ beta, double *__restrict__ p,
                                                    float *__restrict__ pf) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t < n * NR) {
    const double v = z[t] + beta[t % NR] * p[t];
    p[t] = v;
    pf[t] = (float)v;
  }
}
template <int NR>
__global__ __launch_bounds__(TPB) void k_mr_xpby(int64_t n, const double *__restrict__ z, const double *__restrict__ beta,
                                                 double *__restrict__ p) {
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t < n * NR) p[t] = z[t] + beta[t % NR] * p[t];
}
// per-RHS scalars: alpha = rz / pq on the active right-hand sides, else 0
__global__ void k_mr_alpha(int nr, unsigned mask, const double *__restrict__ rz, const double *__restrict__ pq,
                           double *__restrict__ alpha) {
  const int r = threadIdx.x;
  if (r < nr) alpha[r] = ((mask >> r) & 1u) && pq[r] != 0.0 ? rz[r] / pq[r] : 0.0;
}
// Polak-Ribiere: beta = (rz - rzo) / rzprev ; rzprev = rz (inactive: 0)
__global__ void k_mr_beta(int nr, unsigned mask, const double *__restrict__ rzn, double *__restrict__ rzprev,
                          double *__restrict__ beta) {
  const int r = threadIdx.x;
  if (r >= nr) return;
  const double rz = rzn[r], rzo = rzn[nr + r];
  beta[r] = ((mask >> r) & 1u) && rzprev[r] != 0.0 ? (rz - rzo) / rzprev[r] : 0.0;
  rzprev[r] = rz;
}

// ---------------------------------------------------------------- sensitivities
// lane (e, rr): uKu of right-hand side rr; the NR values of an element are
// summed over its lanes in a fixed order; fx partial, dfdx.
template <int NPE, int NR>
__global__ __launch_bounds__(TPB) void k_mr_sens(int ne, const int32_t *__restrict__ elem,
                                                 const double2 *__restrict__ KE, const double *__restrict__ xphys,
                                                 SimpCoef sc, const double *__restrict__ u, double *__restrict__ dfdx,
                                                 double *__restrict__ part) {
  constexpr int ND = 3 * NPE, NP = ND * (ND + 1) / 2, NQ = NP / 2;
  const int64_t t = uftr::xcd_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
  const int64_t e = t / NR;
  const int rr = (int)(t % NR);
  double uku = 0.0;
  if (e < ne) {
    double ue[ND];
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
      const int n = elem[(size_t)a * ne + e];
#pragma unroll
      for (int c = 0; c < 3; ++c) ue[c * NPE + a] = u[(3 * (size_t)n + c) * NR + rr];
    }
    double dg = 0.0, off = 0.0;
    static_for<0, NQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const double2 kk = KE[(size_t)q * ne + e];
      constexpr int r0 = pk_row(2 * q, ND), c0 = pk_col(2 * q, ND);
      constexpr int r1 = pk_row(2 * q + 1, ND), c1 = pk_col(2 * q + 1, ND);
      if constexpr (r0 == c0) dg += kk.x * ue[r0] * ue[c0]; else off += kk.x * ue[r0] * ue[c0];
      if constexpr (r1 == c1) dg += kk.y * ue[r1] * ue[c1]; else off += kk.y * ue[r1] * ue[c1];
    });
    uku = dg + 2.0 * off;
  }
  // sum the element's NR lanes (adjacent lanes, fixed tree)
#pragma unroll
  for (int off = NR / 2; off > 0; off >>= 1) uku += __shfl_down(uku, off, NR);
  double v[1] = {0.0};
  if (e < ne && rr == 0) {
    const double x = xphys[e];
    v[0] = simpE(sc, x) * uku;
    dfdx[e] = -1.0 * sc.penal * pow(x, sc.penal - 1.0) * sc.dE * uku;
  }
  uftr::block_partials<1>(v, part);
}

// case k of an interleaved vector <-> a plain vector
__global__ __launch_bounds__(TPB) void k_mr_extract(int64_t n, int nr, int k, const double *__restrict__ src,
                                                    double *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) dst[i] = src[i * nr + k];
}
// F[idx * nr + k] = val (INSERT of load case k)
__global__ __launch_bounds__(TPB) void k_mr_set_dofs(int n, int nr, int k, const int32_t *__restrict__ idx,
                                                     const double *__restrict__ val, double *__restrict__ F) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i < n) F[(size_t)idx[i] * nr + k] = val ? val[i] : 0.0;
}

}  // namespace

// ---------------------------------------------------------------- host wrappers
namespace uftk {

template <typename T>
static MrOp<T> to_mr(const RowOp<T> &o) {
  MrOp<T> m;
  m.mode = o.mode == ROW_STORE ? MR_STORE
         : o.mode == ROW_STEP ? MR_STEP
         : o.mode == ROW_RES ? MR_RES
         : o.mode == ROW_RESFIRST ? MR_RESFIRST
         : o.mode == ROW_RESFIRSTX ? MR_RESFIRSTX
         : o.mode == ROW_DOTS2 ? MR_DOTS2
         : o.mode == ROW_DOTS3 ? MR_DOTS3
                               : MR_STOREDOT;
  m.addprev = o.addprev;
  m.y = o.y; m.r = o.r; m.dout = o.dout; m.x = o.x; m.kr = o.kr; m.kd = o.kd;
  m.base = o.base; m.dinv = o.dinv; m.w1 = o.w1; m.w2 = o.w2;
  m.c1 = o.c1; m.c2 = o.c2; m.it = o.it; m.part = o.part;
  return m;
}

#define UFT_NR_SWITCH(nr, CALL) \
  switch (nr) {                 \
    case 2: { constexpr int NR = 2; CALL; } break; \
    case 4: { constexpr int NR = 4; CALL; } break; \
    default: { constexpr int NR = 8; CALL; } break; \
  }

template <typename T>
static void mr_cheb_first_t(hipStream_t s, int nr, int64_t n, const T *b, const T *dinv, double it, T *r, T *d, T *x) {
  UFT_NR_SWITCH(nr, (k_mr_cheb_first<T, NR><<<nblk(n * NR), TPB, 0, s>>>(n, b, dinv, it, r, d, x)));
}
void mr_cheb_first(hipStream_t s, int nr, int64_t n, const double *b, const double *dinv, double it, double *r,
                   double *d, double *x) { mr_cheb_first_t<double>(s, nr, n, b, dinv, it, r, d, x); }
void mr_cheb_first(hipStream_t s, int nr, int64_t n, const float *b, const float *dinv, double it, float *r, float *d,
                   float *x) { mr_cheb_first_t<float>(s, nr, n, b, dinv, it, r, d, x); }

void mr_elem_phase(hipStream_t s, const uft_ctx *C, int nr, const double *xphys, const double *u, double *Ye) {
  const SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
  const int ne = (int)C->ne;
  const double2 *KE = (const double2 *)C->d_KE;
  if (C->npe == 8) {
    UFT_NR_SWITCH(nr, (k_mr_elem<8, double, double2, NR><<<nblk((int64_t)ne * NR), TPB, 0, s>>>(
                          ne, C->d_elem, KE, xphys, sc, C->d_fixm, u, Ye)));
  } else {
    UFT_NR_SWITCH(nr, (k_mr_elem<4, double, double2, NR><<<nblk((int64_t)ne * NR), TPB, 0, s>>>(
                          ne, C->d_elem, KE, xphys, sc, C->d_fixm, u, Ye)));
  }
}
void mr_elem_phase(hipStream_t s, const uft_ctx *C, int nr, const double *xphys, const float *u, float *Ye,
                   bool blocks32) {
  const SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
  const int ne = (int)C->ne;
  if (!blocks32 && nr == 8 && C->d_KEb) {  // matrix-core form
    const unsigned g = (unsigned)((ne + MF_EL * MF_WAVES - 1) / (MF_EL * MF_WAVES));
    if (C->npe == 8)
      k_mr_elem_mfma<8><<<g, 64 * MF_WAVES, 0, s>>>(ne, C->d_elem, (const uint4 *)C->d_KEb, C->d_Ee, C->d_fixm, u,
                                                    Ye);
    else
      k_mr_elem_mfma<4><<<g, 64 * MF_WAVES, 0, s>>>(ne, C->d_elem, (const uint4 *)C->d_KEb, C->d_Ee, C->d_fixm, u,
                                                    Ye);
    return;
  }
  const float2 *KE = (const float2 *)C->d_KEf;
  if (C->npe == 8) {
    UFT_NR_SWITCH(nr, (k_mr_elem<8, float, float2, NR><<<nblk((int64_t)ne * NR), TPB, 0, s>>>(
                          ne, C->d_elem, KE, xphys, sc, C->d_fixm, u, Ye)));
  } else {
    UFT_NR_SWITCH(nr, (k_mr_elem<4, float, float2, NR><<<nblk((int64_t)ne * NR), TPB, 0, s>>>(
                          ne, C->d_elem, KE, xphys, sc, C->d_fixm, u, Ye)));
  }
}
void ke_to_bf16(hipStream_t s, const uft_ctx *C) {
  const int ne = (int)C->ne;
  if (C->npe == 8)
    k_ke_to_bf16_rows<8><<<nblk((int64_t)ne * 24), TPB, 0, s>>>(ne, (const double2 *)C->d_KE, (__bf16 *)C->d_KEb);
  else
    k_ke_to_bf16_rows<4><<<nblk((int64_t)ne * 12), TPB, 0, s>>>(ne, (const double2 *)C->d_KE, (__bf16 *)C->d_KEb);
}
int ke_bf16_rows_words(int npe) {  // 32-bit words per element of the row layout
  const int nd = 3 * npe, ndp = (nd + 7) / 8 * 8;
  return nd * ndp / 2;
}

template <typename T>
static void mr_gather_op_t(hipStream_t s, const uft_ctx *C, int nr, const T *u, const T *Ye, const RowOp<T> &o) {
  const MrOp<T> op = to_mr(o);
  const int nn = (int)C->nn, ne = (int)C->ne;
  if (C->npe == 8) {
    UFT_NR_SWITCH(nr, (k_mr_gather_op<8, T, NR><<<nblk((int64_t)nn * NR), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, Ye,
                                                                                     C->d_fixm, u, op)));
  } else {
    UFT_NR_SWITCH(nr, (k_mr_gather_op<4, T, NR><<<nblk((int64_t)nn * NR), TPB, 0, s>>>(nn, ne, C->d_iptr, C->d_inc, Ye,
                                                                                     C->d_fixm, u, op)));
  }
}
void mr_gather_op(hipStream_t s, const uft_ctx *C, int nr, const double *u, const double *Ye, const RowOp<double> &o) {
  mr_gather_op_t<double>(s, C, nr, u, Ye, o);
}
void mr_gather_op(hipStream_t s, const uft_ctx *C, int nr, const float *u, const float *Ye, const RowOp<float> &o) {
  mr_gather_op_t<float>(s, C, nr, u, Ye, o);
}
int mr_gather_blocks(int64_t nn, int nr) { return (int)nblk(nn * nr); }
static bool mr_ell_node_mode(int mode, bool h) {
  return mode != MR_DOTS2 && mode != MR_DOTS3 && !(h && mode == MR_STEP);
}
template <bool H, typename CI>
static void mr_ell_node_t(hipStream_t s, const uft_ctx *C, int nr, const CI *col, const float *u, const MrOp<float> &op) {
  const auto &F = C->fell;
  const int nn = (int)C->nn;
#define UFT_MR_NODE(M)                                                                                              \
  UFT_NR_SWITCH(nr, (k_mr_ell_node<NR, H, CI, M><<<nblk(nn), TPB, 0, s>>>(nn, F.d_off, F.d_width, col, F.d_v8, \
                                                                        F.d_v1, F.d_dg, F.d_w8, F.d_w1, C->d_fixm, u, op)))
  switch (op.mode) {
    case MR_STORE: UFT_MR_NODE(MR_STORE); break;
    case MR_STEP: UFT_MR_NODE(MR_STEP); break;
    case MR_RES: UFT_MR_NODE(MR_RES); break;
    case MR_RESFIRST: UFT_MR_NODE(MR_RESFIRST); break;
    case MR_RESFIRSTX: UFT_MR_NODE(MR_RESFIRSTX); break;
    default: UFT_MR_NODE(MR_STOREDOT); break;
  }
#undef UFT_MR_NODE
}
// the node-tile kernel: eight right-hand sides, every epilogue but the
// K-cycle's dot modes (not used on the fine level)
static bool mr_tile_mode(const uft_ctx *C, int nr, int mode, const float *u) {
  return C->fell.d_tnode && nr == 8 && mode != MR_DOTS2 && mode != MR_DOTS3 && ((uintptr_t)u % 16) == 0;
}
template <bool H>
static void mr_tile_op_t(hipStream_t s, const uft_ctx *C, const float *u, const MrOp<float> &op) {
  const auto &F = C->fell;
#define UFT_MR_TILE(M)                                                                                          \
  k_mr_tile_op<H, M><<<(unsigned)F.ntile, UFT_TILE, 0, s>>>(F.d_tnode, F.d_thoff, F.d_thnode, F.d_toff,       \
                                                           F.d_twidth, F.d_tlc, F.d_tv8, F.d_tv1, F.d_dg,    \
                                                           F.d_tw8, F.d_tw1, C->d_fixm, u, op)
  switch (op.mode) {
    case MR_STORE: UFT_MR_TILE(MR_STORE); break;
    case MR_STEP: UFT_MR_TILE(MR_STEP); break;
    case MR_RES: UFT_MR_TILE(MR_RES); break;
    case MR_RESFIRST: UFT_MR_TILE(MR_RESFIRST); break;
    case MR_RESFIRSTX: UFT_MR_TILE(MR_RESFIRSTX); break;
    default: UFT_MR_TILE(MR_STOREDOT); break;
  }
#undef UFT_MR_TILE
}
template <bool H>
static void mr_ell_op_t(hipStream_t s, const uft_ctx *C, int nr, const float *u, const RowOp<float> &o) {
  const auto &F = C->fell;
  const int nn = (int)C->nn;
  const MrOp<float> op = to_mr(o);
  if (mr_tile_mode(C, nr, op.mode, u)) {
    mr_tile_op_t<H>(s, C, u, op);
    return;
  }
  // the node-lane kernel: every epilogue but the K-cycle's dot modes (not used
  // on the fine level) and the FP16 blocks' Chebyshev step, vector entries
  // 8- or 16-byte aligned
  if (mr_ell_node_mode(op.mode, H) && ((uintptr_t)u % 16) == 0) {
    if (F.d_cd)
      mr_ell_node_t<H, int16_t>(s, C, nr, F.d_cd, u, op);
    else
      mr_ell_node_t<H, int32_t>(s, C, nr, F.d_col, u, op);
    return;
  }
  if (F.d_cd) {
    UFT_NR_SWITCH(nr, (k_mr_ell_op<NR, H, int16_t><<<nblk((int64_t)nn * NR), TPB, 0, s>>>(
                          nn, F.d_off, F.d_width, F.d_cd, F.d_v8, F.d_v1, F.d_dg, F.d_w8, F.d_w1, C->d_fixm, u, op)));
  } else {
    UFT_NR_SWITCH(nr, (k_mr_ell_op<NR, H, int32_t><<<nblk((int64_t)nn * NR), TPB, 0, s>>>(
                          nn, F.d_off, F.d_width, F.d_col, F.d_v8, F.d_v1, F.d_dg, F.d_w8, F.d_w1, C->d_fixm, u, op)));
  }
}
void mr_ell_op64(hipStream_t s, const uft_ctx *C, int nr, const double *u, const RowOp<double> &o) {
  const auto &F = C->fell;
  const int nn = (int)C->nn;
  const MrOp<double> op = to_mr(o);
  if (F.d_cd) {
    UFT_NR_SWITCH(nr, (k_mr_ell_op64<NR, int16_t><<<nblk((int64_t)nn * NR), TPB, 0, s>>>(
                          nn, F.nslot, F.d_off, F.d_width, F.d_cd, F.d_e8, F.d_e1, C->d_fixm, u, op)));
  } else {
    UFT_NR_SWITCH(nr, (k_mr_ell_op64<NR, int32_t><<<nblk((int64_t)nn * NR), TPB, 0, s>>>(
                          nn, F.nslot, F.d_off, F.d_width, F.d_col, F.d_e8, F.d_e1, C->d_fixm, u, op)));
  }
}
// blocks (and STOREDOT partials per right-hand side) of mr_ell_op's launch
int mr_ell_blocks(const uft_ctx *C, const float *u, int mode, bool fp32) {
  RowOp<float> o;
  o.mode = mode;
  if (mr_tile_mode(C, C->mr.nr, to_mr(o).mode, u)) return (int)C->fell.ntile;
  const bool node = mr_ell_node_mode(to_mr(o).mode, !fp32) && ((uintptr_t)u % 16) == 0;
  return node ? (int)nblk(C->nn) : (int)nblk((int64_t)C->nn * C->mr.nr);
}
void mr_ell_op(hipStream_t s, const uft_ctx *C, int nr, const float *u, const RowOp<float> &o, bool fp32) {
  if (fp32)
    mr_ell_op_t<false>(s, C, nr, u, o);
  else
    mr_ell_op_t<true>(s, C, nr, u, o);
}

template <typename T>
static void mr_bsr_op_t(hipStream_t s, int nr, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const T *val,
                        const T *u, const RowOp<T> &o) {
  const MrOp<T> op = to_mr(o);
  UFT_NR_SWITCH(nr, (k_mr_bsr_op<T, NR><<<nblk(nb_rows * 64), TPB, 0, s>>>(nb_rows, rp, ci, val, u, op)));
}
void mr_bsr_op(hipStream_t s, int nr, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const double *val,
               const double *u, const RowOp<double> &o) { mr_bsr_op_t<double>(s, nr, nb_rows, rp, ci, val, u, o); }
void mr_bsr_op(hipStream_t s, int nr, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const float *val,
               const float *u, const RowOp<float> &o) { mr_bsr_op_t<float>(s, nr, nb_rows, rp, ci, val, u, o); }
void mr_bsr_op_h(hipStream_t s, int nr, int64_t nb_rows, const int32_t *rp, const int32_t *ci, const uint16_t *valh,
                 const float *dh, const float *u, const RowOp<float> &o) {
  const MrOp<float> op = to_mr(o);
  if (nr == 8 && ((uintptr_t)u % 16) == 0) {
    const uint2 *vh = reinterpret_cast<const uint2 *>(valh);
    switch (op.mode) {
#define UFT_H8(M) \
  case M: k_mr_bsr_op_h8<M><<<nblk(nb_rows * 64), TPB, 0, s>>>(nb_rows, rp, ci, vh, dh, u, op); break;
      UFT_H8(MR_STORE) UFT_H8(MR_STEP) UFT_H8(MR_RES) UFT_H8(MR_RESFIRST) UFT_H8(MR_RESFIRSTX) UFT_H8(MR_DOTS2)
      UFT_H8(MR_DOTS3) UFT_H8(MR_STOREDOT)
#undef UFT_H8
      default: break;
    }
    return;
  }
  UFT_NR_SWITCH(nr, (k_mr_bsr_op_h<NR><<<nblk(nb_rows * 64), TPB, 0, s>>>(
                        nb_rows, rp, ci, reinterpret_cast<const uint2 *>(valh), dh, u, op)));
}
int mr_bsr_blocks(int64_t nb_rows) { return (int)nblk(nb_rows * 64); }

template <typename T>
static void mr_restrict_first_t(hipStream_t s, int nr, int bs, int64_t nc, const int32_t *cptr, const int32_t *child,
                                const T *P, const T *r, T *bc, const T *dinv, double it, T *rc, T *dc, T *xc) {
  if (bs == 3) {
    UFT_NR_SWITCH(nr, (k_mr_restrict_first<3, T, NR><<<nblk(nc * 64), TPB, 0, s>>>(nc, cptr, child, P, r, bc, dinv, it,
                                                                                   rc, dc, xc)));
  } else {
    UFT_NR_SWITCH(nr, (k_mr_restrict_first<6, T, NR><<<nblk(nc * 64), TPB, 0, s>>>(nc, cptr, child, P, r, bc, dinv, it,
                                                                                   rc, dc, xc)));
  }
}
void mr_restrict_first(hipStream_t s, int nr, int bs, int64_t nc, const int32_t *cptr, const int32_t *child,
                       const double *P, const double *r, double *bc, const double *dinv, double it, double *rc,
                       double *dc, double *xc) {
  mr_restrict_first_t<double>(s, nr, bs, nc, cptr, child, P, r, bc, dinv, it, rc, dc, xc);
}
void mr_restrict_first(hipStream_t s, int nr, int bs, int64_t nc, const int32_t *cptr, const int32_t *child,
                       const float *P, const float *r, float *bc, const float *dinv, double it, float *rc, float *dc,
                       float *xc) {
  mr_restrict_first_t<float>(s, nr, bs, nc, cptr, child, P, r, bc, dinv, it, rc, dc, xc);
}

template <typename T>
static void mr_prolong_comb_t(hipStream_t s, int nr, int bs, int64_t nu, const int32_t *parent, const T *P,
                              const double *sc, int ia, int ib, const T *xa, const T *xb, T *x) {
  if (bs == 3) {
    UFT_NR_SWITCH(nr, (k_mr_prolong_comb<3, T, NR><<<nblk(nu * 3 * NR), TPB, 0, s>>>(nu, parent, P, sc, ia, ib, xa, xb,
                                                                                     x)));
  } else {
    UFT_NR_SWITCH(nr, (k_mr_prolong_comb<6, T, NR><<<nblk(nu * 6 * NR), TPB, 0, s>>>(nu, parent, P, sc, ia, ib, xa, xb,
                                                                                     x)));
  }
}
void mr_prolong_comb(hipStream_t s, int nr, int bs, int64_t nu, const int32_t *parent, const double *P,
                     const double *sc, int ia, int ib, const double *xa, const double *xb, double *x) {
  mr_prolong_comb_t<double>(s, nr, bs, nu, parent, P, sc, ia, ib, xa, xb, x);
}
void mr_prolong_comb(hipStream_t s, int nr, int bs, int64_t nu, const int32_t *parent, const float *P, const double *sc,
                     int ia, int ib, const float *xa, const float *xb, float *x) {
  mr_prolong_comb_t<float>(s, nr, bs, nu, parent, P, sc, ia, ib, xa, xb, x);
}

template <typename T>
static void mr_lincomb_first_t(hipStream_t s, int nr, int64_t n, const double *sc, const T *b, const T *v, T *kr, T *r,
                               const T *dinv, double it, T *d, T *kd) {
  UFT_NR_SWITCH(nr, (k_mr_lincomb_first<T, NR><<<nblk(n * NR), TPB, 0, s>>>(n, sc, b, v, kr, r, dinv, it, d, kd)));
}
void mr_lincomb_first(hipStream_t s, int nr, int64_t n, const double *sc, const double *b, const double *v, double *kr,
                      double *r, const double *dinv, double it, double *d, double *kd) {
  mr_lincomb_first_t<double>(s, nr, n, sc, b, v, kr, r, dinv, it, d, kd);
}
void mr_lincomb_first(hipStream_t s, int nr, int64_t n, const double *sc, const float *b, const float *v, float *kr,
                      float *r, const float *dinv, double it, float *d, float *kd) {
  mr_lincomb_first_t<float>(s, nr, n, sc, b, v, kr, r, dinv, it, d, kd);
}

// many partial rows (the level-1 operator): a first stage of MR_STAGE_BLOCKS
// blocks into the space after them, then the one-block final kernel
void mr_final_kc1(hipStream_t s, int nr, double *part, int nb, double *sc) {
  if (nb > 8 * MR_STAGE_BLOCKS) {
    const int chunk = (nb + MR_STAGE_BLOCKS - 1) / MR_STAGE_BLOCKS;
    double *p2 = part + (size_t)nb * 2 * nr;
    UFT_NR_SWITCH(nr, (k_part_stage<2 * NR><<<MR_STAGE_BLOCKS, TPB, 0, s>>>(part, nb, chunk, p2)));
    UFT_NR_SWITCH(nr, (k_mr_final_kc<2, NR><<<1, TPB, 0, s>>>(p2, MR_STAGE_BLOCKS, sc)));
    return;
  }
  UFT_NR_SWITCH(nr, (k_mr_final_kc<2, NR><<<1, TPB, 0, s>>>(part, nb, sc)));
}
void mr_final_kc2(hipStream_t s, int nr, double *part, int nb, double *sc) {
  if (nb > 8 * MR_STAGE_BLOCKS) {
    const int chunk = (nb + MR_STAGE_BLOCKS - 1) / MR_STAGE_BLOCKS;
    double *p2 = part + (size_t)nb * 3 * nr;
    UFT_NR_SWITCH(nr, (k_part_stage<3 * NR><<<MR_STAGE_BLOCKS, TPB, 0, s>>>(part, nb, chunk, p2)));
    UFT_NR_SWITCH(nr, (k_mr_final_kc<3, NR><<<1, TPB, 0, s>>>(p2, MR_STAGE_BLOCKS, sc)));
    return;
  }
  UFT_NR_SWITCH(nr, (k_mr_final_kc<3, NR><<<1, TPB, 0, s>>>(part, nb, sc)));
}

template <typename T>
static void mr_tail_solve_t(hipStream_t s, int nr, int bs, int nc_units, const int32_t *cptr, const int32_t *child,
                            const T *P, const T *Ainv, T *bc, const T *r, T *xc) {
  // restriction over the whole chip, then the dense product with the inverse
  mr_restrict_first_t<T>(s, nr, bs, nc_units, cptr, child, P, r, bc, nullptr, 0.0, nullptr, nullptr, nullptr);
  const int n = 6 * nc_units;
  const int64_t waves = (n + MR_DM_ROWS - 1) / MR_DM_ROWS;
  UFT_NR_SWITCH(nr, (k_mr_dense_mv<T, NR><<<nblk(waves * 64), TPB, 0, s>>>(n, Ainv, bc, xc)));
}
void mr_tail_solve(hipStream_t s, int nr, int bs, int nc_units, const int32_t *cptr, const int32_t *child,
                   const double *P, const double *Ainv, double *bc, const double *r, double *xc) {
  mr_tail_solve_t<double>(s, nr, bs, nc_units, cptr, child, P, Ainv, bc, r, xc);
}
void mr_tail_solve(hipStream_t s, int nr, int bs, int nc_units, const int32_t *cptr, const int32_t *child,
                   const float *P, const float *Ainv, float *bc, const float *r, float *xc) {
  mr_tail_solve_t<float>(s, nr, bs, nc_units, cptr, child, P, Ainv, bc, r, xc);
}

// ---- outer loop
void mr_dot(hipStream_t s, int nr, int64_t n, const double *a, const double *b, double *part, double *out) {
  UFT_NR_SWITCH(nr, (k_mr_dot<NR><<<mr_red_grid(n * NR), TPB, 0, s>>>(n, a, b, part)));
  final_sum(s, part, (int)mr_red_grid(n * nr), nr, out);
}
void mr_dot_pr(hipStream_t s, int nr, int64_t n, const double *r, const double *z, const double *zo, double *part,
               double *out) {
  UFT_NR_SWITCH(nr, (k_mr_dot_pr<NR><<<mr_red_grid(n * NR), TPB, 0, s>>>(n, r, z, zo, part)));
  final_sum(s, part, (int)mr_red_grid(n * nr), 2 * nr, out);
}
void mr_cg_update(hipStream_t s, int nr, int64_t n, const double *alpha, double *x, const double *p, double *r,
                  const double *q, double *part, double *rr_out) {
  UFT_NR_SWITCH(nr, (k_mr_cg_update<NR><<<mr_red_grid(n * NR), TPB, 0, s>>>(n, alpha, x, p, r, q, part)));
  final_sum(s, part, (int)mr_red_grid(n * nr), nr, rr_out);
}
void mr_cg_update_lo(hipStream_t s, int nr, int64_t n, const double *alpha, double *x, const double *p, double *r,
                     const float *q, float *rf, double *part, double *rr_out) {
  UFT_NR_SWITCH(nr, (k_mr_cg_update_lo<NR><<<mr_red_grid(n * NR), TPB, 0, s>>>(n, alpha, x, p, r, q, rf, part)));
  final_sum(s, part, (int)mr_red_grid(n * nr), nr, rr_out);
}
void mr_take_res(hipStream_t s, int nr, int64_t n, unsigned mask, const double *res, double *r, float *rf,
                 double *part, double *rr_out) {
  UFT_NR_SWITCH(nr, (k_mr_take_res<NR><<<mr_red_grid(n * NR), TPB, 0, s>>>(n, mask, res, r, rf, part)));
  final_sum(s, part, (int)mr_red_grid(n * nr), nr, rr_out);
}
void mr_xpby_lo(hipStream_t s, int nr, int64_t n, const double *z, const double *beta, double *p, float *pf) {
  UFT_NR_SWITCH(nr, (k_mr_xpby_lo<NR><<<nblk(n * NR), TPB, 0, s>>>(n, z, beta, p, pf)));
}
void mr_xpby(hipStream_t s, int nr, int64_t n, const double *z, const double *beta, double *p) {
  UFT_NR_SWITCH(nr, (k_mr_xpby<NR><<<nblk(n * NR), TPB, 0, s>>>(n, z, beta, p)));
}
void mr_alpha(hipStream_t s, int nr, unsigned mask, const double *rz, const double *pq, double *alpha) {
  k_mr_alpha<<<1, 64, 0, s>>>(nr, mask, rz, pq, alpha);
}
void mr_beta(hipStream_t s, int nr, unsigned mask, const double *rzn, double *rzprev, double *beta) {
  k_mr_beta<<<1, 64, 0, s>>>(nr, mask, rzn, rzprev, beta);
}
void mr_sens(hipStream_t s, const uft_ctx *C, int nr, const double *xphys, const double *u, double *dfdx,
             double *part) {
  const SimpCoef sc{C->P.Emin, C->P.Emax - C->P.Emin, C->P.penal};
  const int ne = (int)C->ne;
  const double2 *KE = (const double2 *)C->d_KE;
  if (C->npe == 8) {
    UFT_NR_SWITCH(nr, (k_mr_sens<8, NR><<<nblk((int64_t)ne * NR), TPB, 0, s>>>(ne, C->d_elem, KE, xphys, sc, u, dfdx,
                                                                              part)));
  } else {
    UFT_NR_SWITCH(nr, (k_mr_sens<4, NR><<<nblk((int64_t)ne * NR), TPB, 0, s>>>(ne, C->d_elem, KE, xphys, sc, u, dfdx,
                                                                              part)));
  }
}
int mr_sens_blocks(int64_t ne, int nr) { return (int)nblk(ne * nr); }
void mr_extract(hipStream_t s, int64_t n, int nr, int k, const double *src, double *dst) {
  k_mr_extract<<<nblk(n), TPB, 0, s>>>(n, nr, k, src, dst);
}
void mr_set_dofs(hipStream_t s, int n, int nr, int k, const int32_t *idx, const double *val, double *F) {
  if (n > 0) k_mr_set_dofs<<<nblk(n), TPB, 0, s>>>(n, nr, k, idx, val, F);
}

}  // namespace uftk
