// Device helpers shared by the element kernels: compile-time loops, the
// row-major packed upper triangle of a symmetric nd x nd element matrix,
// the SIMP modulus, and the symmetric multiply-accumulate of one packed entry.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "uft_internal.h"

namespace uftd {

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// row / column of entry p of the row-major packed upper triangle of an nd x nd matrix
__host__ __device__ constexpr int pk_row(int p, int nd) {
  int r = 0;
  while (p >= nd - r) {
    p -= nd - r;
    ++r;
  }
  return r;
}
__host__ __device__ constexpr int pk_col(int p, int nd) {
  int r = 0;
  while (p >= nd - r) {
    p -= nd - r;
    ++r;
  }
  return r + p;
}
__host__ __device__ inline int pk_index(int r, int c, int nd) {
  if (r > c) {
    int t = r;
    r = c;
    c = t;
  }
  return r * nd - (r * (r - 1)) / 2 + (c - r);
}

__device__ __forceinline__ double simpE(const SimpCoef &sc, double x) { return sc.Emin + pow(x, sc.penal) * sc.dE; }

// y += k * u for packed entry (I, J) of a symmetric matrix (both halves)
template <int I, int J, typename T>
__device__ __forceinline__ void kacc(T k, const T *ue, T *ye) {
  if constexpr (I == J) {
    ye[I] += k * ue[I];
  } else {
    ye[I] += k * ue[J];
    ye[J] += k * ue[I];
  }
}

}  // namespace uftd
