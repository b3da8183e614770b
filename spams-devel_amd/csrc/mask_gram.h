// Gram columns over a signal's observed rows (masked solvers).
#pragma once

#include <hip/hip_runtime.h>

namespace spx {

template <typename T>
struct GramVec;
template <>
struct GramVec<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int N = 4;
};
template <>
struct GramVec<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int N = 2;
};

// Gram column c over a signal's observed rows, G_M(:, c) = sum over the
// observed rows r of D(r, :)' D(r, c) (ProdMatrix::extract_rawCol of the
// masked copy of D, linalg/linalg.h:5157-5167), plus `diag` at row c
// (ProdMatrix::addDiag), into out (rows 0..ldg-1, zero past p).  Each of the
// GT threads forms its NQ chunks of rows (chunk q of thread t covers rows
// N (t + GT q) .. + N - 1, the solvers' row ownership).  Dt is D' (p x m,
// leading dimension p), Dd is D (m x p); rows outside the mask add exact
// zeros.  Ends with a barrier.
template <typename T, int GT, int NQ>
__device__ __forceinline__ void masked_gram_col(const T* __restrict__ Dt, const T* __restrict__ Dd,
                                                int m, int p, int ldg,
                                                const unsigned char* __restrict__ msk, int c,
                                                T* out, T diag) {
  constexpr int VEC = GramVec<T>::N;
  typedef typename GramVec<T>::type V;
  const int tid = threadIdx.x;
  V acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = V(0);
  const T* dc = Dd + (size_t)c * m;
  const bool vec = p % VEC == 0;
#pragma unroll 4
  for (int r = 0; r < m; ++r) {
    const T w = msk[r] ? dc[r] : T(0);
    const T* row = Dt + (size_t)r * p;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int j0 = VEC * (tid + GT * q);
      V x;
      if (vec && j0 + VEC <= p) {
        x = *reinterpret_cast<const V*>(row + j0);
      } else {
#pragma unroll
        for (int e = 0; e < VEC; ++e) x[e] = j0 + e < p ? row[j0 + e] : T(0);
      }
      acc[q] += x * w;
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int j0 = VEC * (tid + GT * q);
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int j = j0 + e;
      if (j < ldg) out[j] = j < p ? acc[q][e] + (j == c ? diag : T(0)) : T(0);
    }
  }
  __syncthreads();
}

}  // namespace spx
