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
// (ProdMatrix::addDiag), into out (rows 0..ldg-1, zero past p).  `rows`
// lists the nsel rows to visit in ascending order: the observed rows, or,
// when `comp` (more rows observed than not), the unobserved ones, whose
// terms are subtracted from the resident Gram column G(:, c) -- fewer rows
// of D' to read.  Each of the GT threads forms its NQ chunks of rows (chunk
// q of thread t covers rows N (t + GT q) .. + N - 1, the solvers' row
// ownership).  Dt is D' (p x m, leading dimension p), Dd is D (m x p), G
// the resident Gram (leading dimension ldg, no ridge).  Ends with a barrier.
template <typename T, int GT, int NQ>
__device__ __forceinline__ void masked_gram_col(const T* __restrict__ Dt, const T* __restrict__ Dd,
                                                const T* __restrict__ G, int m, int p, int ldg,
                                                const int* __restrict__ rows, int nsel, bool comp,
                                                int c, T* out, T diag) {
  constexpr int VEC = GramVec<T>::N;
  typedef typename GramVec<T>::type V;
  const int tid = threadIdx.x;
  const bool vec = p % VEC == 0;
  V acc[NQ];
  const T* gc = G + (size_t)c * ldg;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int j0 = VEC * (tid + GT * q);
    if (comp && j0 + VEC <= ldg) {
      acc[q] = *reinterpret_cast<const V*>(gc + j0);
    } else {
#pragma unroll
      for (int e = 0; e < VEC; ++e) acc[q][e] = comp && j0 + e < ldg ? gc[j0 + e] : T(0);
    }
  }
  const T* dc = Dd + (size_t)c * m;
  const T sgn = comp ? T(-1) : T(1);
#pragma unroll 4
  for (int k = 0; k < nsel; ++k) {
    const int r = rows[k];
    const T w = sgn * dc[r];
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

// The rows masked_gram_col visits for a signal with nobs of m rows observed.
__device__ __forceinline__ bool masked_gram_comp(int nobs, int m) { return 2 * nobs > m; }
__device__ __forceinline__ int masked_gram_nsel(int nobs, int m) {
  return masked_gram_comp(nobs, m) ? m - nobs : nobs;
}

}  // namespace spx
