// Shared device/host helpers for the spx engine (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace spx {

// ------------------------------------------------------------------ errors

// Records a message in the thread-local error slot read by spx_last_error().
void set_error(const char* fmt, ...);

#define SPX_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::spx::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                  \
      return SPX_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

// ---------------------------------------------------------------- layouts

// Internal leading dimension of every p-length column (DtX, G, solver state):
// rounded up to 16 elements so that columns start 64-byte aligned and can be
// read with 128-bit loads; padding rows are kept at zero.
inline int pad_ld(int p) { return (p + 15) & ~15; }

// ------------------------------------------------------- wavefront helpers

constexpr int kWave = 64;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// (key, idx) arg-min with the lowest index winning ties; NaN keys never win
// against a number.  Every lane returns the wave-wide winner.
template <typename T>
__device__ __forceinline__ bool key_better(T k1, int i1, T k2, int i2) {
  // true when (k1, i1) beats (k2, i2)
  if (k1 < k2) return true;
  if (k1 == k2) return i1 < i2;
  return (k2 != k2) && (k1 == k1);  // k2 is NaN, k1 is not
}

template <typename T>
__device__ __forceinline__ void wave_argmin(T& key, int& idx, T& payload) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T k2 = __shfl_xor(key, o, kWave);
    int i2 = __shfl_xor(idx, o, kWave);
    T p2 = __shfl_xor(payload, o, kWave);
    if (key_better(k2, i2, key, idx)) {
      key = k2;
      idx = i2;
      payload = p2;
    }
  }
}

template <typename T>
struct KeyIdx {
  T key;
  int idx;
  T payload;
};

// Group-wide arg-min for a group of W wavefronts that is exactly one
// workgroup.  `red` is LDS scratch of at least W entries.  All threads return
// the winner.  Contains two barriers when W > 1 and none when W == 1.
template <typename T, int W>
__device__ __forceinline__ void group_argmin(T& key, int& idx, T& payload,
                                             KeyIdx<T>* red) {
  wave_argmin(key, idx, payload);
  if constexpr (W > 1) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (lane == 0) red[wid] = KeyIdx<T>{key, idx, payload};
    __syncthreads();
    KeyIdx<T> b = red[0];
#pragma unroll
    for (int w = 1; w < W; ++w) {
      KeyIdx<T> c = red[w];
      if (key_better(c.key, c.idx, b.key, b.idx)) b = c;
    }
    key = b.key;
    idx = b.idx;
    payload = b.payload;
    __syncthreads();
  }
}

// Group-wide sums of up to three values (same conventions as group_argmin).
template <typename T, int W>
__device__ __forceinline__ void group_sum3(T& a, T& b, T& c, T* red) {
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  if constexpr (W > 1) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (lane == 0) {
      red[3 * wid + 0] = a;
      red[3 * wid + 1] = b;
      red[3 * wid + 2] = c;
    }
    __syncthreads();
    a = b = c = T(0);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      a += red[3 * w + 0];
      b += red[3 * w + 1];
      c += red[3 * w + 2];
    }
    __syncthreads();
  }
}

template <typename T>
__device__ __forceinline__ T dev_inf() {
  return __builtin_huge_val();
}
template <>
__device__ __forceinline__ float dev_inf<float>() {
  return __builtin_huge_valf();
}

}  // namespace spx
