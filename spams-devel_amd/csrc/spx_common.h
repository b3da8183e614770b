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

// ---- cross-lane moves (all 64 lanes active) --------------------------------
// DPP controls used below (gfx9 encoding): quad_perm [1,0,3,2] = 0xB1,
// quad_perm [2,3,0,1] = 0x4E, row_ror:4 = 0x124, row_ror:8 = 0x128.  After
// the four steps every lane holds the reduction of its row of 16 lanes; the
// four row results are then combined from lanes 0, 16, 32 and 48.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_v(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_v(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i<CTRL>(static_cast<int>(b & 0xffffffffLL));
  const int hi = dpp_i<CTRL>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}
__device__ __forceinline__ int lane_i(int v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}
__device__ __forceinline__ float lane_v(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_v(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

template <typename T, int CTRL>
__device__ __forceinline__ T sum_step(T v) {
  return v + dpp_v<CTRL>(v);
}

// Wave-wide sum; every lane returns the same value and the association
// order is fixed (deterministic).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v = sum_step<T, 0xB1>(v);
  v = sum_step<T, 0x4E>(v);
  v = sum_step<T, 0x124>(v);
  v = sum_step<T, 0x128>(v);
  return (lane_v(v, 0) + lane_v(v, 16)) + (lane_v(v, 32) + lane_v(v, 48));
}

// (key, idx) ordering of the arg-min searches: smaller key wins, the lower
// index wins ties, a NaN key loses to any number.  Branch-free.
template <typename T>
__device__ __forceinline__ bool key_better(T k1, int i1, T k2, int i2) {
  const bool lt = k1 < k2;
  const bool tie = (k1 == k2) & (i1 < i2);
  const bool nan2 = (k2 != k2) & (k1 == k1);
  return lt | tie | nan2;
}

template <typename T>
__device__ __forceinline__ void take_if_better(T& key, int& idx, T& pay, T k2,
                                               int i2, T p2) {
  const bool b = key_better(k2, i2, key, idx);
  key = b ? k2 : key;
  idx = b ? i2 : idx;
  pay = b ? p2 : pay;
}

template <typename T, int CTRL>
__device__ __forceinline__ void argmin_step(T& key, int& idx, T& pay) {
  const T k2 = dpp_v<CTRL>(key);
  const int i2 = dpp_i<CTRL>(idx);
  const T p2 = dpp_v<CTRL>(pay);
  take_if_better(key, idx, pay, k2, i2, p2);
}

// Wave-wide arg-min; every lane returns the winner.
template <typename T>
__device__ __forceinline__ void wave_argmin(T& key, int& idx, T& payload) {
  argmin_step<T, 0xB1>(key, idx, payload);
  argmin_step<T, 0x4E>(key, idx, payload);
  argmin_step<T, 0x124>(key, idx, payload);
  argmin_step<T, 0x128>(key, idx, payload);
  T k = lane_v(key, 0), p = lane_v(payload, 0);
  int i = lane_i(idx, 0);
#pragma unroll
  for (int r = 16; r < 64; r += 16)
    take_if_better(k, i, p, lane_v(key, r), lane_i(idx, r), lane_v(payload, r));
  key = k;
  idx = i;
  payload = p;
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
      const KeyIdx<T> c = red[w];
      take_if_better(b.key, b.idx, b.payload, c.key, c.idx, c.payload);
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
