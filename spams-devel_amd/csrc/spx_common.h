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

// A value the caller knows to be the same in every lane (an LDS word read at
// a uniform address), moved to a scalar register: arithmetic on it (column
// addresses, multipliers) then costs no vector instructions.
__device__ __forceinline__ int uniform_v(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform_v(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform_v(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32));
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
    key = uniform_v(b.key);
    idx = uniform_v(b.idx);
    payload = uniform_v(b.payload);
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
    a = uniform_v(a);
    b = uniform_v(b);
    c = uniform_v(c);
    __syncthreads();
  }
}

// ---- reductions with fewer instructions and barriers ----------------------
//
// A non-negative FP32 key (sign bit clear; +INF; NaN orders after +INF) and a
// 32-bit index pack into one unsigned 64-bit word whose unsigned order is the
// order of the arg-min searches: the smaller key first, the lower index on
// ties, a NaN key after every number.  One 64-bit minimum then replaces the
// compare-and-select of key, index and payload.
__device__ __forceinline__ unsigned long long pack_key(float key, unsigned idx) {
  return (static_cast<unsigned long long>(__float_as_uint(key)) << 32) | idx;
}
__device__ __forceinline__ float packed_key(unsigned long long v) {
  return __uint_as_float(static_cast<unsigned>(v >> 32));
}

template <int CTRL>
__device__ __forceinline__ void min_step_u64(unsigned long long& v) {
  const unsigned lo = static_cast<unsigned>(dpp_i<CTRL>(static_cast<int>(v)));
  const unsigned hi = static_cast<unsigned>(dpp_i<CTRL>(static_cast<int>(v >> 32)));
  const unsigned long long o = (static_cast<unsigned long long>(hi) << 32) | lo;
  v = o < v ? o : v;
}
__device__ __forceinline__ unsigned long long lane_u64(unsigned long long v, int lane) {
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), lane));
  const unsigned hi =
      static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), lane));
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Wave-wide unsigned 64-bit minimum; every lane returns it.
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
  min_step_u64<0xB1>(v);
  min_step_u64<0x4E>(v);
  min_step_u64<0x124>(v);
  min_step_u64<0x128>(v);
  unsigned long long a = lane_u64(v, 0), b = lane_u64(v, 16);
  unsigned long long c = lane_u64(v, 32), d = lane_u64(v, 48);
  a = b < a ? b : a;
  c = d < c ? d : c;
  return c < a ? c : a;
}

// Wave-wide minimum of FP32 values (a NaN never wins over a number); every
// lane returns it.
__device__ __forceinline__ float wave_min_f32(float v) {
  v = fminf(v, dpp_v<0xB1>(v));
  v = fminf(v, dpp_v<0x4E>(v));
  v = fminf(v, dpp_v<0x124>(v));
  v = fminf(v, dpp_v<0x128>(v));
  return fminf(fminf(lane_v(v, 0), lane_v(v, 16)), fminf(lane_v(v, 32), lane_v(v, 48)));
}

// Group-wide reductions for a group of W wavefronts that is exactly one
// workgroup, with ONE barrier each: the per-wave results go to one of two LDS
// buffers, alternately (`parity`, the same in every thread).  A buffer is
// written again only two reductions later, and every thread has passed the
// barrier of the reduction in between, which it reached after reading the
// buffer.  All threads return the result.
template <int W>
__device__ __forceinline__ unsigned long long group_min_u64(unsigned long long v,
                                                            unsigned long long* red,
                                                            int& parity) {
  v = wave_min_u64(v);
  if constexpr (W > 1) {
    unsigned long long* r = red + parity * W;
    parity ^= 1;
    if ((threadIdx.x & (kWave - 1)) == 0) r[threadIdx.x / kWave] = v;
    __syncthreads();
    v = r[0];
#pragma unroll
    for (int w = 1; w < W; ++w) v = r[w] < v ? r[w] : v;
    v = (static_cast<unsigned long long>(static_cast<unsigned>(uniform_v(static_cast<int>(v >> 32))))
         << 32) |
        static_cast<unsigned>(uniform_v(static_cast<int>(v)));
  }
  return v;
}

template <int W>
__device__ __forceinline__ float group_min_f32(float v, float* red, int& parity) {
  v = wave_min_f32(v);
  if constexpr (W > 1) {
    float* r = red + parity * 2 * W;
    parity ^= 1;
    if ((threadIdx.x & (kWave - 1)) == 0) r[threadIdx.x / kWave] = v;
    __syncthreads();
    v = r[0];
#pragma unroll
    for (int w = 1; w < W; ++w) v = fminf(v, r[w]);
    v = uniform_v(v);
  }
  return v;
}

// Sums of one or two values, in the order of group_sum3 (wave sums, then the
// waves in order from zero), so the results are those of group_sum3.
template <typename T, int W>
__device__ __forceinline__ void group_sum2(T& a, T& b, T* red, int& parity) {
  a = wave_sum(a);
  b = wave_sum(b);
  if constexpr (W > 1) {
    T* r = red + parity * 2 * W;
    parity ^= 1;
    const int wid = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
      r[2 * wid] = a;
      r[2 * wid + 1] = b;
    }
    __syncthreads();
    a = b = T(0);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      a += r[2 * w];
      b += r[2 * w + 1];
    }
    a = uniform_v(a);
    b = uniform_v(b);
  }
}
template <typename T, int W>
__device__ __forceinline__ T group_sum1(T a, T* red, int& parity) {
  a = wave_sum(a);
  if constexpr (W > 1) {
    T* r = red + parity * 2 * W;
    parity ^= 1;
    if ((threadIdx.x & (kWave - 1)) == 0) r[threadIdx.x / kWave] = a;
    __syncthreads();
    a = T(0);
#pragma unroll
    for (int w = 0; w < W; ++w) a += r[w];
    a = uniform_v(a);
  }
  return a;
}

// Sub-array of the dynamic LDS block `base` starting at the first 16-byte
// boundary at or after `end` (a pointer into the same block).  The address is
// formed by pointer arithmetic on the block, so the compiler keeps it in the
// LDS address space and emits ds_* instructions; a round trip through an
// integer would make it a generic pointer, accessed with flat_* instructions
// (longer latency, and each access waits on the vector-memory counter too).
template <typename U, typename P>
__device__ __forceinline__ U* lds_next16(char* base, P* end) {
  const int off = static_cast<int>(reinterpret_cast<char*>(end) - base);
  return reinterpret_cast<U*>(base + ((off + 15) & ~15));
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
