// Work ordering for the persistent solvers: signal indices sorted by a key
// (the per-signal constraint), so that the longest homotopy paths are taken
// first from the work queue (longest-processing-time-first scheduling).
#include <hipcub/hipcub.hpp>

#include "spx_common.h"
#include "spx_internal.h"

namespace spx {

namespace {

__global__ void iota_kernel(int* v, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) v[i] = i;
}

// Keys of the entries past the device count sort last.
template <typename T>
__global__ void pad_keys_kernel(const T* key, const int* count, int n, T* out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = i < *count ? key[i] : T(__builtin_huge_val());
}

}  // namespace

template <typename T>
int sort_list(Ctx* ctx, const T* key, const int* list, const int* count_dev, int n, int* out) {
  if (n <= 0) return SPX_OK;
  hipStream_t st = ctx->stream;
  void* kb = nullptr;
  int rc = ctx->scratch_get(11, 2 * sizeof(T) * (size_t)n + 256, &kb);
  if (rc) return rc;
  T* kin = static_cast<T*>(kb);
  T* kout = reinterpret_cast<T*>(
      (reinterpret_cast<uintptr_t>(kin + n) + 255) & ~uintptr_t(255));
  hipLaunchKernelGGL(pad_keys_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, key, count_dev,
                     n, kin);
  SPX_HIP(hipGetLastError());
  size_t tmp_bytes = 0;
  SPX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kin, kout, list, out, n, 0,
                                             (int)(8 * sizeof(T)), st));
  void* tmp = nullptr;
  rc = ctx->scratch_get(9, tmp_bytes + 16, &tmp);
  if (rc) return rc;
  SPX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, list, out, n, 0,
                                             (int)(8 * sizeof(T)), st));
  return SPX_OK;
}

template int sort_list<float>(Ctx*, const float*, const int*, const int*, int, int*);
template int sort_list<double>(Ctx*, const double*, const int*, const int*, int, int*);

template <typename T>
int sort_order(Ctx* ctx, const T* key, int n, int* order) {
  if (n <= 0) return SPX_OK;
  hipStream_t st = ctx->stream;
  void* kb = nullptr;
  int rc = ctx->scratch_get(11, (sizeof(T) + sizeof(int)) * (size_t)n + 256, &kb);
  if (rc) return rc;
  T* keys_out = static_cast<T*>(kb);
  int* idx_in = reinterpret_cast<int*>(
      (reinterpret_cast<uintptr_t>(keys_out + n) + 255) & ~uintptr_t(255));
  hipLaunchKernelGGL(iota_kernel, dim3((n + 255) / 256), dim3(256), 0, st, idx_in, n);
  SPX_HIP(hipGetLastError());
  size_t tmp_bytes = 0;
  SPX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, keys_out, idx_in,
                                             order, n, 0, (int)(8 * sizeof(T)), st));
  void* tmp = nullptr;
  rc = ctx->scratch_get(9, tmp_bytes + 16, &tmp);
  if (rc) return rc;
  SPX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, keys_out, idx_in, order,
                                             n, 0, (int)(8 * sizeof(T)), st));
  return SPX_OK;
}

template int sort_order<float>(Ctx*, const float*, int, int*);
template int sort_order<double>(Ctx*, const double*, int, int*);

}  // namespace spx
