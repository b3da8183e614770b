// Work ordering for the persistent solvers: signal indices sorted by a key
// (the per-signal constraint), so that the longest homotopy paths are taken
// first from the work queue (longest-processing-time-first scheduling).
//
// A stable least-significant-digit radix sort of 32-bit key images on
// their 16 most significant bits (sign, exponent and 7 mantissa bits of the
// FP32 key: a 1 % resolution, ample for scheduling; keys equal there keep
// their input order), 8 bits per pass (2 passes):
//   1. keys -> order-preserving unsigned images (FP64 keys are rounded to
//      FP32 first: the order only schedules work, it never changes a
//      signal's result); NaN maps to 0xffffffff and so sorts last, in input
//      order; with a device count only the leading entries take part;
//   2. per pass: each tile of kTile keys counts its digits (LDS counters),
//      an exclusive scan over (digit, tile) gives every tile its output
//      base per digit, and the tile scatters its keys and values: within
//      each round of 256 keys a key's rank among equal digits comes from
//      wave ballots over the digit's bits (lanes in order) plus the counts
//      of the earlier waves and rounds, so equal keys keep their order.
#include "spx_common.h"
#include "spx_internal.h"

namespace spx {

namespace {

constexpr int kThreads = 256;          // 4 wavefronts
constexpr int kItems = 8;              // keys per thread per tile
constexpr int kTile = kThreads * kItems;
constexpr int kBins = 256;
constexpr int kScan = 1024;            // scan workgroup

__device__ __forceinline__ unsigned f2ord(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// key images and initial values: value = index (sort_order) or list entry
template <typename T>
__global__ __launch_bounds__(kThreads) void keys_init_kernel(
    const T* __restrict__ key, const int* __restrict__ list, const int* __restrict__ count,
    int n, unsigned* __restrict__ kout, int* __restrict__ vout) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int lim = count ? *count : n;
  unsigned u = 0xffffffffu;
  if (i < lim) {
    const float f = static_cast<float>(key[i]);
    u = f == f ? f2ord(f) : 0xffffffffu;
  }
  kout[i] = u;
  vout[i] = list ? (i < lim ? list[i] : -1) : i;
}

// digit counts of one tile: hist[d * ntiles + tile]; only the first *count
// entries (all n without a count) take part
__global__ __launch_bounds__(kThreads) void digit_hist_kernel(
    const unsigned* __restrict__ k, int n, const int* __restrict__ count, int shift, int ntiles,
    int* __restrict__ hist) {
  __shared__ int cnt[kBins];
  const int tid = threadIdx.x;
  cnt[tid] = 0;
  __syncthreads();
  if (count) n = *count < n ? *count : n;
  const int base = blockIdx.x * kTile;
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int i = base + it * kThreads + tid;
    if (i < n) atomicAdd(&cnt[(k[i] >> shift) & 0xff], 1);
  }
  __syncthreads();
  hist[tid * ntiles + blockIdx.x] = cnt[tid];
}

// exclusive scan of n ints in place, one workgroup, chunks of kScan in order
__global__ __launch_bounds__(kScan) void excl_scan_kernel(int* __restrict__ v, int n) {
  __shared__ int s[kScan];
  __shared__ int carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int b = 0; b < n; b += kScan) {
    const int i = b + tid;
    const int x = i < n ? v[i] : 0;
    s[tid] = x;
    __syncthreads();
    for (int o = 1; o < kScan; o <<= 1) {
      const int t = tid >= o ? s[tid - o] : 0;
      __syncthreads();
      s[tid] += t;
      __syncthreads();
    }
    if (i < n) v[i] = carry + s[tid] - x;
    __syncthreads();
    if (tid == kScan - 1) carry += s[tid];
    __syncthreads();
  }
}

// per-tile partial scan: block b scans hist[b * kScan ...] and writes its
// total to sums[b] (first level of a two-level scan for long inputs)
__global__ __launch_bounds__(kScan) void scan_part_kernel(int* __restrict__ v, int n,
                                                          int* __restrict__ sums) {
  __shared__ int s[kScan];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * kScan + tid;
  const int x = i < n ? v[i] : 0;
  s[tid] = x;
  __syncthreads();
  for (int o = 1; o < kScan; o <<= 1) {
    const int t = tid >= o ? s[tid - o] : 0;
    __syncthreads();
    s[tid] += t;
    __syncthreads();
  }
  if (i < n) v[i] = s[tid] - x;
  if (tid == kScan - 1) sums[blockIdx.x] = s[tid];
}

__global__ __launch_bounds__(kScan) void scan_add_kernel(int* __restrict__ v, int n,
                                                         const int* __restrict__ sums) {
  const int i = blockIdx.x * kScan + threadIdx.x;
  if (i < n) v[i] += sums[blockIdx.x];
}

// stable scatter of one tile by the digit at `shift` (first *count entries)
__global__ __launch_bounds__(kThreads) void digit_scatter_kernel(
    const unsigned* __restrict__ kin, const int* __restrict__ vin, int n,
    const int* __restrict__ count, int shift, int ntiles, const int* __restrict__ offs,
    unsigned* __restrict__ kout, int* __restrict__ vout) {
  __shared__ int base[kBins];          // tile output base per digit (running)
  __shared__ int wcnt[kThreads / 64][kBins];
  if (count) n = *count < n ? *count : n;
  if ((int)blockIdx.x * kTile >= n) return;  // uniform: the whole tile is past the count
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  base[tid] = offs[tid * ntiles + blockIdx.x];
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int t0 = blockIdx.x * kTile;
  for (int it = 0; it < kItems; ++it) {
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const int i = t0 + it * kThreads + tid;
    const bool ok = i < n;
    const unsigned key = ok ? kin[i] : 0u;
    const int d = (key >> shift) & 0xff;
    // lanes of this wave holding the same digit (invalid lanes excluded)
    unsigned long long peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const unsigned long long m = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? m : ~m;
    }
    const int rank = __popcll(peers & lt);
    // the highest peer lane records the wave's count of the digit
    if (ok && (peers >> lane) == 1ull) wcnt[wv][d] = __popcll(peers);
    __syncthreads();
    if (ok) {
      int off = base[d] + rank;
      for (int w = 0; w < wv; ++w) off += wcnt[w][d];
      kout[off] = key;
      vout[off] = vin[i];
    }
    __syncthreads();
    int tot = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) tot += wcnt[w][tid];
    base[tid] += tot;
    __syncthreads();
  }
}

// scan of hist (kBins * ntiles entries) in place
int scan_counts(hipStream_t st, int* hist, int len, int* sums) {
  if (len <= 8 * kScan) {
    hipLaunchKernelGGL(excl_scan_kernel, dim3(1), dim3(kScan), 0, st, hist, len);
    SPX_HIP(hipGetLastError());
    return SPX_OK;
  }
  const int nb = (len + kScan - 1) / kScan;
  hipLaunchKernelGGL(scan_part_kernel, dim3(nb), dim3(kScan), 0, st, hist, len, sums);
  SPX_HIP(hipGetLastError());
  hipLaunchKernelGGL(excl_scan_kernel, dim3(1), dim3(kScan), 0, st, sums, nb);
  SPX_HIP(hipGetLastError());
  hipLaunchKernelGGL(scan_add_kernel, dim3(nb), dim3(kScan), 0, st, hist, len, sums);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

// Sorts n entries (values = indices, or `list` entries) by `key`,
// ascending and stable; the sorted values land in `out`.  With `count`
// (device), only the first *count entries are sorted (out past them is left
// unspecified).
template <typename T>
int radix_sort(Ctx* ctx, const T* key, const int* list, const int* count, int n, int* out) {
  if (n <= 0) return SPX_OK;
  hipStream_t st = ctx->stream;
  const int ntiles = (n + kTile - 1) / kTile;
  const int hlen = kBins * ntiles;
  const int nsum = (hlen + kScan - 1) / kScan;
  void* buf = nullptr;
  const size_t need = sorted_scratch_bytes(n);
  int rc = ctx->scratch_get(9, need, &buf);
  if (rc) return rc;
  char* p = static_cast<char*>(buf);
  auto take = [&](size_t bytes) {
    char* r = p;
    p += (bytes + 255) & ~size_t(255);
    return r;
  };
  unsigned* k0 = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * (size_t)n));
  unsigned* k1 = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * (size_t)n));
  int* v0 = reinterpret_cast<int*>(take(sizeof(int) * (size_t)n));
  int* v1 = reinterpret_cast<int*>(take(sizeof(int) * (size_t)n));
  int* hist = reinterpret_cast<int*>(take(sizeof(int) * (size_t)hlen));
  int* sums = reinterpret_cast<int*>(take(sizeof(int) * (size_t)(nsum > 0 ? nsum : 1)));
  hipLaunchKernelGGL(keys_init_kernel<T>, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     st, key, list, count, n, k0, v0);
  SPX_HIP(hipGetLastError());
  for (int pass = 0; pass < 2; ++pass) {
    const int shift = 16 + 8 * pass;
    const bool last = pass == 1;
    unsigned* ko = k1;
    int* vo = last ? out : v1;
    hipLaunchKernelGGL(digit_hist_kernel, dim3(ntiles), dim3(kThreads), 0, st, k0, n, count,
                       shift, ntiles, hist);
    SPX_HIP(hipGetLastError());
    rc = scan_counts(st, hist, hlen, sums);
    if (rc) return rc;
    hipLaunchKernelGGL(digit_scatter_kernel, dim3(ntiles), dim3(kThreads), 0, st, k0, v0, n,
                       count, shift, ntiles, hist, ko, vo);
    SPX_HIP(hipGetLastError());
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  return SPX_OK;
}

}  // namespace

size_t sorted_scratch_bytes(int n) {
  if (n <= 0) return 256;
  const size_t ntiles = ((size_t)n + kTile - 1) / kTile;
  const size_t hlen = kBins * ntiles;
  const size_t nsum = (hlen + kScan - 1) / kScan;
  auto pad = [](size_t b) { return (b + 255) & ~size_t(255); };
  return 4 * pad(sizeof(int) * (size_t)n) + pad(sizeof(int) * hlen) +
         pad(sizeof(int) * (nsum > 0 ? nsum : 1));
}

template <typename T>
int sort_list(Ctx* ctx, const T* key, const int* list, const int* count_dev, int n, int* out) {
  return radix_sort<T>(ctx, key, list, count_dev, n, out);
}

template int sort_list<float>(Ctx*, const float*, const int*, const int*, int, int*);
template int sort_list<double>(Ctx*, const double*, const int*, const int*, int, int*);

template <typename T>
int sort_order(Ctx* ctx, const T* key, int n, int* order) {
  return radix_sort<T>(ctx, key, nullptr, nullptr, n, order);
}

template int sort_order<float>(Ctx*, const float*, int, int*);
template int sort_order<double>(Ctx*, const double*, int, int*);

}  // namespace spx
