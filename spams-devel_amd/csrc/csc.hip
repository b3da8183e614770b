// CSC assembly from the solvers' per-signal (value, index) lists:
// SpMatrix<T>::convert + the per-column index sort (linalg/linalg.h:
// 4840-4862, linalg/misc.h:138-166).  Column i of the output holds signal i's
// active set with row indices ascending; explicit zeros are kept, as in the
// reference.
//
//   1. exclusive scan of the per-signal counts -> pB (n + 1)
//   2. one wavefront per column: every lane ranks its entries by counting the
//      smaller indices of the column (indices within a column are distinct),
//      then scatters (row, value) to pB[col] + rank.
#include "spx_common.h"
#include "spx_internal.h"

namespace spx {

namespace {

constexpr int kScanBlock = 1024;

// per-block inclusive scan of count[], block totals to blockSums
__global__ __launch_bounds__(kScanBlock) void scan_blocks(
    const int* __restrict__ count, int n, long long* __restrict__ out,
    long long* __restrict__ blockSums) {
  __shared__ long long s[kScanBlock];
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  long long v = i < n ? count[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < kScanBlock; o <<= 1) {
    long long t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x];
  if (threadIdx.x == kScanBlock - 1) blockSums[blockIdx.x] = s[threadIdx.x];
}

// sequential-by-chunks scan of the block totals (one workgroup)
__global__ __launch_bounds__(kScanBlock) void scan_sums(long long* sums, int nb) {
  __shared__ long long s[kScanBlock];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += kScanBlock) {
    const int i = base + threadIdx.x;
    s[threadIdx.x] = i < nb ? sums[i] : 0;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
      long long t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) sums[i] = s[threadIdx.x] + carry;  // inclusive
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1) carry += s[threadIdx.x];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void scan_finish(
    const long long* __restrict__ incl, const long long* __restrict__ sums,
    int n, int* __restrict__ pB, long long* __restrict__ total) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i > n) return;
  long long v = 0;
  if (i > 0) {
    const int k = i - 1;
    const int b = k / kScanBlock;
    v = incl[k] + (b > 0 ? sums[b - 1] : 0);
  }
  pB[i] = (int)v;
  if (i == n) *total = v;
}

template <typename T>
__global__ __launch_bounds__(256) void fill_columns(
    const T* __restrict__ v, const int* __restrict__ r, int L, int n,
    const int* __restrict__ pB, int* __restrict__ rOut, T* __restrict__ vOut) {
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (col >= n) return;
  const int base = pB[col];
  const int cnt = pB[col + 1] - base;
  const int* rc = r + (size_t)col * L;
  const T* vc = v + (size_t)col * L;
  for (int k = lane; k < cnt; k += 64) {
    const int rk = rc[k];
    int rank = 0;
    for (int q = 0; q < cnt; ++q) rank += rc[q] < rk;
    rOut[base + rank] = rk;
    vOut[base + rank] = vc[k];
  }
}

}  // namespace

template <typename T>
int csc_count_scan(Ctx* ctx, const int* count, int n, int* pB, long long* nnz) {
  hipStream_t st = ctx->stream;
  const int nb = (n + kScanBlock - 1) / kScanBlock;
  void* tmp = nullptr;
  const size_t need = sizeof(long long) * ((size_t)(nb > 0 ? nb : 1) * kScanBlock + nb + 2);
  int rc = ctx->scratch_get(2, need, &tmp);
  if (rc) return rc;
  long long* incl = static_cast<long long*>(tmp);
  long long* sums = incl + (size_t)(nb > 0 ? nb : 1) * kScanBlock;
  long long* total = sums + nb + 1;
  {
  ScopedTimer tm(ctx, SPX_TIMER_CSC);
  if (n > 0) {
    hipLaunchKernelGGL(scan_blocks, dim3(nb), dim3(kScanBlock), 0, st, count,
                       n, incl, sums);
    SPX_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanBlock), 0, st, sums, nb);
    SPX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(scan_finish, dim3((n + 1 + 255) / 256), dim3(256), 0, st,
                     incl, sums, n, pB, total);
  SPX_HIP(hipGetLastError());
  }
  long long h = 0;
  SPX_HIP(hipMemcpyAsync(&h, total, sizeof(long long), hipMemcpyDeviceToHost, st));
  SPX_HIP(hipStreamSynchronize(st));
  *nnz = h;
  return SPX_OK;
}

template <typename T>
int csc_fill(Ctx* ctx, const T* v, const int* r, int L, int n, const int* pB,
             int* rOut, T* vOut) {
  if (n <= 0) return SPX_OK;
  ScopedTimer tm(ctx, SPX_TIMER_CSC);
  hipLaunchKernelGGL(fill_columns<T>, dim3((n + 3) / 4), dim3(256), 0,
                     ctx->stream, v, r, L, n, pB, rOut, vOut);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

template <typename T>
int lists_to_csc(Ctx* ctx, const T* v, const int* r, const int* count, int L,
                 int n, int* pB, int* rOut, T* vOut, long long* nnz) {
  int rc = csc_count_scan<T>(ctx, count, n, pB, nnz);
  if (rc) return rc;
  return csc_fill<T>(ctx, v, r, L, n, pB, rOut, vOut);
}

template int csc_count_scan<float>(Ctx*, const int*, int, int*, long long*);
template int csc_count_scan<double>(Ctx*, const int*, int, int*, long long*);
template int csc_fill<float>(Ctx*, const float*, const int*, int, int, const int*, int*, float*);
template int csc_fill<double>(Ctx*, const double*, const int*, int, int, const int*, int*, double*);
template int lists_to_csc<float>(Ctx*, const float*, const int*, const int*, int, int, int*, int*, float*, long long*);
template int lists_to_csc<double>(Ctx*, const double*, const int*, const int*, int, int, int*, int*, double*, long long*);

}  // namespace spx
