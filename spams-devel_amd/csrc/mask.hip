// Set-up of the masked-signal solvers (lasso_mask, decomp/decomp.h:1466-1561;
// omp_mask, :406-519): per signal the count of observed rows (0 = none, m =
// all, else some: the solvers' three cases), the signal with its unobserved
// rows zeroed (so that D'x and ||x||^2 are those of the masked copies,
// Vector::copyMask, linalg/linalg.h:4080-4090) and, for a partly observed
// signal, the constraint scaled by the observed fraction (constraint * |M| /
// m, decomp/decomp.h:1534 and :491).
#include "mask_gram.h"
#include "spx_common.h"
#include "spx_internal.h"

namespace spx {

namespace {

// one wavefront per signal
template <typename T>
__global__ __launch_bounds__(256) void mask_prep_kernel(
    const T* __restrict__ X, int ldx, int m, int n, const unsigned char* __restrict__ B, int ldb,
    T* __restrict__ Xm, int* __restrict__ obs, int* __restrict__ rows,
    const T* __restrict__ s_in, int ns, T* __restrict__ s_out, int scale) {
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (col >= n) return;
  const T* x = X + (size_t)col * ldx;
  const unsigned char* b = B + (size_t)col * ldb;
  T* xm = Xm + (size_t)col * m;
  int cnt = 0;
  for (int r = lane; r < m; r += 64) {
    const bool o = b[r] != 0;
    xm[r] = o ? x[r] : T(0);
    cnt += o ? 1 : 0;
  }
  cnt = static_cast<int>(wave_sum(static_cast<float>(cnt)));  // exact below 2^24
  // the rows the Gram columns visit, ascending: the observed ones, or the
  // unobserved ones when those are fewer
  const bool comp = masked_gram_comp(cnt, m);
  int* rl = rows + (size_t)col * m;
  int base = 0;
  for (int r0 = 0; r0 < m; r0 += 64) {
    const int r = r0 + lane;
    const bool sel = r < m && ((b[r] != 0) != comp);
    const unsigned long long bal = __ballot(sel);
    if (sel) rl[base + __popcll(bal & ((1ull << lane) - 1))] = r;
    base += __popcll(bal);
  }
  if (lane == 0) {
    obs[col] = cnt;
    if (s_out) {
      const T v = ns == 0 ? T(0) : s_in[ns > 1 ? col : 0];
      s_out[col] = (scale && cnt > 0 && cnt < m) ? v * T(cnt) / T(m) : v;
    }
  }
}

}  // namespace

template <typename T>
int mask_prep(hipStream_t st, const T* X, int ldx, int m, int n, const unsigned char* B, int ldb,
              T* Xm, int* obs, int* rows, const T* s_in, int ns, T* s_out, int scale) {
  if (n <= 0) return SPX_OK;
  hipLaunchKernelGGL(mask_prep_kernel<T>, dim3((n + 3) / 4), dim3(256), 0, st, X, ldx, m, n, B,
                     ldb, Xm, obs, rows, s_in, ns, s_out, scale);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

template int mask_prep<float>(hipStream_t, const float*, int, int, int, const unsigned char*, int,
                              float*, int*, int*, const float*, int, float*, int);
template int mask_prep<double>(hipStream_t, const double*, int, int, int, const unsigned char*,
                               int, double*, int*, int*, const double*, int, double*, int);

}  // namespace spx
