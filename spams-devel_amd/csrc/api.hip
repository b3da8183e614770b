// C ABI of the engine (declared in include/spx.h).
//
// Host entry points follow the reference facade (cpp_library/spams.h):
// argument normalisation identical to _lassoD / _omp / _fistaFlat, then
// upload, device products (D'X, ||x||^2) on MFMA, the solver kernel, device
// CSC assembly and download.  No CPU arithmetic path exists: every result is
// produced by the HIP kernels of this library.
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>

#include "spx_common.h"
#include "spx_internal.h"

namespace spx {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

namespace {

// Makes ctx's device current for the lifetime of the guard.
struct DeviceGuard {
  int prev = -1;
  bool ok = true;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev)
      (void)hipSetDevice(prev);
  }
};

__global__ void zero_ints_kernel(int* p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <typename T>
__global__ void transpose_kernel(const T* __restrict__ A, int m, int n,
                                 T* __restrict__ B) {
  // B (n x m) = A' for A m x n, both column-major
  __shared__ T tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int i = bx + tx, j = by + k;
    tile[k][tx] = (i < m && j < n) ? A[(size_t)i + (size_t)j * m] : T(0);
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int j = by + tx, i = bx + k;
    if (i < m && j < n) B[(size_t)j + (size_t)i * n] = tile[tx][k];
  }
}

// part[s][i] = sum over the columns j of slice s of G(i, j) x(j), for a p x p
// column-major G (leading dimension ld): block (b, s) takes rows 64b..64b+63
// and the slice's columns, its four waves every fourth column, combined in
// wave order (deterministic).
template <typename T>
__global__ __launch_bounds__(256) void gram_gemv_part_kernel(int p, int cs,
                                                             const T* __restrict__ G, int ld,
                                                             const T* __restrict__ x,
                                                             double* __restrict__ part) {
  __shared__ double red[4][64];
  const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + r;
  const int c0 = blockIdx.y * cs;
  const int c1 = c0 + cs < p ? c0 + cs : p;
  double acc = 0;
  if (i < p)
    for (int j = c0 + q; j < c1; j += 4) acc += static_cast<double>(G[i + (size_t)j * ld]) * x[j];
  red[q][r] = acc;
  __syncthreads();
  if (q == 0 && i < p)
    part[(size_t)blockIdx.y * p + i] = ((red[0][r] + red[1][r]) + red[2][r]) + red[3][r];
}

// y = scale * (sum over the S slices of part, in slice order)
template <typename T>
__global__ __launch_bounds__(256) void gram_gemv_sum_kernel(int p, int S,
                                                            const double* __restrict__ part,
                                                            const double* __restrict__ nrm,
                                                            T* __restrict__ y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p) return;
  double acc = 0;
  for (int z = 0; z < S; ++z) acc += part[(size_t)z * p + i];
  const double sc = nrm ? 1.0 / *nrm : 1.0;
  y[i] = static_cast<T>(acc * sc);
}

template <typename T>
__global__ __launch_bounds__(1024) void norm2_kernel(int p, const T* __restrict__ y,
                                                     double* __restrict__ out) {
  __shared__ double red[1024];
  double s = 0;
  for (int i = threadIdx.x; i < p; i += 1024) s += static_cast<double>(y[i]) * y[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sqrt(red[0]);
}

// G(j, j) += v for j < p (leading dimension ldg)
template <typename T>
__global__ void add_diag_kernel(T* G, int ldg, int p, T v) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < p) G[(size_t)j * ldg + j] += v;
}

// G(i, j) = G(j, i) for i > j: the Gram matrix exactly symmetric whatever
// the product's rounding (the BF16x3 split rounds G(i,j) and G(j,i) apart),
// so a solver may read G(c, k) from column k, which it has just streamed,
// instead of from column c.
template <typename T>
__global__ void sym_lower_kernel(T* G, int ldg, int p) {
  // column j = blockIdx.x, rows i of the 256-row tile blockIdx.y
  const int j = blockIdx.x;
  const int i = blockIdx.y * 256 + threadIdx.x;
  if (i < p && i > j) G[(size_t)j * ldg + i] = G[(size_t)i * ldg + j];
}

// dst = src + v I (p x p, leading dimension ldg, padding rows included)
template <typename T>
static int gram_with_ridge(hipStream_t st, const T* src, T* dst, int ldg, int p, T v) {
  SPX_HIP(hipMemcpyAsync(dst, src, sizeof(T) * (size_t)ldg * p, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(add_diag_kernel<T>, dim3((p + 255) / 256), dim3(256), 0, st, dst, ldg, p, v);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

template <typename T>
__global__ void fill_kernel(T* v, int n, T val) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) v[i] = val;
}

}  // namespace

template <typename T>
static int lipschitz(spx_ctx* ctx, const spx_dict* d, int iters, double* L) {
  const int p = d->p;
  // G x in column slices of at least 128 columns, at most 16 (the chip then
  // holds (p / 64) x 16 workgroups)
  int S = (p + 127) / 128;
  if (S > 16) S = 16;
  if (S < 1) S = 1;
  const int cs = (p + S - 1) / S;
  void* b = nullptr;
  int rc = ctx->scratch_get(12, 2 * sizeof(T) * (size_t)p + sizeof(double) * ((size_t)S * p + 1) + 64,
                            &b);
  if (rc) return rc;
  T* x = static_cast<T*>(b);
  T* y = x + p;
  double* nrm = reinterpret_cast<double*>(
      (reinterpret_cast<uintptr_t>(y + p) + 15) & ~uintptr_t(15));
  double* part = nrm + 1;
  hipStream_t st = ctx->stream;
  const dim3 g((p + 255) / 256);
  hipLaunchKernelGGL(fill_kernel<T>, g, dim3(256), 0, st, x, p, T(1));
  hipLaunchKernelGGL(norm2_kernel<T>, dim3(1), dim3(1024), 0, st, p, x, nrm);
  for (int it = 0; it < iters; ++it) {
    // x <- G (x / ||x||), then ||x|| -> the eigenvalue estimate
    hipLaunchKernelGGL(gram_gemv_part_kernel<T>, dim3((p + 63) / 64, S), dim3(256), 0, st, p,
                       cs, static_cast<const T*>(d->G), d->ldg, x, part);
    hipLaunchKernelGGL(gram_gemv_sum_kernel<T>, g, dim3(256), 0, st, p, S,
                       static_cast<const double*>(part), nrm, y);
    hipLaunchKernelGGL(norm2_kernel<T>, dim3(1), dim3(1024), 0, st, p, y, nrm);
    T* t = x;
    x = y;
    y = t;
  }
  SPX_HIP(hipGetLastError());
  SPX_HIP(hipMemcpyAsync(L, nrm, sizeof(double), hipMemcpyDeviceToHost, st));
  SPX_HIP(hipStreamSynchronize(st));
  return SPX_OK;
}

// Clears n ints with a kernel (capture-safe replacement for a memset node).
int zero_ints(hipStream_t st, int* p, int n) {
  if (n <= 0) return SPX_OK;
  hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

int tuned_waves(const Dict* d, int solver) {
  if (d->waves[solver] > 0) return d->waves[solver];
  const int per_wave = 64 * 16 * (16 / d->elem);
  if (d->p <= per_wave) return 1;
  if (d->p <= 4 * per_wave) return 4;
  return 16;
}

// ------------------------------------------------------------------ timers

ScopedTimer::ScopedTimer(Ctx* ctx, int id) : ctx_(ctx), id_(id) {
  if (!ctx_->timing || id_ < 0 || ctx_->arena) return;
  a_ = ctx_->event_get();
  if (a_ && hipEventRecord(a_, ctx_->stream) != hipSuccess) a_ = nullptr;
}

ScopedTimer::~ScopedTimer() {
  if (!a_) return;
  hipEvent_t b = ctx_->event_get();
  if (!b || hipEventRecord(b, ctx_->stream) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(ctx_->mu);
  ctx_->pending.push_back({id_, a_, b});
}

template <typename T>
static int dict_create(spx_ctx* ctx, const T* D, int m, int p, int on_device,
                       int precision, spx_dict** out) {
  if (!ctx || !out || !D || m <= 0 || p <= 0) {
    set_error("dict_create: invalid arguments (m=%d, p=%d)", m, p);
    return SPX_ERR_ARG;
  }
  DeviceGuard g(ctx->device);
  spx_dict* d = new spx_dict();
  d->ctx = ctx;
  d->device = ctx->device;
  d->m = m;
  d->p = p;
  d->ldg = pad_ld(p);
  d->elem = sizeof(T);
  d->precision = precision;
  hipStream_t st = ctx->stream;
  const size_t nD = (size_t)m * p;
  if (hipMalloc(&d->D, nD * sizeof(T)) != hipSuccess ||
      hipMalloc(&d->Dt, nD * sizeof(T)) != hipSuccess ||
      hipMalloc(&d->G, (size_t)d->ldg * p * sizeof(T)) != hipSuccess ||
      hipMalloc(&d->Gr, (size_t)d->ldg * p * sizeof(T)) != hipSuccess) {
    spx_dict_destroy(d);
    set_error("dict_create: device allocation failed (p=%d)", p);
    return SPX_ERR_NOMEM;
  }
  SPX_HIP(hipMemcpyAsync(d->D, D, nD * sizeof(T),
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  dim3 tg((m + 31) / 32, (p + 31) / 32);
  hipLaunchKernelGGL(transpose_kernel<T>, tg, dim3(256), 0, st,
                     static_cast<const T*>(d->D), m, p, static_cast<T*>(d->Dt));
  SPX_HIP(hipGetLastError());
  // zero the padding rows once, then G = D'D on MFMA (Matrix::XtX)
  SPX_HIP(hipMemsetAsync(d->G, 0, (size_t)d->ldg * p * sizeof(T), st));
  int rc = gemm_tn<T>(st, p, p, m, static_cast<const T*>(d->D), m,
                      static_cast<const T*>(d->D), m, static_cast<T*>(d->G),
                      d->ldg, T(1), T(0), nullptr, 0, precision);
  if (!rc) {
    // a 2-D grid: p columns x row tiles (p * p threads in one dimension
    // would pass 2^32 at p = 65536)
    hipLaunchKernelGGL(sym_lower_kernel<T>, dim3((unsigned)p, (unsigned)((p + 255) / 256)),
                       dim3(256), 0, st, static_cast<T*>(d->G), d->ldg, p);
    SPX_HIP(hipGetLastError());
  }
  if (!rc)
    rc = gram_with_ridge<T>(st, static_cast<const T*>(d->G), static_cast<T*>(d->Gr), d->ldg, p,
                            static_cast<T>(1e-10));
  if (rc) {
    spx_dict_destroy(d);
    return rc;
  }
  SPX_HIP(hipStreamSynchronize(st));
  *out = d;
  return SPX_OK;
}

// A resident Gram matrix without a dictionary (the Gram-given form of the
// reference, _lassoQq / lasso2(X, Q, q, ...), cpp_library/spams.h:145-188):
// G = Q (p x p, leading dimension p), used as given (no diagonal ridge);
// m is the signal length (for the facade's L clamp).
template <typename T>
static int dict_create_gram(spx_ctx* ctx, const T* Q, int m, int p, int on_device,
                            spx_dict** out) {
  if (!ctx || !out || !Q || m <= 0 || p <= 0) {
    set_error("dict_create_gram: invalid arguments (m=%d, p=%d)", m, p);
    return SPX_ERR_ARG;
  }
  DeviceGuard g(ctx->device);
  spx_dict* d = new spx_dict();
  d->ctx = ctx;
  d->device = ctx->device;
  d->m = m;
  d->p = p;
  d->ldg = pad_ld(p);
  d->elem = sizeof(T);
  d->gram_only = 1;
  hipStream_t st = ctx->stream;
  if (hipMalloc(&d->G, (size_t)d->ldg * p * sizeof(T)) != hipSuccess) {
    spx_dict_destroy(d);
    set_error("dict_create_gram: device allocation failed (p=%d)", p);
    return SPX_ERR_NOMEM;
  }
  SPX_HIP(hipMemsetAsync(d->G, 0, (size_t)d->ldg * p * sizeof(T), st));
  SPX_HIP(hipMemcpy2DAsync(d->G, sizeof(T) * d->ldg, Q, sizeof(T) * p, sizeof(T) * p, p,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  SPX_HIP(hipStreamSynchronize(st));
  *out = d;
  return SPX_OK;
}

static int need_dictionary(const spx_dict* d) {
  if (d && d->gram_only) {
    set_error("this entry point needs the dictionary D; the object holds only a Gram matrix");
    return SPX_ERR_ARG;
  }
  return SPX_OK;
}

// D'X and ||x||^2 for n signals into context scratch; returns pointers.
template <typename T>
static int products(spx_ctx* ctx, const spx_dict* d, const T* X, int ldx, int n,
                    int precision, T** DtX, T** norms) {
  if (int rc = need_dictionary(d)) return rc;
  void* b1 = nullptr;
  void* b2 = nullptr;
  int rc = ctx->scratch_get(3, sizeof(T) * (size_t)d->ldg * (n > 0 ? n : 1), &b1);
  if (rc) return rc;
  rc = ctx->scratch_get(4, sizeof(T) * (size_t)(n > 0 ? n : 1), &b2);
  if (rc) return rc;
  *DtX = static_cast<T*>(b1);
  *norms = static_cast<T*>(b2);
  rc = gemm_tn<T>(ctx->stream, d->p, n, d->m, static_cast<const T*>(d->D), d->m,
                  X, ldx, *DtX, d->ldg, T(1), T(0), nullptr, 0, precision);
  if (rc) return rc;
  return col_sqnorms<T>(ctx->stream, d->m, n, X, ldx, *norms);
}

// Masked-signal inputs of the solvers (device): count of observed rows per
// signal and the rows its masked Gram columns visit (m x n).
struct MaskIn {
  const int* obs;
  const int* rows;
};

// W (device, p x n, leading dimension ldw) selects the weighted penalty of
// lassoWeight / coreLARS2W (decomp/decomp.h:1315-1391, :1917-2107); mi the
// masked signals of lasso_mask (decomp/decomp.h:1466-1561).
template <typename T>
static int lasso_dtx_dev(spx_ctx* ctx, const spx_dict* d, const T* DtX, int ldd,
                         const T* norms, int n, const T* lambda, int nlam, T lambda2,
                         const spx_lasso_opts* o, spx_lists_dev* out,
                         const T* W = nullptr, int ldw = 0, const MaskIn* mi = nullptr) {
  if (!ctx || !d || !o || !out || n < 0 || (n > 0 && (!DtX || !norms || !lambda)) ||
      ldd < d->p || !(nlam == 1 || nlam == n)) {
    set_error("lasso_dev: invalid arguments");
    return SPX_ERR_ARG;
  }
  if ((int)sizeof(T) != d->elem) {
    set_error("lasso_dev: dictionary precision does not match the entry point");
    return SPX_ERR_ARG;
  }
  if (o->mode < 0 || o->mode > 2) {
    set_error("lasso: unsupported mode %d", o->mode);
    return SPX_ERR_UNSUP;
  }
  if (out->L <= 0) {
    set_error("lasso_dev: lists have L=%d", out->L);
    return SPX_ERR_ARG;
  }
  if (W && (ldw < d->p || o->ols || (o->precision & SPX_PREC_REFIT64) ||
            (out->path && out->path_cols > 0))) {
    set_error("lassoWeighted: invalid arguments (ldw < p, or ols / refit / path requested)");
    return SPX_ERR_ARG;
  }
  if (mi && (o->ols || (o->precision & SPX_PREC_REFIT64) || (out->path && out->path_cols > 0))) {
    set_error("lassoMask: ols / refit / path are not options of the masked solver");
    return SPX_ERR_ARG;
  }
  if (n == 0) return SPX_OK;
  DeviceGuard g(ctx->device);
  LarsArgs<T> a{};
  a.tol = T(1e-15);
  a.ldg = d->ldg;
  a.gsym = !d->gram_only;  // the engine's Gram matrices are exactly symmetric
  // G + max(lambda2, 1e-10) I (decomp/decomp.h:1566, ProdMatrix::addDiag);
  // a Gram given as such is used unchanged (decomp/decomp.h:1574).  The
  // dictionary holds G + 1e-10 I; another ridge is formed in scratch.
  a.diag = d->gram_only ? T(0) : static_cast<T>(lambda2 > T(1e-10) ? double(lambda2) : 1e-10);
  if (d->gram_only || W) {
    // lassoWeight's ridge is 1e-10 (decomp/decomp.h:1337), set below
    a.G = static_cast<const T*>(d->gram_only ? d->G : d->Gr);
  } else if (!(lambda2 > T(1e-10))) {
    a.G = static_cast<const T*>(d->Gr);
  } else {
    T* Gl = nullptr;
    int rc = ctx->scratch_get(22, sizeof(T) * (size_t)d->ldg * d->p, reinterpret_cast<void**>(&Gl));
    if (rc) return rc;
    rc = gram_with_ridge<T>(ctx->stream, static_cast<const T*>(d->G), Gl, d->ldg, d->p, a.diag);
    if (rc) return rc;
    a.G = Gl;
  }
  a.DtX = DtX;
  a.ldd = ldd;
  a.norms = norms;
  a.lambda = lambda;
  a.nlam = nlam;
  a.p = d->p;
  a.n = n;
  a.L = out->L;
  a.mode = o->mode;
  a.pos = o->pos;
  const int Lp = o->L < 0 ? d->p : o->L;
  a.length_path = o->max_length_path < 0 ? 4 * Lp : o->max_length_path;
  // coreLARS2: length_path <= 1 means 4 * min(L, p) (decomp/decomp.h:1724)
  if (a.length_path <= 1) a.length_path = 4 * (a.L < a.p ? a.L : a.p);
  a.outV = static_cast<T*>(out->v);
  a.outR = out->r;
  a.outCount = out->count;
  a.stats = out->stats;
  a.refit = (o->precision & SPX_PREC_REFIT64) ? 1 : 0;
  a.ols = o->ols ? 1 : 0;
  // cholesky selects the path of coreLARS (decomp/decomp.h:844-1093), and
  // ols implies it (cpp_library/spams.h:136-137); with ols that path tracks
  // the least-squares residual of the active set (read by the L2ERROR and
  // normX stops) and has no exit on a residual increase.  The weighted and
  // masked solvers have no Cholesky variant (coreLARS2W, coreLARS2 with a
  // mask): the option is read by the plain solver only.
  a.chol = ((o->cholesky || o->ols) && !W && !mi) ? 1 : 0;
  if (out->path && out->path_cols > 0) {
    a.path = static_cast<T*>(out->path);
    a.path_cols = out->path_cols;
    // cleared by a kernel (no memset node when a caller captures the call)
    const size_t np = (size_t)d->p * out->path_cols;
    hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)((np + 255) / 256)), dim3(256), 0,
                       ctx->stream, static_cast<T*>(out->path), (int)np, T(0));
    SPX_HIP(hipGetLastError());
  }
  if (W) {
    // lassoWeight: G + 1e-10 I (decomp/decomp.h:1337), tolerance 1e-10 and
    // the step cap iter > 4 min(L, p) (decomp/decomp.h:2099)
    a.weights = W;
    a.ldw = ldw;
    a.diag = d->gram_only ? T(0) : T(1e-10);
    a.tol = T(1e-10);
    a.length_path = 4 * (a.L < a.p ? a.L : a.p) + 2;
  }
  if (mi) {
    // coreLARS2's default step cap 4 min(L, p) (decomp/decomp.h:1724)
    a.mobs = mi->obs;
    a.mrows = mi->rows;
    a.Dd = static_cast<const T*>(d->D);
    a.Dt = static_cast<const T*>(d->Dt);
    a.m = d->m;
    a.length_path = 4 * (a.L < a.p ? a.L : a.p);
  }
  const int w = o->waves_per_signal > 0 ? o->waves_per_signal
                : o->form == SPX_FORM_PLAIN ? 4 : tuned_waves(d, 0);
  return lars_solve<T>(ctx, a, o->form, w);
}

template <typename T>
static int lasso_dev(spx_ctx* ctx, const spx_dict* d, const T* X, int ldx,
                     int n, const T* lambda, int nlam, T lambda2,
                     const spx_lasso_opts* o, spx_lists_dev* out,
                     const T* W = nullptr, int ldw = 0) {
  if (!ctx || !d || !o || !out || n < 0 || (n > 0 && (!X || !lambda)) ||
      ldx < d->m || !(nlam == 1 || nlam == n)) {
    set_error("lasso_dev: invalid arguments");
    return SPX_ERR_ARG;
  }
  if ((int)sizeof(T) != d->elem) {
    set_error("lasso_dev: dictionary precision does not match the entry point");
    return SPX_ERR_ARG;
  }
  if (n == 0) return SPX_OK;
  DeviceGuard g(ctx->device);
  T* DtX = nullptr;
  T* norms = nullptr;
  int rc = products<T>(ctx, d, X, ldx, n, o->precision & SPX_PREC_BF16X3, &DtX, &norms);
  if (rc) return rc;
  return lasso_dtx_dev<T>(ctx, d, DtX, d->ldg, norms, n, lambda, nlam, lambda2, o, out, W, ldw);
}

template <typename T>
static int omp_solve_dtx(spx_ctx* ctx, const spx_dict* d, const T* DtX, const T* norms, int n,
                         const int* L, int nL, const T* eps, int nEps, const T* lambda,
                         int nLambda, const spx_omp_opts* o, spx_lists_dev* out,
                         const MaskIn* mi = nullptr);

template <typename T>
// B (device m x n bytes, leading dimension ldb, nonzero = observed): the
// masked signals of omp_mask (decomp/decomp.h:406-519), or null.
static int omp_dev(spx_ctx* ctx, const spx_dict* d, const T* X, int ldx, int n,
                   const int* L, int nL, const T* eps, int nEps,
                   const T* lambda, int nLambda, const spx_omp_opts* o,
                   spx_lists_dev* out, const unsigned char* B = nullptr, int ldb = 0) {
  if (!ctx || !d || !out || n < 0 || ldx < d->m || (B && ldb < d->m)) {
    set_error("omp_dev: invalid arguments");
    return SPX_ERR_ARG;
  }
  if ((int)sizeof(T) != d->elem) {
    set_error("omp_dev: dictionary precision does not match the entry point");
    return SPX_ERR_ARG;
  }
  if (!L) nL = 0;
  if (!eps) nEps = 0;
  if (!lambda) nLambda = 0;
  if (nL == 0 && nEps == 0 && nLambda == 0) {
    // cpp_library/spams.h:267-268
    set_error("omp : You should either provide L, eps or lambda");
    return SPX_ERR_ARG;
  }
  if ((nL > 1 && nL != n) || (nEps > 1 && nEps != n) || (nLambda > 1 && nLambda != n)) {
    set_error("omp: per-signal vectors must have n entries");
    return SPX_ERR_ARG;
  }
  if (n == 0 || out->L <= 0) return SPX_OK;
  DeviceGuard g(ctx->device);
  spx_omp_opts def{SPX_PREC_FP32, SPX_FORM_BATCHED, 0};
  if (!o) o = &def;
  T* DtX = nullptr;
  T* norms = nullptr;
  if (B) {
    // masked copies, observed counts, row lists and eps * |M| / m
    // (decomp/decomp.h:491)
    if (int rc = need_dictionary(d)) return rc;
    const int m = d->m;
    void *bx = nullptr, *bc = nullptr, *be = nullptr;
    int rc = ctx->scratch_get(18, sizeof(T) * (size_t)m * n, &bx);
    if (!rc) rc = ctx->scratch_get(19, sizeof(int) * (size_t)n * (m + 1), &bc);
    if (!rc) rc = ctx->scratch_get(20, sizeof(T) * (size_t)n, &be);
    if (rc) return rc;
    T* Xm = static_cast<T*>(bx);
    int* obs = static_cast<int*>(bc);
    int* rows = obs + n;
    T* epsv = static_cast<T*>(be);
    rc = mask_prep<T>(ctx->stream, X, ldx, m, n, B, ldb, Xm, obs, rows, eps, nEps, epsv, 1);
    if (!rc) rc = products<T>(ctx, d, Xm, m, n, o->precision, &DtX, &norms);
    if (rc) return rc;
    const MaskIn mi{obs, rows};
    return omp_solve_dtx<T>(ctx, d, DtX, norms, n, L, nL, nEps ? epsv : nullptr, nEps ? n : 0,
                            lambda, nLambda, o, out, &mi);
  }
  int rc = products<T>(ctx, d, X, ldx, n, o->precision, &DtX, &norms);
  if (rc) return rc;
  return omp_solve_dtx<T>(ctx, d, DtX, norms, n, L, nL, eps, nEps, lambda, nLambda, o, out);
}

// ORMP from the correlations D'X (leading dimension ldg) and ||x||^2 after
// omp_dev's argument checks.
template <typename T>
static int omp_solve_dtx(spx_ctx* ctx, const spx_dict* d, const T* DtX, const T* norms, int n,
                         const int* L, int nL, const T* eps, int nEps, const T* lambda,
                         int nLambda, const spx_omp_opts* o, spx_lists_dev* out,
                         const MaskIn* mi) {
  OmpArgs<T> a{};
  if (mi) {
    a.mobs = mi->obs;
    a.mrows = mi->rows;
    a.Dd = static_cast<const T*>(d->D);
    a.Dt = static_cast<const T*>(d->Dt);
    a.m = d->m;
  }
  a.G = static_cast<const T*>(d->G);
  a.gsym = !d->gram_only;
  a.ldg = d->ldg;
  a.DtX = DtX;
  a.ldd = d->ldg;
  a.norms = norms;
  a.Ls = L;
  a.nL = nL;
  a.eps = eps;
  a.nEps = nEps;
  a.lambda = lambda;
  a.nLambda = nLambda;
  a.scalarL = d->m < d->p ? d->m : d->p;
  a.p = d->p;
  a.n = n;
  a.L = out->L;
  a.outV = static_cast<T*>(out->v);
  a.outR = out->r;
  a.outCount = out->count;
  if (out->path && out->path_cols > 0) {
    a.path = static_cast<T*>(out->path);
    a.path_cols = out->path_cols;
    SPX_HIP(hipMemsetAsync(out->path, 0, sizeof(T) * (size_t)d->p * out->path_cols, ctx->stream));
  }
  const int w = o->waves_per_signal > 0 ? o->waves_per_signal
                : o->form == SPX_FORM_PLAIN ? 4 : tuned_waves(d, 1);
  return omp_solve<T>(ctx, a, o->form, w);
}

// ------------------------------------------------------------- autotuner

constexpr int kTuneSignals = 65536;
constexpr int kTuneWidths[4] = {1, 2, 4, 16};

// Device lists for the autotuner's trial runs (freed on scope exit).
template <typename T>
struct TrialLists {
  spx_lists_dev l{};
  int alloc(int L, int n) {
    l.L = L;
    SPX_HIP(hipMalloc(&l.v, sizeof(T) * (size_t)L * n));
    SPX_HIP(hipMalloc(reinterpret_cast<void**>(&l.r), sizeof(int) * (size_t)L * n));
    SPX_HIP(hipMalloc(reinterpret_cast<void**>(&l.count), sizeof(int) * (size_t)n));
    return SPX_OK;
  }
  ~TrialLists() {
    if (l.v) (void)hipFree(l.v);
    if (l.r) (void)hipFree(l.r);
    if (l.count) (void)hipFree(l.count);
  }
};

// Times run(w) for each width that covers p (one warm-up launch, then the
// best of two timed launches on the context's stream) and stores the fastest
// width on the dictionary.  A wider width must beat the narrower best by more
// than 3 %: narrower groups keep more signals in flight, and a full batch
// hides the grid's tail that the sample shows (m=256, p=1024, n=200000: the
// sample put 1 and 2 wavefronts within 0.6 %, the batch ran 1.4x faster at 1).
template <typename Run>
static int tune_widths(spx_ctx* ctx, spx_dict* d, int solver, bool (*fits)(int, int), Run run,
                       int* waves, float* ms_out) {
  const bool timing = ctx->timing;
  ctx->timing = false;
  hipEvent_t e0 = ctx->event_get(), e1 = ctx->event_get();
  int rc = (e0 && e1) ? SPX_OK : SPX_ERR_HIP;
  if (rc) set_error("autotune: event creation failed");
  float ms[4] = {0, 0, 0, 0};
  float best = 0.f;
  int bw = 0;
  for (int i = 0; i < 4 && rc == SPX_OK; ++i) {
    const int w = kTuneWidths[i];
    if (!fits(w, d->p)) continue;
    rc = run(w);
    if (rc) break;
    float t = 0.f;
    for (int rep = 0; rep < 2 && rc == SPX_OK; ++rep) {
      float x = 0.f;
      if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = SPX_ERR_HIP;
      if (!rc) rc = run(w);
      if (!rc && (hipEventRecord(e1, ctx->stream) != hipSuccess ||
                  hipEventSynchronize(e1) != hipSuccess ||
                  hipEventElapsedTime(&x, e0, e1) != hipSuccess)) {
        set_error("autotune: event timing failed");
        rc = SPX_ERR_HIP;
      }
      t = rep == 0 ? x : (x < t ? x : t);
    }
    ms[i] = t;
    if (!rc && (bw == 0 || t < 0.97f * best)) {
      best = t;
      bw = w;
    }
  }
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (e0) ctx->free_events.push_back(e0);
    if (e1) ctx->free_events.push_back(e1);
  }
  ctx->timing = timing;
  if (rc) return rc;
  if (bw == 0) {
    set_error("autotune: p=%d exceeds every supported width", d->p);
    return SPX_ERR_UNSUP;
  }
  d->waves[solver] = bw;
  for (int i = 0; i < 4; ++i) d->tune_ms[solver][i] = ms[i];
  if (waves) *waves = bw;
  if (ms_out)
    for (int i = 0; i < 4; ++i) ms_out[i] = ms[i];
  return SPX_OK;
}

template <typename T>
static int lasso_autotune(spx_ctx* ctx, spx_dict* d, const T* X, int ldx, int n,
                          const T* lambda, int nlam, T lambda2, const spx_lasso_opts* o,
                          int* waves, float* ms) {
  if (!ctx || !d || !o || !X || !lambda || n <= 0 || ldx < d->m ||
      !(nlam == 1 || nlam == n) || (int)sizeof(T) != d->elem) {
    set_error("autotune_lasso: invalid arguments");
    return SPX_ERR_ARG;
  }
  DeviceGuard g(ctx->device);
  const int ns = n < kTuneSignals ? n : kTuneSignals;
  int L = o->L < 0 ? d->p : o->L;
  if (L > d->m) L = d->m;
  if (L > d->p) L = d->p;
  if (L < 1) L = 1;
  TrialLists<T> tl;
  int rc = tl.alloc(L, ns);
  if (rc) return rc;
  T* DtX = nullptr;
  T* norms = nullptr;
  rc = products<T>(ctx, d, X, ldx, ns, o->precision & SPX_PREC_BF16X3, &DtX, &norms);
  if (rc) return rc;
  spx_lasso_opts oo = *o;
  oo.form = SPX_FORM_BATCHED;
  auto run = [&](int w) {
    oo.waves_per_signal = w;
    return lasso_dtx_dev<T>(ctx, d, DtX, d->ldg, norms, ns, lambda, nlam == 1 ? 1 : ns,
                            lambda2, &oo, &tl.l);
  };
  return tune_widths(ctx, d, 0, &lars_fits<T>, run, waves, ms);
}

template <typename T>
static int omp_autotune(spx_ctx* ctx, spx_dict* d, const T* X, int ldx, int n, const int* L,
                        int nL, const T* eps, int nEps, const T* lambda, int nLambda,
                        int* waves, float* ms) {
  if (!ctx || !d || !X || n <= 0 || ldx < d->m || (int)sizeof(T) != d->elem) {
    set_error("autotune_omp: invalid arguments");
    return SPX_ERR_ARG;
  }
  if (!L) nL = 0;
  if (!eps) nEps = 0;
  if (!lambda) nLambda = 0;
  if ((nL == 0 && nEps == 0 && nLambda == 0) || (nL > 1 && nL != n) ||
      (nEps > 1 && nEps != n) || (nLambda > 1 && nLambda != n)) {
    set_error("autotune_omp: invalid stopping-rule vectors");
    return SPX_ERR_ARG;
  }
  DeviceGuard g(ctx->device);
  const int ns = n < kTuneSignals ? n : kTuneSignals;
  // list width: the largest L any signal may use (the scalar default when L
  // is not given), read back from the device vector
  int Lmax = d->m < d->p ? d->m : d->p;
  if (nL > 0) {
    std::vector<int> hl((size_t)(nL > 1 ? ns : 1));
    SPX_HIP(hipMemcpy(hl.data(), L, sizeof(int) * hl.size(), hipMemcpyDeviceToHost));
    Lmax = spx_omp_L(d->m, d->p, hl.data(), (int)hl.size(), nL > 1 ? ns : 1);
  }
  if (Lmax < 1) Lmax = 1;
  TrialLists<T> tl;
  int rc = tl.alloc(Lmax, ns);
  if (rc) return rc;
  T* DtX = nullptr;
  T* norms = nullptr;
  rc = products<T>(ctx, d, X, ldx, ns, SPX_PREC_FP32, &DtX, &norms);
  if (rc) return rc;
  spx_omp_opts oo{SPX_PREC_FP32, SPX_FORM_BATCHED, 0};
  auto run = [&](int w) {
    oo.waves_per_signal = w;
    return omp_solve_dtx<T>(ctx, d, DtX, norms, ns, L, nL > 1 ? ns : nL, eps,
                            nEps > 1 ? ns : nEps, lambda, nLambda > 1 ? ns : nLambda, &oo,
                            &tl.l);
  };
  return tune_widths(ctx, d, 1, &ormp_fits<T>, run, waves, ms);
}

// Device CSC of the lists, then download into a host spx_csc.
// lasso_mask (decomp/decomp.h:1466-1561): B (device, m x n bytes, leading
// dimension ldb; nonzero = observed).  A signal with every row observed runs
// the plain path; with none, its column is empty; otherwise coreLARS2 over
// its observed rows with the constraint scaled by |M| / m (PENALTY and
// L2ERROR) and the Gram columns of its rows, ridge max(lambda2, 1e-10).
template <typename T>
static int lasso_mask_dev(spx_ctx* ctx, const spx_dict* d, const T* X, int ldx, int n,
                          const unsigned char* B, int ldb, const T* lambda, int nlam, T lambda2,
                          const spx_lasso_opts* o, spx_lists_dev* out) {
  if (!ctx || !d || !o || !out || n < 0 || (n > 0 && (!X || !lambda || !B)) || ldx < d->m ||
      ldb < d->m || !(nlam == 1 || nlam == n)) {
    set_error("lassoMask: invalid arguments");
    return SPX_ERR_ARG;
  }
  if ((int)sizeof(T) != d->elem) {
    set_error("lasso_dev: dictionary precision does not match the entry point");
    return SPX_ERR_ARG;
  }
  if (int rc = need_dictionary(d)) return rc;
  if (n == 0) return SPX_OK;
  DeviceGuard g(ctx->device);
  const int m = d->m;
  void *bx = nullptr, *bc = nullptr, *bl = nullptr;
  int rc = ctx->scratch_get(18, sizeof(T) * (size_t)m * n, &bx);
  if (!rc) rc = ctx->scratch_get(19, sizeof(int) * (size_t)n * (m + 1), &bc);
  if (!rc) rc = ctx->scratch_get(20, sizeof(T) * (size_t)n, &bl);
  if (rc) return rc;
  T* Xm = static_cast<T*>(bx);
  int* obs = static_cast<int*>(bc);
  int* rows = obs + n;
  T* lamv = static_cast<T*>(bl);
  const int scale = o->mode == SPX_PENALTY || o->mode == SPX_L2ERROR;
  rc = mask_prep<T>(ctx->stream, X, ldx, m, n, B, ldb, Xm, obs, rows, lambda, nlam, lamv, scale);
  if (rc) return rc;
  T* DtX = nullptr;
  T* norms = nullptr;
  rc = products<T>(ctx, d, Xm, m, n, o->precision & SPX_PREC_BF16X3, &DtX, &norms);
  if (rc) return rc;
  const MaskIn mi{obs, rows};
  return lasso_dtx_dev<T>(ctx, d, DtX, d->ldg, norms, n, lamv, n, lambda2, o, out, nullptr, 0,
                          &mi);
}

template <typename T>
static int lists_to_host_csc(spx_ctx* ctx, const spx_lists_dev& l, int p, int n,
                             spx_csc* out) {
  hipStream_t st = ctx->stream;
  int* dpB = nullptr;
  void* b = nullptr;
  int rc = ctx->scratch_get(5, sizeof(int) * (size_t)(n + 1), &b);
  if (rc) return rc;
  dpB = static_cast<int*>(b);
  long long nnz = 0;
  rc = csc_count_scan<T>(ctx, l.count, n, dpB, &nnz);
  if (rc) return rc;
  void* b2 = nullptr;
  rc = ctx->scratch_get(6, (sizeof(int) + sizeof(T)) * (size_t)(nnz > 0 ? nnz : 1), &b2);
  if (rc) return rc;
  int* dr = static_cast<int*>(b2);
  T* dv = reinterpret_cast<T*>(static_cast<char*>(b2) + sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  dv = reinterpret_cast<T*>((reinterpret_cast<uintptr_t>(dv) + 7) & ~uintptr_t(7));
  rc = csc_fill<T>(ctx, static_cast<const T*>(l.v), l.r, l.L, n, dpB, dr, dv);
  if (rc) return rc;
  out->m = p;
  out->n = n;
  out->nnz = nnz;
  out->pB = static_cast<int*>(malloc(sizeof(int) * (size_t)(n + 1)));
  out->r = static_cast<int*>(malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1)));
  out->v = malloc(sizeof(T) * (size_t)(nnz > 0 ? nnz : 1));
  if (!out->pB || !out->r || !out->v) {
    set_error("out of host memory");
    return SPX_ERR_NOMEM;
  }
  SPX_HIP(hipMemcpyAsync(out->pB, dpB, sizeof(int) * (size_t)(n + 1), hipMemcpyDeviceToHost, st));
  if (nnz > 0) {
    SPX_HIP(hipMemcpyAsync(out->r, dr, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost, st));
    SPX_HIP(hipMemcpyAsync(out->v, dv, sizeof(T) * (size_t)nnz, hipMemcpyDeviceToHost, st));
  }
  SPX_HIP(hipStreamSynchronize(st));
  return SPX_OK;
}

// RAII device buffer for the host entry points.
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
      p = nullptr;
      set_error("device allocation of %zu bytes failed", bytes);
      return SPX_ERR_NOMEM;
    }
    return SPX_OK;
  }
};

// W: host p x n weights of the weighted penalty (_lassoWeighted), or null;
// B: host m x n observation mask of _lassoMask, or null.
template <typename T>
static int lasso_host(spx_ctx* ctx, const T* X, int m, int n, const T* D, int p,
                      const T* lambda, int nlam, T lambda2,
                      const spx_lasso_opts* opts, spx_csc* out, T* path = nullptr,
                      const T* W = nullptr, const unsigned char* B = nullptr) {
  if (!ctx || !out || !D || !lambda || m <= 0 || p <= 0 || n < 0 || (n > 0 && !X)) {
    set_error("lasso : invalid arguments");
    return SPX_ERR_ARG;
  }
  if (!(nlam == 1 || nlam == n)) {
    set_error("lasso : lambda must have 1 or n entries");
    return SPX_ERR_ARG;
  }
  memset(out, 0, sizeof(*out));
  spx_lasso_opts o{-1, SPX_PENALTY, 0, -1, SPX_PREC_FP32, SPX_FORM_BATCHED, 0};
  if (opts) o = *opts;
  const int Le = spx_lasso_L(m, p, o.L, o.mode, double(lambda[0]), o.pos, double(lambda2));
  if (o.max_length_path < 0) o.max_length_path = 4 * (o.L < 0 ? p : o.L);
  o.L = Le;
  DeviceGuard g(ctx->device);
  const int pcols = o.max_length_path > 0 ? o.max_length_path : 0;
  if (path && pcols > 0) memset(path, 0, sizeof(T) * (size_t)p * pcols);
  if (Le <= 0 || n == 0) {
    // decomp/decomp.h:1587: L <= 0 -> empty result
    out->m = p;
    out->n = n;
    out->nnz = 0;
    out->pB = static_cast<int*>(calloc((size_t)n + 1, sizeof(int)));
    out->r = static_cast<int*>(malloc(sizeof(int)));
    out->v = malloc(sizeof(T));
    return SPX_OK;
  }
  spx_dict* d = nullptr;
  int rc = dict_create<T>(ctx, D, m, p, 0, o.precision & SPX_PREC_BF16X3, &d);
  if (rc) return rc;
  DevBuf bX, bL, bV, bR, bC, bP, bW, bB;
  if ((rc = bX.alloc(sizeof(T) * (size_t)m * n)) || (rc = bL.alloc(sizeof(T) * nlam)) ||
      (W && (rc = bW.alloc(sizeof(T) * (size_t)p * n))) ||
      (B && (rc = bB.alloc((size_t)m * n))) ||
      (rc = bV.alloc(sizeof(T) * (size_t)Le * n)) ||
      (rc = bR.alloc(sizeof(int) * (size_t)Le * n)) || (rc = bC.alloc(sizeof(int) * (size_t)n)) ||
      (path && pcols > 0 && (rc = bP.alloc(sizeof(T) * (size_t)p * pcols)))) {
    spx_dict_destroy(d);
    return rc;
  }
  hipStream_t st = ctx->stream;
  if (hipMemcpyAsync(bX.p, X, sizeof(T) * (size_t)m * n, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(bL.p, lambda, sizeof(T) * nlam, hipMemcpyHostToDevice, st) != hipSuccess ||
      (W && hipMemcpyAsync(bW.p, W, sizeof(T) * (size_t)p * n, hipMemcpyHostToDevice, st) !=
                hipSuccess) ||
      (B && hipMemcpyAsync(bB.p, B, (size_t)m * n, hipMemcpyHostToDevice, st) != hipSuccess)) {
    spx_dict_destroy(d);
    set_error("lasso: upload failed");
    return SPX_ERR_HIP;
  }
  spx_lists_dev l{bV.p, static_cast<int*>(bR.p), static_cast<int*>(bC.p), Le, nullptr};
  if (path && pcols > 0) {
    l.path = bP.p;
    l.path_cols = pcols;
  }
  if (B)
    rc = lasso_mask_dev<T>(ctx, d, static_cast<const T*>(bX.p), m, n,
                           static_cast<const unsigned char*>(bB.p), m,
                           static_cast<const T*>(bL.p), nlam, lambda2, &o, &l);
  else
    rc = lasso_dev<T>(ctx, d, static_cast<const T*>(bX.p), m, n,
                      static_cast<const T*>(bL.p), nlam, lambda2, &o, &l,
                      static_cast<const T*>(bW.p), p);
  if (!rc) rc = lists_to_host_csc<T>(ctx, l, p, n, out);
  if (!rc && path && pcols > 0 &&
      hipMemcpy(path, bP.p, sizeof(T) * (size_t)p * pcols, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("lasso: path download failed");
    rc = SPX_ERR_HIP;
  }
  spx_dict_destroy(d);
  return rc;
}

// Gram-given LARS (_lassoQq, cpp_library/spams.h:145-188): X (m x n) only
// gives ||x_i||^2, Q (p x p) is the Gram matrix used as given (lambda2 is
// ignored except in the facade's L clamp), q (p x n) the correlations.
template <typename T>
static int lasso_qq_host(spx_ctx* ctx, const T* X, int m, int n, const T* Q, const T* q,
                         int p, const T* lambda, int nlam, T lambda2,
                         const spx_lasso_opts* opts, spx_csc* out, T* path) {
  if (!ctx || !out || !Q || !lambda || m <= 0 || p <= 0 || n < 0 || (n > 0 && (!X || !q))) {
    set_error("lasso : invalid arguments");
    return SPX_ERR_ARG;
  }
  if (!(nlam == 1 || nlam == n)) {
    set_error("lasso : lambda must have 1 or n entries");
    return SPX_ERR_ARG;
  }
  memset(out, 0, sizeof(*out));
  spx_lasso_opts o{-1, SPX_PENALTY, 0, -1, SPX_PREC_FP32, SPX_FORM_BATCHED, 0};
  if (opts) o = *opts;
  const int Le = spx_lasso_L(m, p, o.L, o.mode, double(lambda[0]), o.pos, double(lambda2));
  if (o.max_length_path < 0) o.max_length_path = 4 * (o.L < 0 ? p : o.L);
  o.L = Le;
  DeviceGuard g(ctx->device);
  const int pcols = o.max_length_path > 0 ? o.max_length_path : 0;
  if (path && pcols > 0) memset(path, 0, sizeof(T) * (size_t)p * pcols);
  if (Le <= 0 || n == 0) {
    out->m = p;
    out->n = n;
    out->pB = static_cast<int*>(calloc((size_t)n + 1, sizeof(int)));
    out->r = static_cast<int*>(malloc(sizeof(int)));
    out->v = malloc(sizeof(T));
    return SPX_OK;
  }
  spx_dict* d = nullptr;
  int rc = dict_create_gram<T>(ctx, Q, m, p, 0, &d);
  if (rc) return rc;
  DevBuf bX, bq, bN, bL, bV, bR, bC, bP;
  if ((rc = bX.alloc(sizeof(T) * (size_t)m * n)) || (rc = bq.alloc(sizeof(T) * (size_t)p * n)) ||
      (rc = bN.alloc(sizeof(T) * (size_t)n)) || (rc = bL.alloc(sizeof(T) * nlam)) ||
      (rc = bV.alloc(sizeof(T) * (size_t)Le * n)) ||
      (rc = bR.alloc(sizeof(int) * (size_t)Le * n)) || (rc = bC.alloc(sizeof(int) * (size_t)n)) ||
      (path && pcols > 0 && (rc = bP.alloc(sizeof(T) * (size_t)p * pcols)))) {
    spx_dict_destroy(d);
    return rc;
  }
  hipStream_t st = ctx->stream;
  if (hipMemcpyAsync(bX.p, X, sizeof(T) * (size_t)m * n, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(bq.p, q, sizeof(T) * (size_t)p * n, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(bL.p, lambda, sizeof(T) * nlam, hipMemcpyHostToDevice, st) != hipSuccess) {
    spx_dict_destroy(d);
    set_error("lasso: upload failed");
    return SPX_ERR_HIP;
  }
  rc = col_sqnorms<T>(st, m, n, static_cast<const T*>(bX.p), m, static_cast<T*>(bN.p));
  spx_lists_dev l{bV.p, static_cast<int*>(bR.p), static_cast<int*>(bC.p), Le, nullptr};
  if (path && pcols > 0) {
    l.path = bP.p;
    l.path_cols = pcols;
  }
  if (!rc)
    rc = lasso_dtx_dev<T>(ctx, d, static_cast<const T*>(bq.p), p, static_cast<const T*>(bN.p), n,
                          static_cast<const T*>(bL.p), nlam, lambda2, &o, &l);
  if (!rc) rc = lists_to_host_csc<T>(ctx, l, p, n, out);
  if (!rc && path && pcols > 0 &&
      hipMemcpy(path, bP.p, sizeof(T) * (size_t)p * pcols, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("lasso: path download failed");
    rc = SPX_ERR_HIP;
  }
  spx_dict_destroy(d);
  return rc;
}

template <typename T>
static int omp_host(spx_ctx* ctx, const T* X, int m, int n, const T* D, int p,
                    const int* L, int nL, const T* eps, int nEps, const T* lambda,
                    int nLambda, const spx_omp_opts* opts, spx_csc* out,
                    T* path = nullptr, const unsigned char* B = nullptr) {
  if (!ctx || !out || !D || m <= 0 || p <= 0 || n < 0 || (n > 0 && !X)) {
    set_error("omp : invalid arguments");
    return SPX_ERR_ARG;
  }
  memset(out, 0, sizeof(*out));
  // path: p x min(m, p), zero where never written (cpp_library/spams.h:283-285)
  const int pcols = m < p ? m : p;
  if (path) memset(path, 0, sizeof(T) * (size_t)p * pcols);
  if (!L) nL = 0;
  if (!eps) nEps = 0;
  if (!lambda) nLambda = 0;
  if (nL == 0 && nEps == 0 && nLambda == 0) {
    set_error("omp : You should either provide L, eps or lambda");
    return SPX_ERR_ARG;
  }
  if ((nL > 1 && nL != n) || (nEps > 1 && nEps != n) || (nLambda > 1 && nLambda != n)) {
    set_error("omp : per-signal vectors must have n entries");
    return SPX_ERR_ARG;
  }
  const int Le = spx_omp_L(m, p, L, nL, n);
  DeviceGuard g(ctx->device);
  if (Le <= 0 || n == 0) {
    out->m = p;
    out->n = n;
    out->pB = static_cast<int*>(calloc((size_t)n + 1, sizeof(int)));
    out->r = static_cast<int*>(malloc(sizeof(int)));
    out->v = malloc(sizeof(T));
    return SPX_OK;
  }
  spx_omp_opts o{SPX_PREC_FP32, SPX_FORM_BATCHED, 0};
  if (opts) o = *opts;
  spx_dict* d = nullptr;
  int rc = dict_create<T>(ctx, D, m, p, 0, o.precision, &d);
  if (rc) return rc;
  DevBuf bX, bLs, bE, bLa, bV, bR, bC, bP, bB;
  if ((rc = bX.alloc(sizeof(T) * (size_t)m * n)) || (rc = bLs.alloc(sizeof(int) * (nL ? nL : 1))) ||
      (B && (rc = bB.alloc((size_t)m * n))) ||
      (rc = bE.alloc(sizeof(T) * (nEps ? nEps : 1))) || (rc = bLa.alloc(sizeof(T) * (nLambda ? nLambda : 1))) ||
      (rc = bV.alloc(sizeof(T) * (size_t)Le * n)) || (rc = bR.alloc(sizeof(int) * (size_t)Le * n)) ||
      (rc = bC.alloc(sizeof(int) * (size_t)n)) ||
      (path && (rc = bP.alloc(sizeof(T) * (size_t)p * pcols)))) {
    spx_dict_destroy(d);
    return rc;
  }
  hipStream_t st = ctx->stream;
  bool ok = hipMemcpyAsync(bX.p, X, sizeof(T) * (size_t)m * n, hipMemcpyHostToDevice, st) == hipSuccess;
  if (nL) ok = ok && hipMemcpyAsync(bLs.p, L, sizeof(int) * nL, hipMemcpyHostToDevice, st) == hipSuccess;
  if (nEps) ok = ok && hipMemcpyAsync(bE.p, eps, sizeof(T) * nEps, hipMemcpyHostToDevice, st) == hipSuccess;
  if (nLambda) ok = ok && hipMemcpyAsync(bLa.p, lambda, sizeof(T) * nLambda, hipMemcpyHostToDevice, st) == hipSuccess;
  if (B) ok = ok && hipMemcpyAsync(bB.p, B, (size_t)m * n, hipMemcpyHostToDevice, st) == hipSuccess;
  if (!ok) {
    spx_dict_destroy(d);
    set_error("omp: upload failed");
    return SPX_ERR_HIP;
  }
  spx_lists_dev l{bV.p, static_cast<int*>(bR.p), static_cast<int*>(bC.p), Le, nullptr};
  if (path) {
    l.path = bP.p;
    l.path_cols = pcols;
  }
  rc = omp_dev<T>(ctx, d, static_cast<const T*>(bX.p), m, n,
                  nL ? static_cast<const int*>(bLs.p) : nullptr, nL,
                  nEps ? static_cast<const T*>(bE.p) : nullptr, nEps,
                  nLambda ? static_cast<const T*>(bLa.p) : nullptr, nLambda, &o, &l,
                  static_cast<const unsigned char*>(bB.p), m);
  if (!rc) rc = lists_to_host_csc<T>(ctx, l, p, n, out);
  if (!rc && path &&
      hipMemcpy(path, bP.p, sizeof(T) * (size_t)p * pcols, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("omp: path download failed");
    rc = SPX_ERR_HIP;
  }
  spx_dict_destroy(d);
  return rc;
}

template <typename T>
static int fista_host(spx_ctx* ctx, const T* X, int m, int n, const T* D, int p,
                      const T* W0, T* W, const spx_fista_opts* o, T* info) {
  if (!ctx || !o || !D || !W0 || !W || !info || m <= 0 || p <= 0 || n < 0 || (n > 0 && !X)) {
    set_error("fistaFlat : Argument sizes are not consistent");
    return SPX_ERR_ARG;
  }
  if (o->regul != 0 && o->regul != 1) {
    set_error("fistaFlat: Unknown regularization.");
    return SPX_ERR_ARG;
  }
  if (n == 0) return SPX_OK;
  DeviceGuard g(ctx->device);
  spx_dict* d = nullptr;
  int rc = dict_create<T>(ctx, D, m, p, 0, o->precision, &d);
  if (rc) return rc;
  DevBuf bX, bW0, bW, bI;
  if ((rc = bX.alloc(sizeof(T) * (size_t)m * n)) || (rc = bW0.alloc(sizeof(T) * (size_t)p * n)) ||
      (rc = bW.alloc(sizeof(T) * (size_t)p * n)) || (rc = bI.alloc(sizeof(T) * 4 * (size_t)n))) {
    spx_dict_destroy(d);
    return rc;
  }
  hipStream_t st = ctx->stream;
  if (hipMemcpyAsync(bX.p, X, sizeof(T) * (size_t)m * n, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(bW0.p, W0, sizeof(T) * (size_t)p * n, hipMemcpyHostToDevice, st) != hipSuccess) {
    spx_dict_destroy(d);
    set_error("fistaFlat: upload failed");
    return SPX_ERR_HIP;
  }
  rc = fista_run<T>(ctx, d, static_cast<const T*>(bX.p), n, static_cast<const T*>(bW0.p),
                    static_cast<T*>(bW.p), o, static_cast<T*>(bI.p));
  if (!rc) {
    if (hipMemcpyAsync(W, bW.p, sizeof(T) * (size_t)p * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(info, bI.p, sizeof(T) * 4 * (size_t)n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      set_error("fistaFlat: download failed");
      rc = SPX_ERR_HIP;
    }
  }
  spx_dict_destroy(d);
  return rc;
}

}  // namespace spx

// ------------------------------------------------------------------ ctx

int spx_ctx::scratch_get(int id, size_t bytes, void** out) {
  std::lock_guard<std::mutex> lk(mu);
  if (arena) {
    if ((size_t)id >= arena_slots.size()) arena_slots.resize(id + 1);
    Slot& a = arena_slots[id];
    if (!a.ptr || a.bytes < bytes) {
      const size_t want = (bytes + 255) & ~size_t(255);
      if (want > arena_bytes - arena_used) {
        spx::set_error("allocation-free call: scratch of %zu bytes exceeds the %zu bytes left "
                       "of the %zu-byte static scratch (use the allocating entry point, or "
                       "fewer signals per call)", want, arena_bytes - arena_used, arena_bytes);
        return SPX_ERR_NOMEM;
      }
      a.ptr = arena + arena_used;
      a.bytes = want;
      arena_used += want;
    }
    *out = a.ptr;
    return SPX_OK;
  }
  if ((size_t)id >= slots.size()) slots.resize(id + 1);
  Slot& s = slots[id];
  if (s.bytes < bytes) {
    if (s.ptr) {
      if (hipStreamSynchronize(stream) != hipSuccess || hipFree(s.ptr) != hipSuccess) {
        spx::set_error("scratch: release failed");
        return SPX_ERR_HIP;
      }
      s.ptr = nullptr;
      s.bytes = 0;
    }
    const size_t want = bytes + bytes / 8;
    if (hipMalloc(&s.ptr, want) != hipSuccess) {
      s.ptr = nullptr;
      spx::set_error("scratch: device allocation of %zu bytes failed", want);
      return SPX_ERR_NOMEM;
    }
    s.bytes = want;
  }
  *out = s.ptr;
  return SPX_OK;
}

hipEvent_t spx_ctx::event_get() {
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!free_events.empty()) {
      hipEvent_t e = free_events.back();
      free_events.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int spx_ctx::timers_fold() {
  std::vector<Pending> p;
  {
    std::lock_guard<std::mutex> lk(mu);
    p.swap(pending);
  }
  int rc = SPX_OK;
  for (auto& e : p) {
    float ms = 0.f;
    if (hipEventSynchronize(e.b) != hipSuccess || hipEventElapsedTime(&ms, e.a, e.b) != hipSuccess) {
      spx::set_error("timing: event query failed");
      rc = SPX_ERR_HIP;
    } else if (e.id >= 0 && e.id < SPX_TIMER_COUNT) {
      timer_ms[e.id] += ms;
      timer_n[e.id] += 1;
    }
    std::lock_guard<std::mutex> lk(mu);
    free_events.push_back(e.a);
    free_events.push_back(e.b);
  }
  return rc;
}

spx_ctx::~spx_ctx() {
  for (auto& g : graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (capture_stream) (void)hipStreamDestroy(capture_stream);
  for (auto& e : pending) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  for (auto e : free_events) (void)hipEventDestroy(e);
  for (auto& s : slots)
    if (s.ptr) (void)hipFree(s.ptr);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

extern "C" {

const char* spx_last_error(void) { return spx::g_err; }
const char* spx_version(void) { return "spx 0.1 (gfx950)"; }

// ---------------------------------------------- allocation-free entry point
// Static scratch of the allocation-free entry point: one copy per device,
// held by the loaded code object until the process ends.  Host code uses
// only the device address of each copy (taken once per device at context
// creation), never the symbol itself.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// the device address of this device's copy (the device must be current)
int scratch_address(int device, char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    spx::set_error("spx: device %d beyond the supported %d", device, kMaxDevices);
    return SPX_ERR_ARG;
  }
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      spx::set_error("spx: address of the static scratch on device %d: %s", device,
                     hipGetErrorString(e));
      return SPX_ERR_HIP;
    }
    g_scratch_ptr[device] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[device];
  return SPX_OK;
}
}  // namespace

int spx_lasso_dtx_dev_noalloc_f32(spx_ctx* ctx, const spx_dict* d, const float* DtX, int ldd,
                                  const float* norms, int n, const float* lambda, int n_lambda,
                                  float lambda2, const spx_lasso_opts* o, spx_lists_dev* out) {
  if (!ctx) {
    spx::set_error("lasso_dtx_dev_noalloc: no context");
    return SPX_ERR_ARG;
  }
  spx::DeviceGuard g(ctx->device);
  char* base = nullptr;
  int rc = scratch_address(ctx->device, &base);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->arena = base;
    ctx->arena_bytes = kNoAllocScratchBytes;
    ctx->arena_used = 0;
    for (auto& sl : ctx->arena_slots) sl = spx_ctx::Slot{};
  }
  rc = spx::lasso_dtx_dev<float>(ctx, d, DtX, ldd, norms, n, lambda, n_lambda, lambda2, o, out);
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->arena = nullptr;
  ctx->arena_used = 0;
  return rc;
}

int spx_ctx_create(spx_ctx** out, int device) {
  if (!out) return SPX_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    spx::set_error("spx_ctx_create: no HIP device %d", device);
    return SPX_ERR_HIP;
  }
  spx::DeviceGuard g(device);
  {
    // once per device, with the device current: the address of its copy of
    // the allocation-free entry point's static scratch
    char* sp = nullptr;
    const int rs = scratch_address(device, &sp);
    if (rs) return rs;
  }
  spx_ctx* c = new spx_ctx();
  c->device = device;
  if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    spx::set_error("spx_ctx_create: device query / stream creation failed");
    return SPX_ERR_HIP;
  }
  c->own_stream = true;
  // slot table of allocation-free calls sized once here, so that such a call
  // does not grow it
  c->arena_slots.resize(32);
  *out = c;
  return SPX_OK;
}

void spx_ctx_destroy(spx_ctx* ctx) {
  if (!ctx) return;
  spx::DeviceGuard g(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
}

int spx_ctx_set_stream(spx_ctx* ctx, void* s) {
  if (!ctx) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  ctx->stream = static_cast<hipStream_t>(s);
  ctx->own_stream = false;
  return SPX_OK;
}

int spx_ctx_own_stream(spx_ctx* ctx) {
  if (!ctx) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  if (ctx->own_stream) return SPX_OK;
  (void)hipStreamSynchronize(ctx->stream);
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    spx::set_error("spx_ctx_own_stream: stream creation failed");
    return SPX_ERR_HIP;
  }
  ctx->stream = st;
  ctx->own_stream = true;
  return SPX_OK;
}

int spx_ctx_synchronize(spx_ctx* ctx) {
  if (!ctx) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    spx::set_error("spx_ctx_synchronize: %s", hipGetErrorString(hipGetLastError()));
    return SPX_ERR_HIP;
  }
  return SPX_OK;
}

int spx_ctx_set_timing(spx_ctx* ctx, int on) {
  if (!ctx) return SPX_ERR_ARG;
  ctx->timing = on != 0;
  return SPX_OK;
}

int spx_ctx_read_timing(spx_ctx* ctx, double* ms, long long* launches, int count) {
  if (!ctx || count < 0) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  int rc = ctx->timers_fold();
  for (int i = 0; i < count; ++i) {
    if (ms) ms[i] = i < SPX_TIMER_COUNT ? ctx->timer_ms[i] : 0.0;
    if (launches) launches[i] = i < SPX_TIMER_COUNT ? ctx->timer_n[i] : 0;
  }
  return rc;
}

size_t spx_ctx_scratch_bytes(const spx_ctx* ctx) {
  if (!ctx) return 0;
  size_t b = 0;
  for (const auto& sl : ctx->slots) b += sl.bytes;
  return b;
}

int spx_ctx_release_scratch(spx_ctx* ctx) {
  if (!ctx) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (ctx->stream && hipStreamSynchronize(ctx->stream) != hipSuccess) {
    spx::set_error("release_scratch: stream synchronisation failed");
    return SPX_ERR_HIP;
  }
  // captured FISTA loops address the scratch: they go with it
  for (auto& gr : ctx->graphs)
    if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
  ctx->graphs.clear();
  for (auto& sl : ctx->slots) {
    if (sl.ptr && hipFree(sl.ptr) != hipSuccess) {
      spx::set_error("release_scratch: device free failed");
      return SPX_ERR_HIP;
    }
    sl.ptr = nullptr;
    sl.bytes = 0;
  }
  return SPX_OK;
}

int spx_ctx_reset_timing(spx_ctx* ctx) {
  if (!ctx) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  int rc = ctx->timers_fold();
  for (int i = 0; i < SPX_TIMER_COUNT; ++i) {
    ctx->timer_ms[i] = 0.0;
    ctx->timer_n[i] = 0;
  }
  return rc;
}

int spx_dict_create_f32(spx_ctx* ctx, const float* D, int m, int p, int on_device,
                        int precision, spx_dict** out) {
  return spx::dict_create<float>(ctx, D, m, p, on_device, precision, out);
}
int spx_dict_create_f64(spx_ctx* ctx, const double* D, int m, int p, int on_device,
                        spx_dict** out) {
  return spx::dict_create<double>(ctx, D, m, p, on_device, SPX_PREC_FP32, out);
}

void spx_dict_destroy(spx_dict* d) {
  if (!d) return;
  // the creating context may already be gone: only the device is used, and
  // every stream using the dictionary is drained before the buffers go
  spx::DeviceGuard g(d->device);
  (void)hipDeviceSynchronize();
  if (d->D) (void)hipFree(d->D);
  if (d->Dt) (void)hipFree(d->Dt);
  if (d->G) (void)hipFree(d->G);
  if (d->Gr) (void)hipFree(d->Gr);
  delete d;
}

int spx_dict_dims(const spx_dict* d, int* m, int* p) {
  if (!d) return SPX_ERR_ARG;
  if (m) *m = d->m;
  if (p) *p = d->p;
  return SPX_OK;
}

size_t spx_dict_bytes(const spx_dict* d) {
  if (!d) return 0;
  return (size_t)d->elem * ((size_t)2 * d->m * d->p +
                            (size_t)d->ldg * d->p * (d->Gr ? 2 : 1));
}

int spx_dict_autotune_lasso_f32(spx_ctx* ctx, spx_dict* d, const float* X, int ldx, int n,
                                const float* lambda, int nlam, float lambda2,
                                const spx_lasso_opts* o, int* waves, float* ms) {
  return spx::lasso_autotune<float>(ctx, d, X, ldx, n, lambda, nlam, lambda2, o, waves, ms);
}
int spx_dict_autotune_lasso_f64(spx_ctx* ctx, spx_dict* d, const double* X, int ldx, int n,
                                const double* lambda, int nlam, double lambda2,
                                const spx_lasso_opts* o, int* waves, float* ms) {
  return spx::lasso_autotune<double>(ctx, d, X, ldx, n, lambda, nlam, lambda2, o, waves, ms);
}
int spx_dict_autotune_omp_f32(spx_ctx* ctx, spx_dict* d, const float* X, int ldx, int n,
                              const int* L, int nL, const float* eps, int nEps,
                              const float* lambda, int nLambda, int* waves, float* ms) {
  return spx::omp_autotune<float>(ctx, d, X, ldx, n, L, nL, eps, nEps, lambda, nLambda, waves,
                                  ms);
}
int spx_dict_autotune_omp_f64(spx_ctx* ctx, spx_dict* d, const double* X, int ldx, int n,
                              const int* L, int nL, const double* eps, int nEps,
                              const double* lambda, int nLambda, int* waves, float* ms) {
  return spx::omp_autotune<double>(ctx, d, X, ldx, n, L, nL, eps, nEps, lambda, nLambda, waves,
                                   ms);
}

int spx_dict_get_waves(const spx_dict* d, int solver, int* waves) {
  if (!d || solver < 0 || solver > 1 || !waves) return SPX_ERR_ARG;
  *waves = spx::tuned_waves(d, solver);
  return SPX_OK;
}

int spx_dict_set_waves(spx_dict* d, int solver, int waves) {
  if (!d || solver < 0 || solver > 1 || !(waves == 0 || waves == 1 || waves == 2 ||
                                          waves == 4 || waves == 16)) {
    spx::set_error("spx_dict_set_waves: solver must be 0 or 1 and waves one of 0, 1, 2, 4, 16");
    return SPX_ERR_ARG;
  }
  d->waves[solver] = waves;
  return SPX_OK;
}

int spx_lasso_L(int m, int p, int L, int mode, double lambda1, int pos, double lambda2) {
  // cpp_library/spams.h:120-131
  if (L < 0) L = p;
  if (L > m && !(mode == SPX_PENALTY && lambda1 == 0 && !pos && lambda2 > 0)) L = m;
  if (L > p) L = p;
  return L;
}

int spx_omp_L(int m, int p, const int* L, int nL, int n) {
  // cpp_library/spams.h:269-271, decomp/decomp.h:326-335
  int Lv;
  if (!L || nL == 0) Lv = m < p ? m : p;
  else if (nL == 1) Lv = L[0];
  else {
    Lv = L[0];
    for (int i = 1; i < n && i < nL; ++i) Lv = L[i] > Lv ? L[i] : Lv;
  }
  if (Lv <= 0) return 0;
  if (Lv > p) Lv = p;
  if (Lv > m) Lv = m;
  return Lv;
}

int spx_lasso_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D, int p,
                  const float* lambda, int n_lambda, float lambda2,
                  const spx_lasso_opts* opts, spx_csc* out) {
  return spx::lasso_host<float>(ctx, X, m, n, D, p, lambda, n_lambda, lambda2, opts, out);
}
int spx_lasso_f64(spx_ctx* ctx, const double* X, int m, int n, const double* D, int p,
                  const double* lambda, int n_lambda, double lambda2,
                  const spx_lasso_opts* opts, spx_csc* out) {
  return spx::lasso_host<double>(ctx, X, m, n, D, p, lambda, n_lambda, lambda2, opts, out);
}
int spx_lasso_qq_f32(spx_ctx* ctx, const float* X, int m, int n, const float* Q,
                     const float* q, int p, const float* lambda, int n_lambda, float lambda2,
                     const spx_lasso_opts* opts, spx_csc* out, float* path) {
  return spx::lasso_qq_host<float>(ctx, X, m, n, Q, q, p, lambda, n_lambda, lambda2, opts, out,
                                   path);
}
int spx_lasso_qq_f64(spx_ctx* ctx, const double* X, int m, int n, const double* Q,
                     const double* q, int p, const double* lambda, int n_lambda,
                     double lambda2, const spx_lasso_opts* opts, spx_csc* out, double* path) {
  return spx::lasso_qq_host<double>(ctx, X, m, n, Q, q, p, lambda, n_lambda, lambda2, opts, out,
                                    path);
}
int spx_dict_create_gram_f32(spx_ctx* ctx, const float* Q, int m, int p, int on_device,
                             spx_dict** out) {
  return spx::dict_create_gram<float>(ctx, Q, m, p, on_device, out);
}
int spx_dict_create_gram_f64(spx_ctx* ctx, const double* Q, int m, int p, int on_device,
                             spx_dict** out) {
  return spx::dict_create_gram<double>(ctx, Q, m, p, on_device, out);
}
int spx_lasso_mask_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D, int p,
                       const unsigned char* B, const float* lambda, int n_lambda, float lambda2,
                       const spx_lasso_opts* opts, spx_csc* out) {
  if (!B) {
    spx::set_error("lassoMask: B is required");
    return SPX_ERR_ARG;
  }
  return spx::lasso_host<float>(ctx, X, m, n, D, p, lambda, n_lambda, lambda2, opts, out,
                                nullptr, nullptr, B);
}
int spx_lasso_mask_f64(spx_ctx* ctx, const double* X, int m, int n, const double* D, int p,
                       const unsigned char* B, const double* lambda, int n_lambda, double lambda2,
                       const spx_lasso_opts* opts, spx_csc* out) {
  if (!B) {
    spx::set_error("lassoMask: B is required");
    return SPX_ERR_ARG;
  }
  return spx::lasso_host<double>(ctx, X, m, n, D, p, lambda, n_lambda, lambda2, opts, out,
                                 nullptr, nullptr, B);
}
int spx_lasso_mask_dev_f32(spx_ctx* ctx, const spx_dict* d, const float* X, int ldx, int n,
                           const unsigned char* B, int ldb, const float* lambda, int n_lambda,
                           float lambda2, const spx_lasso_opts* o, spx_lists_dev* out) {
  return spx::lasso_mask_dev<float>(ctx, d, X, ldx, n, B, ldb, lambda, n_lambda, lambda2, o, out);
}
int spx_lasso_mask_dev_f64(spx_ctx* ctx, const spx_dict* d, const double* X, int ldx, int n,
                           const unsigned char* B, int ldb, const double* lambda, int n_lambda,
                           double lambda2, const spx_lasso_opts* o, spx_lists_dev* out) {
  return spx::lasso_mask_dev<double>(ctx, d, X, ldx, n, B, ldb, lambda, n_lambda, lambda2, o,
                                     out);
}
int spx_lasso_weighted_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D, int p,
                           const float* W, const float* lambda, int n_lambda,
                           const spx_lasso_opts* opts, spx_csc* out) {
  if (!W) {
    spx::set_error("lassoWeighted: W is required");
    return SPX_ERR_ARG;
  }
  return spx::lasso_host<float>(ctx, X, m, n, D, p, lambda, n_lambda, 0.f, opts, out, nullptr, W);
}
int spx_lasso_weighted_f64(spx_ctx* ctx, const double* X, int m, int n, const double* D, int p,
                           const double* W, const double* lambda, int n_lambda,
                           const spx_lasso_opts* opts, spx_csc* out) {
  if (!W) {
    spx::set_error("lassoWeighted: W is required");
    return SPX_ERR_ARG;
  }
  return spx::lasso_host<double>(ctx, X, m, n, D, p, lambda, n_lambda, 0.0, opts, out, nullptr, W);
}
int spx_lasso_path_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D, int p,
                       const float* lambda, int n_lambda, float lambda2,
                       const spx_lasso_opts* opts, spx_csc* out, float* path) {
  return spx::lasso_host<float>(ctx, X, m, n, D, p, lambda, n_lambda, lambda2, opts, out, path);
}
int spx_lasso_path_f64(spx_ctx* ctx, const double* X, int m, int n, const double* D, int p,
                       const double* lambda, int n_lambda, double lambda2,
                       const spx_lasso_opts* opts, spx_csc* out, double* path) {
  return spx::lasso_host<double>(ctx, X, m, n, D, p, lambda, n_lambda, lambda2, opts, out,
                                 path);
}

int spx_omp_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D, int p,
                const int* L, int nL, const float* eps, int n_eps, const float* lambda,
                int n_lambda, const spx_omp_opts* opts, spx_csc* out) {
  return spx::omp_host<float>(ctx, X, m, n, D, p, L, nL, eps, n_eps, lambda, n_lambda, opts, out);
}
int spx_omp_f64(spx_ctx* ctx, const double* X, int m, int n, const double* D, int p,
                const int* L, int nL, const double* eps, int n_eps, const double* lambda,
                int n_lambda, const spx_omp_opts* opts, spx_csc* out) {
  return spx::omp_host<double>(ctx, X, m, n, D, p, L, nL, eps, n_eps, lambda, n_lambda, opts, out);
}
int spx_omp_path_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D, int p,
                     const int* L, int nL, const float* eps, int n_eps, const float* lambda,
                     int n_lambda, const spx_omp_opts* opts, spx_csc* out, float* path) {
  return spx::omp_host<float>(ctx, X, m, n, D, p, L, nL, eps, n_eps, lambda, n_lambda, opts, out,
                              path);
}
int spx_omp_path_f64(spx_ctx* ctx, const double* X, int m, int n, const double* D, int p,
                     const int* L, int nL, const double* eps, int n_eps, const double* lambda,
                     int n_lambda, const spx_omp_opts* opts, spx_csc* out, double* path) {
  return spx::omp_host<double>(ctx, X, m, n, D, p, L, nL, eps, n_eps, lambda, n_lambda, opts,
                               out, path);
}
int spx_omp_mask_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D, int p,
                     const unsigned char* B, const int* L, int nL, const float* eps, int n_eps,
                     const float* lambda, int n_lambda, const spx_omp_opts* opts, spx_csc* out,
                     float* path) {
  if (!B) {
    spx::set_error("ompMask: B is required");
    return SPX_ERR_ARG;
  }
  return spx::omp_host<float>(ctx, X, m, n, D, p, L, nL, eps, n_eps, lambda, n_lambda, opts, out,
                              path, B);
}
int spx_omp_mask_f64(spx_ctx* ctx, const double* X, int m, int n, const double* D, int p,
                     const unsigned char* B, const int* L, int nL, const double* eps, int n_eps,
                     const double* lambda, int n_lambda, const spx_omp_opts* opts, spx_csc* out,
                     double* path) {
  if (!B) {
    spx::set_error("ompMask: B is required");
    return SPX_ERR_ARG;
  }
  return spx::omp_host<double>(ctx, X, m, n, D, p, L, nL, eps, n_eps, lambda, n_lambda, opts,
                               out, path, B);
}
int spx_omp_mask_dev_f32(spx_ctx* ctx, const spx_dict* d, const float* X, int ldx, int n,
                         const unsigned char* B, int ldb, const int* L, int nL, const float* eps,
                         int n_eps, const float* lambda, int n_lambda, const spx_omp_opts* o,
                         spx_lists_dev* out) {
  if (!B) {
    spx::set_error("ompMask: B is required");
    return SPX_ERR_ARG;
  }
  return spx::omp_dev<float>(ctx, d, X, ldx, n, L, nL, eps, n_eps, lambda, n_lambda, o, out, B,
                             ldb);
}
int spx_omp_mask_dev_f64(spx_ctx* ctx, const spx_dict* d, const double* X, int ldx, int n,
                         const unsigned char* B, int ldb, const int* L, int nL,
                         const double* eps, int n_eps, const double* lambda, int n_lambda,
                         const spx_omp_opts* o, spx_lists_dev* out) {
  if (!B) {
    spx::set_error("ompMask: B is required");
    return SPX_ERR_ARG;
  }
  return spx::omp_dev<double>(ctx, d, X, ldx, n, L, nL, eps, n_eps, lambda, n_lambda, o, out, B,
                              ldb);
}

int spx_fista_flat_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D, int p,
                       const float* W0, float* W, const spx_fista_opts* opts,
                       float* optim_info) {
  return spx::fista_host<float>(ctx, X, m, n, D, p, W0, W, opts, optim_info);
}
int spx_fista_flat_f64(spx_ctx* ctx, const double* X, int m, int n, const double* D, int p,
                       const double* W0, double* W, const spx_fista_opts* opts,
                       double* optim_info) {
  return spx::fista_host<double>(ctx, X, m, n, D, p, W0, W, opts, optim_info);
}

void spx_csc_free(spx_csc* c) {
  if (!c) return;
  free(c->pB);
  free(c->r);
  free(c->v);
  c->pB = nullptr;
  c->r = nullptr;
  c->v = nullptr;
}

int spx_lasso_dev_f32(spx_ctx* ctx, const spx_dict* d, const float* X, int ldx, int n,
                      const float* lambda, int n_lambda, float lambda2,
                      const spx_lasso_opts* o, spx_lists_dev* out) {
  return spx::lasso_dev<float>(ctx, d, X, ldx, n, lambda, n_lambda, lambda2, o, out);
}
int spx_lasso_dev_f64(spx_ctx* ctx, const spx_dict* d, const double* X, int ldx, int n,
                      const double* lambda, int n_lambda, double lambda2,
                      const spx_lasso_opts* o, spx_lists_dev* out) {
  return spx::lasso_dev<double>(ctx, d, X, ldx, n, lambda, n_lambda, lambda2, o, out);
}
int spx_lasso_weighted_dev_f32(spx_ctx* ctx, const spx_dict* d, const float* X, int ldx, int n,
                               const float* W, int ldw, const float* lambda, int n_lambda,
                               const spx_lasso_opts* o, spx_lists_dev* out) {
  if (!W) {
    spx::set_error("lassoWeighted: W is required");
    return SPX_ERR_ARG;
  }
  return spx::lasso_dev<float>(ctx, d, X, ldx, n, lambda, n_lambda, 0.f, o, out, W, ldw);
}
int spx_lasso_weighted_dev_f64(spx_ctx* ctx, const spx_dict* d, const double* X, int ldx, int n,
                               const double* W, int ldw, const double* lambda, int n_lambda,
                               const spx_lasso_opts* o, spx_lists_dev* out) {
  if (!W) {
    spx::set_error("lassoWeighted: W is required");
    return SPX_ERR_ARG;
  }
  return spx::lasso_dev<double>(ctx, d, X, ldx, n, lambda, n_lambda, 0.0, o, out, W, ldw);
}

int spx_lasso_dtx_dev_f32(spx_ctx* ctx, const spx_dict* d, const float* DtX, int ldd,
                          const float* norms, int n, const float* lambda, int n_lambda,
                          float lambda2, const spx_lasso_opts* o, spx_lists_dev* out) {
  return spx::lasso_dtx_dev<float>(ctx, d, DtX, ldd, norms, n, lambda, n_lambda, lambda2, o,
                                   out);
}
int spx_lasso_dtx_dev_f64(spx_ctx* ctx, const spx_dict* d, const double* DtX, int ldd,
                          const double* norms, int n, const double* lambda, int n_lambda,
                          double lambda2, const spx_lasso_opts* o, spx_lists_dev* out) {
  return spx::lasso_dtx_dev<double>(ctx, d, DtX, ldd, norms, n, lambda, n_lambda, lambda2, o,
                                    out);
}

int spx_sqnorms_dev_f32(spx_ctx* ctx, const float* X, int m, int ldx, int n, float* norms) {
  if (!ctx || ldx < m || n < 0) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  return spx::col_sqnorms<float>(ctx->stream, m, n, X, ldx, norms);
}
int spx_sqnorms_dev_f64(spx_ctx* ctx, const double* X, int m, int ldx, int n, double* norms) {
  if (!ctx || ldx < m || n < 0) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  return spx::col_sqnorms<double>(ctx->stream, m, n, X, ldx, norms);
}

int spx_omp_dev_f32(spx_ctx* ctx, const spx_dict* d, const float* X, int ldx, int n,
                    const int* L, int nL, const float* eps, int n_eps, const float* lambda,
                    int n_lambda, const spx_omp_opts* o, spx_lists_dev* out) {
  return spx::omp_dev<float>(ctx, d, X, ldx, n, L, nL, eps, n_eps, lambda, n_lambda, o, out);
}
int spx_omp_dev_f64(spx_ctx* ctx, const spx_dict* d, const double* X, int ldx, int n,
                    const int* L, int nL, const double* eps, int n_eps, const double* lambda,
                    int n_lambda, const spx_omp_opts* o, spx_lists_dev* out) {
  return spx::omp_dev<double>(ctx, d, X, ldx, n, L, nL, eps, n_eps, lambda, n_lambda, o, out);
}

int spx_fista_flat_dev_f32(spx_ctx* ctx, const spx_dict* d, const float* X, int n,
                           const float* W0, float* W, const spx_fista_opts* o, float* info) {
  if (!ctx || !d || !o || d->elem != 4) {
    spx::set_error("fista_dev: invalid arguments");
    return SPX_ERR_ARG;
  }
  if (int rc = spx::need_dictionary(d)) return rc;
  spx::DeviceGuard g(ctx->device);
  return spx::fista_run<float>(ctx, d, X, n, W0, W, o, info);
}
int spx_fista_flat_dev_f64(spx_ctx* ctx, const spx_dict* d, const double* X, int n,
                           const double* W0, double* W, const spx_fista_opts* o, double* info) {
  if (!ctx || !d || !o || d->elem != 8) {
    spx::set_error("fista_dev: invalid arguments");
    return SPX_ERR_ARG;
  }
  if (int rc = spx::need_dictionary(d)) return rc;
  spx::DeviceGuard g(ctx->device);
  return spx::fista_run<double>(ctx, d, X, n, W0, W, o, info);
}

int spx_lists_to_csc_dev_f32(spx_ctx* ctx, const spx_lists_dev* l, int n, int* pB, int* r,
                             float* v, long long* nnz) {
  if (!ctx || !l || !pB || !nnz) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  int rc = spx::csc_count_scan<float>(ctx, l->count, n, pB, nnz);
  if (rc) return rc;
  return spx::csc_fill<float>(ctx, static_cast<const float*>(l->v), l->r, l->L, n, pB, r, v);
}
int spx_lists_to_csc_dev_f64(spx_ctx* ctx, const spx_lists_dev* l, int n, int* pB, int* r,
                             double* v, long long* nnz) {
  if (!ctx || !l || !pB || !nnz) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  int rc = spx::csc_count_scan<double>(ctx, l->count, n, pB, nnz);
  if (rc) return rc;
  return spx::csc_fill<double>(ctx, static_cast<const double*>(l->v), l->r, l->L, n, pB, r, v);
}

int spx_dtx_dev_f32(spx_ctx* ctx, const spx_dict* d, const float* X, int ldx, int n,
                    float* DtX, int ldc, int precision) {
  if (!ctx || !d || d->elem != 4 || ldc < d->p || ldx < d->m) return SPX_ERR_ARG;
  if (int rc = spx::need_dictionary(d)) return rc;
  spx::DeviceGuard g(ctx->device);
  return spx::gemm_tn<float>(ctx->stream, d->p, n, d->m, static_cast<const float*>(d->D),
                             d->m, X, ldx, DtX, ldc, 1.f, 0.f, nullptr, 0, precision);
}
int spx_dtx_dev_f64(spx_ctx* ctx, const spx_dict* d, const double* X, int ldx, int n,
                    double* DtX, int ldc) {
  if (!ctx || !d || d->elem != 8 || ldc < d->p || ldx < d->m) return SPX_ERR_ARG;
  if (int rc = spx::need_dictionary(d)) return rc;
  spx::DeviceGuard g(ctx->device);
  return spx::gemm_tn<double>(ctx->stream, d->p, n, d->m, static_cast<const double*>(d->D),
                              d->m, X, ldx, DtX, ldc, 1.0, 0.0, nullptr, 0, 0);
}

int spx_dict_lipschitz(spx_ctx* ctx, const spx_dict* d, int iters, double* L) {
  if (!ctx || !d || !L || iters < 1) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  return d->elem == 4 ? spx::lipschitz<float>(ctx, d, iters, L)
                      : spx::lipschitz<double>(ctx, d, iters, L);
}

int spx_dict_gram_f32(spx_ctx* ctx, const spx_dict* d, float* G) {
  if (!ctx || !d || d->elem != 4 || !G) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  if (hipMemcpy2DAsync(G, sizeof(float) * d->p, d->G, sizeof(float) * d->ldg,
                       sizeof(float) * d->p, d->p, hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess) {
    spx::set_error("dict_gram: copy failed");
    return SPX_ERR_HIP;
  }
  return SPX_OK;
}
int spx_dict_gram_f64(spx_ctx* ctx, const spx_dict* d, double* G) {
  if (!ctx || !d || d->elem != 8 || !G) return SPX_ERR_ARG;
  spx::DeviceGuard g(ctx->device);
  if (hipMemcpy2DAsync(G, sizeof(double) * d->p, d->G, sizeof(double) * d->ldg,
                       sizeof(double) * d->p, d->p, hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess) {
    spx::set_error("dict_gram: copy failed");
    return SPX_ERR_HIP;
  }
  return SPX_OK;
}

}  // extern "C"
