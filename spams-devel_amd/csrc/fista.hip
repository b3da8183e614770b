// FISTA-flat for the square loss with the l1 / elastic-net regularisers,
// all signals at once (FISTA_Generic, prox/fista.h:2905-3007; ISTA_Generic
// :2785-2903; SqLoss :361-494; Lasso/Ridge/ComposeProx :1167-1472;
// duality_gap :2653-2687).
//
// Per iteration:
//   R    = D*Y - X        MFMA GEMM (A = D', E = X)
//   Grad = D'*R           MFMA GEMM
//   fused pass, one wavefront per signal: z = prox(y - grad/L, lambda/L),
//   x <- z, y <- z + ((1-t_old)/t)(x_old - z), and the per-signal sums of the
//   relative-change stop test.
// The step sequence t_k does not depend on the signal, so it is a host scalar;
// the Lipschitz constant L is per signal (backtracking multiplies it by
// gamma).  A signal that meets its stop test is frozen (x, y and its
// optim_info column stop changing) while the others continue.
#include <cmath>
#include <vector>

#include <algorithm>
#include <cstring>

#include "spx_common.h"
#include "spx_internal.h"

namespace spx {

namespace {

enum { REG_L1 = 0, REG_EN = 1 };

template <typename T>
struct ProxParams {
  T lambda;     // regularisation weight
  T lambda2d1;  // lambda2 / lambda
  int regul;
  int pos;
};

// prox of the regulariser at threshold thr (Lasso::prox then Ridge::prox for
// the elastic net, prox/fista.h:1173-1178, :1325-1330, :1431-1450)
template <typename T>
__device__ __forceinline__ T prox1(T v, T thr, T scale, int pos, int regul) {
  if (pos && v < 0) v = 0;
  if (v > thr) v -= thr;
  else if (v < -thr) v += thr;
  else v = 0;
  if (regul == REG_EN) {
    if (pos && v < 0) v = 0;
    v *= scale;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T en_scale(T thr, T lambda2d1) {
  return static_cast<T>(1.0 / (1.0 + static_cast<double>(thr * lambda2d1)));
}

// Fused proximal + momentum step.  One wavefront per column j.
// stopsum[2j] = ||x_old - x_new||^2, stopsum[2j+1] = ||x_new||^2.
template <typename T>
__global__ __launch_bounds__(256) void prox_step_kernel(
    int p, int n, T* __restrict__ x, T* __restrict__ y,
    const T* __restrict__ grad, const T* __restrict__ Lsig,
    const int* __restrict__ active, ProxParams<T> pp, T beta, int ista,
    T* __restrict__ stopsum) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n || !active[j]) return;
  const T L = Lsig[j];
  const T ainv = -T(1) / L;
  const T thr = pp.lambda / L;
  const T sc = en_scale(thr, pp.lambda2d1);
  T* xc = x + (size_t)j * p;
  T* yc = y + (size_t)j * p;
  const T* gc = grad + (size_t)j * p;
  T d2 = 0, x2 = 0;
  for (int k = lane; k < p; k += 64) {
    const T base = ista ? xc[k] : yc[k];
    const T z = prox1(base + ainv * gc[k], thr, sc, pp.pos, pp.regul);
    const T dlt = xc[k] - z;
    xc[k] = z;
    if (!ista) yc[k] = z + beta * dlt;
    d2 += dlt * dlt;
    x2 += z * z;
  }
  d2 = wave_sum(d2);
  x2 = wave_sum(x2);
  if (lane == 0) {
    stopsum[2 * j] = d2;
    stopsum[2 * j + 1] = x2;
  }
}

// Backtracking trial: z = prox(base - grad/L) into zbuf, diff = base - z,
// dn[j] = ||diff||^2 (SqLoss::test_backtracking, prox/fista.h:403-416).
template <typename T>
__global__ __launch_bounds__(256) void prox_trial_kernel(
    int p, int n, const T* __restrict__ base, const T* __restrict__ grad,
    const T* __restrict__ Lsig, const int* __restrict__ pending,
    ProxParams<T> pp, T* __restrict__ zbuf, T* __restrict__ diff,
    T* __restrict__ dn) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n || !pending[j]) return;
  const T L = Lsig[j];
  const T ainv = -T(1) / L;
  const T thr = pp.lambda / L;
  const T sc = en_scale(thr, pp.lambda2d1);
  const T* bc = base + (size_t)j * p;
  const T* gc = grad + (size_t)j * p;
  T* zc = zbuf + (size_t)j * p;
  T* dc = diff + (size_t)j * p;
  T s = 0;
  for (int k = lane; k < p; k += 64) {
    const T z = prox1(bc[k] + ainv * gc[k], thr, sc, pp.pos, pp.regul);
    zc[k] = z;
    const T d = bc[k] - z;
    dc[k] = d;
    s += d * d;
  }
  s = wave_sum(s);
  if (lane == 0) dn[j] = s;
}

// Accept/reject a backtracking trial: passes when ||D diff||^2 <= L ||diff||^2.
template <typename T>
__global__ void backtrack_check_kernel(int n, int m, const T* __restrict__ Ud,
                                       const T* __restrict__ Vd,
                                       const T* __restrict__ dn, T gamma,
                                       int last, T* __restrict__ Lsig,
                                       int* __restrict__ pending,
                                       int* __restrict__ npending) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n || !pending[j]) return;
  // ||D (y - z)||^2, or with the Gram matrix (y - z)' G (y - z) = V'U
  // with U = G (y - z), V = y - z (prox/fista.h:404-413)
  const T* u = Ud + (size_t)j * m;
  const T* v = Vd ? Vd + (size_t)j * m : u;
  T s = 0;
  for (int k = lane; k < m; k += 64) s += u[k] * v[k];
  s = wave_sum(s);
  if (lane == 0) {
    // the reference's while loop multiplies L on every failed trial, the
    // last one included, and keeps the last trial's prox
    const bool ok = s <= Lsig[j] * dn[j];
    if (ok) {
      pending[j] = 0;
    } else {
      Lsig[j] *= gamma;
      if (last) pending[j] = 0;
      else atomicAdd(npending, 1);
    }
  }
}

// Momentum step from an accepted z (backtracking path).
template <typename T>
__global__ __launch_bounds__(256) void momentum_kernel(
    int p, int n, T* __restrict__ x, T* __restrict__ y,
    const T* __restrict__ zbuf, const int* __restrict__ active, T beta,
    int ista, T* __restrict__ stopsum) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n || !active[j]) return;
  T* xc = x + (size_t)j * p;
  T* yc = y + (size_t)j * p;
  const T* zc = zbuf + (size_t)j * p;
  T d2 = 0, x2 = 0;
  for (int k = lane; k < p; k += 64) {
    const T z = zc[k];
    const T dlt = xc[k] - z;
    xc[k] = z;
    if (!ista) yc[k] = z + beta * dlt;
    d2 += dlt * dlt;
    x2 += z * z;
  }
  d2 = wave_sum(d2);
  x2 = wave_sum(x2);
  if (lane == 0) {
    stopsum[2 * j] = d2;
    stopsum[2 * j + 1] = x2;
  }
}

// Relative-change stop test for the non-dual regulariser
// (prox/fista.h:2995-2996): sqrt(||dx||^2 / max(EPSILON, ||x||^2)) < tol.
template <typename T>
__global__ void relstop_kernel(int n, const T* __restrict__ stopsum, double tol,
                               int it, int* __restrict__ active,
                               T* __restrict__ info) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n || !active[j]) return;
  const double d2 = stopsum[2 * j], x2 = stopsum[2 * j + 1];
  const double den = x2 > 10e-10 ? x2 : 10e-10;
  if (sqrt(d2 / den) < tol) {
    active[j] = 0;
    info[4 * (size_t)j + 3] = static_cast<T>(it);
  }
}

// Per-signal evaluation: objective and (l1) Fenchel duality gap.
// g1 = D x - x_sig (m), g2 = D' g1 (p).  mode 0: it0 check (updates best
// dual, gap, may stop the signal); mode 1: final evaluation (writes loss).
template <typename T>
__global__ __launch_bounds__(256) void eval_kernel(
    int m, int p, int n, const T* __restrict__ X, const T* __restrict__ x,
    const T* __restrict__ g1, const T* __restrict__ g2, ProxParams<T> pp,
    int duality, int mode, double tol, int it, int* __restrict__ active,
    T* __restrict__ best_dual, T* __restrict__ info) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n) return;
  if (mode == 0 && !active[j]) return;
  const T* g1c = g1 + (size_t)j * m;
  const T* Xc = X + (size_t)j * m;
  const T* xc = x + (size_t)j * p;
  const T* g2c = g2 + (size_t)j * p;
  T r2 = 0, rx = 0, a1 = 0, x2 = 0, mm = 0;
  for (int k = lane; k < m; k += 64) {
    r2 += g1c[k] * g1c[k];
    rx += g1c[k] * Xc[k];
  }
  const T sc = -T(1) / pp.lambda;
  for (int k = lane; k < p; k += 64) {
    a1 += fabs(xc[k]);
    x2 += xc[k] * xc[k];
    if (duality) {
      T v = g2c[k] * sc;
      if (pp.pos && v < 0) v = 0;
      mm = fabs(v) > mm ? fabs(v) : mm;
    }
  }
  r2 = wave_sum(r2);
  rx = wave_sum(rx);
  a1 = wave_sum(a1);
  x2 = wave_sum(x2);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T t = __shfl_xor(mm, o, 64);
    mm = t > mm ? t : mm;
  }
  if (lane != 0) return;
  const T loss = static_cast<T>(0.5 * r2);
  T reg = a1;
  if (pp.regul == REG_EN) reg = a1 + pp.lambda2d1 * static_cast<T>(0.5 * x2);
  T* inf = info + 4 * (size_t)j;
  if (mode == 1) inf[0] = loss + pp.lambda * reg;
  if (!duality) return;
  const T primal = loss + pp.lambda * a1;
  const T scal = mm > T(1) ? T(1) / mm : T(1);
  T dual = -(static_cast<T>(0.5 * (scal * scal) * r2 + scal * rx));
  const T bd = best_dual[j];
  dual = dual > bd ? dual : bd;
  const T delta = primal == 0 ? T(0) : (primal - dual) / fabs(primal);
  best_dual[j] = dual;
  inf[1] = dual;
  inf[2] = delta;
  if (mode == 0 && static_cast<double>(delta) < tol) {
    active[j] = 0;
    inf[3] = static_cast<T>(it);
  }
}

template <typename T>
__global__ void init_state_kernel(int n, T L0, T* Lsig, int* active, T* best,
                                  T* info) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  Lsig[j] = L0;
  active[j] = 1;
  best[j] = -__builtin_huge_val();
  for (int k = 0; k < 4; ++k) info[4 * (size_t)j + k] = T(-1);
}

template <typename T>
__global__ void finish_it_kernel(int n, const int* active, int it, T* info) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n || !active[j]) return;
  info[4 * (size_t)j + 3] = static_cast<T>(it);
}

__global__ void set_pending_kernel(int n, const int* active, int* pending) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < n) pending[j] = active[j];
}

__global__ void zero_one_kernel(int* v) { *v = 0; }

inline dim3 cols4(int n) { return dim3((n + 3) / 4); }
inline dim3 lin(int n) { return dim3((n + 255) / 256); }

}  // namespace

template <typename T>
int fista_run(Ctx* ctx, const Dict* d, const T* X, int n, const T* W0, T* W,
              const spx_fista_opts* o, T* info) {
  if (n <= 0) return SPX_OK;
  if (!(o->lambda1 > 0)) {
    set_error("fistaFlat: lambda1 must be positive");
    return SPX_ERR_ARG;
  }
  const int m = d->m, p = d->p;
  hipStream_t st = ctx->stream;
  const T* D = static_cast<const T*>(d->D);
  const T* Dt = static_cast<const T*>(d->Dt);
  const int prec = o->precision;

  // workspace: y, grad (p x n); R (m x n); z, diff (p x n, backtracking only);
  // per-signal L, active, pending, best dual, stop sums; the slices of the
  // two products when a small batch leaves most of the chip idle (the inner
  // dimension is then split, gemm_tn_product)
  const bool bt = !o->fixed_step;
  // compute_gram: the gradient as G alpha - D'x with G = D'D and D'X formed
  // once, and the backtracking test as (y-z)' G (y-z) (prox/fista.h:364-413)
  const bool gram = o->compute_gram != 0;
  const size_t pn = (size_t)p * n, mn = (size_t)m * n;
  const int sR = gemm_small_split(m, n, p), sG = gemm_small_split(p, n, m);
  const int sQ = gram ? gemm_small_split(p, n, p) : 1;
  const size_t npart = std::max({sR > 1 ? (size_t)sR * mn : 0, sG > 1 ? (size_t)sG * pn : 0,
                                 sQ > 1 ? (size_t)sQ * pn : 0});
  // every piece starts 256-byte aligned (take below)
  auto padded = [](size_t bytes) { return (bytes + 255) & ~size_t(255); };
  const size_t need = padded(sizeof(T) * pn) * ((bt ? 4 : 2) + (gram ? (bt ? 2 : 1) : 0)) +
                      padded(sizeof(T) * mn) +
                      padded(sizeof(T) * n) * 2 + padded(sizeof(T) * 2 * n) +
                      padded(sizeof(int) * n) * 2 + padded(sizeof(int) * 4) +
                      padded(sizeof(T) * npart);
  void* ws = nullptr;
  int rc = ctx->scratch_get(7, need, &ws);
  if (rc) return rc;
  char* cur = static_cast<char*>(ws);
  auto take = [&](size_t bytes) {
    char* r = cur;
    cur += padded(bytes);
    return r;
  };
  T* y = reinterpret_cast<T*>(take(sizeof(T) * pn));
  T* grad = reinterpret_cast<T*>(take(sizeof(T) * pn));
  T* R = reinterpret_cast<T*>(take(sizeof(T) * mn));
  T* zb = bt ? reinterpret_cast<T*>(take(sizeof(T) * pn)) : nullptr;
  T* diff = bt ? reinterpret_cast<T*>(take(sizeof(T) * pn)) : nullptr;
  T* Lsig = reinterpret_cast<T*>(take(sizeof(T) * n));
  T* best = reinterpret_cast<T*>(take(sizeof(T) * n));
  T* stopsum = reinterpret_cast<T*>(take(sizeof(T) * 2 * n));
  int* active = reinterpret_cast<int*>(take(sizeof(int) * n));
  int* pending = reinterpret_cast<int*>(take(sizeof(int) * n));
  int* npend = reinterpret_cast<int*>(take(sizeof(int) * 4));
  T* part = npart ? reinterpret_cast<T*>(take(sizeof(T) * npart)) : nullptr;
  T* DtX = gram ? reinterpret_cast<T*>(take(sizeof(T) * pn)) : nullptr;
  T* Gq = gram && bt ? reinterpret_cast<T*>(take(sizeof(T) * pn)) : nullptr;
  const T* G = static_cast<const T*>(d->G);
  const int ldg = d->ldg;
  (void)cur;

  ProxParams<T> pp;
  pp.lambda = static_cast<T>(o->lambda1);
  pp.lambda2d1 = static_cast<T>(o->lambda2) / pp.lambda;
  pp.regul = o->regul;
  pp.pos = o->pos;
  const int duality = o->regul == REG_L1;
  const int it0 = o->it0 > 1 ? o->it0 : 1;
  const int ista = o->ista;

  // everything after the workspace set-up, on stream `qs` (the context's
  // stream, or the capture stream when the loop is recorded as a graph)
  auto enqueue = [&](hipStream_t qs) -> int {
    SPX_HIP(hipMemcpyAsync(W, W0, sizeof(T) * pn, hipMemcpyDeviceToDevice, qs));
    SPX_HIP(hipMemcpyAsync(y, W0, sizeof(T) * pn, hipMemcpyDeviceToDevice, qs));
    hipLaunchKernelGGL(init_state_kernel<T>, lin(n), dim3(256), 0, qs, n,
                       static_cast<T>(o->L0), Lsig, active, best, info);
    SPX_HIP(hipGetLastError());

    // launch timers only for eager launches on the context's stream (a
    // captured loop records no events)
    const bool timed = qs == st;
    if (gram) {
      // D'X, once (SqLoss::init)
      ScopedTimer tm(ctx, timed ? SPX_TIMER_FISTA_PRODUCTS : -1);
      int r = gemm_tn_product<T>(qs, p, n, m, D, m, X, m, DtX, p, T(1), T(0), nullptr, 0, part,
                                 sG, prec);
      if (r) return r;
    }
    // gradient of the square loss at `at`: R = D*at - X, grad = D'R; with
    // the Gram matrix grad = G at - D'X
    auto gradient = [&](const T* at) -> int {
      ScopedTimer tm(ctx, timed ? SPX_TIMER_FISTA_PRODUCTS : -1);
      if (gram)
        return gemm_tn_product<T>(qs, p, n, p, G, ldg, at, p, grad, p, T(1), T(-1), DtX, p, part,
                                  sQ, prec);
      int r = gemm_tn_product<T>(qs, m, n, p, Dt, p, at, p, R, m, T(1), T(-1), X, m, part, sR,
                                 prec);
      if (r) return r;
      return gemm_tn_product<T>(qs, p, n, m, D, m, R, m, grad, p, T(1), T(0), nullptr, 0, part,
                                sG, prec);
    };
    // objective / gap evaluation at W
    auto evaluate = [&](int mode, int it) -> int {
      int r = gemm_tn_product<T>(qs, m, n, p, Dt, p, W, p, R, m, T(1), T(-1), X, m, part, sR,
                                 prec);
      if (r) return r;
      if (duality) {
        r = gemm_tn_product<T>(qs, p, n, m, D, m, R, m, grad, p, T(1), T(0), nullptr, 0, part,
                               sG, prec);
        if (r) return r;
      }
      hipLaunchKernelGGL(eval_kernel<T>, cols4(n), dim3(256), 0, qs, m, p, n, X,
                         W, R, grad, pp, duality, mode, o->tol, it, active, best, info);
      SPX_HIP(hipGetLastError());
      return SPX_OK;
    };

    double t = 1.0;
    int it;
    int rc2;
    for (it = 1; it <= o->max_it; ++it) {
      rc2 = gradient(ista ? W : y);
      if (rc2) return rc2;
      T beta = 0;
      if (!ista) {
        const T told = static_cast<T>(t);
        const T sq = std::sqrt(static_cast<T>(1 + 4 * told * told));
        const T tn = static_cast<T>((1.0 + static_cast<double>(sq)) / 2);
        beta = (T(1) - told) / tn;
        t = tn;
      }
      if (!bt) {
        ScopedTimer tm(ctx, timed ? SPX_TIMER_FISTA_PROX : -1);
        hipLaunchKernelGGL(prox_step_kernel<T>, cols4(n), dim3(256), 0, qs, p, n,
                           W, y, grad, Lsig, active, pp, beta, ista, stopsum);
        SPX_HIP(hipGetLastError());
      } else {
        const T* base = ista ? W : y;
        hipLaunchKernelGGL(set_pending_kernel, lin(n), dim3(256), 0, qs, n, active, pending);
        SPX_HIP(hipGetLastError());
        for (int trial = 1;; ++trial) {
          const int last = trial + 1 >= o->max_iter_backtracking;
          hipLaunchKernelGGL(prox_trial_kernel<T>, cols4(n), dim3(256), 0, qs, p,
                             n, base, grad, Lsig, pending, pp, zb, diff, stopsum);
          SPX_HIP(hipGetLastError());
          if (gram)
            rc2 = gemm_tn_product<T>(qs, p, n, p, G, ldg, diff, p, Gq, p, T(1), T(0), nullptr,
                                     0, part, sQ, prec);
          else
            rc2 = gemm_tn_product<T>(qs, m, n, p, Dt, p, diff, p, R, m, T(1), T(0), nullptr, 0,
                                     part, sR, prec);
          if (rc2) return rc2;
          hipLaunchKernelGGL(zero_one_kernel, dim3(1), dim3(1), 0, qs, npend);
          hipLaunchKernelGGL(backtrack_check_kernel<T>, cols4(n), dim3(256), 0, qs,
                             n, gram ? p : m, gram ? Gq : R, gram ? diff : nullptr, stopsum,
                             static_cast<T>(o->gamma), last, Lsig, pending, npend);
          SPX_HIP(hipGetLastError());
          int h = 0;
          SPX_HIP(hipMemcpyAsync(&h, npend, sizeof(int), hipMemcpyDeviceToHost, qs));
          SPX_HIP(hipStreamSynchronize(qs));
          if (h == 0 || last) break;
        }
        hipLaunchKernelGGL(momentum_kernel<T>, cols4(n), dim3(256), 0, qs, p, n,
                           W, y, zb, active, beta, ista, stopsum);
        SPX_HIP(hipGetLastError());
      }
      if (duality) {
        if (it % it0 == 0) {
          rc2 = evaluate(0, it);
          if (rc2) return rc2;
        }
      } else {
        hipLaunchKernelGGL(relstop_kernel<T>, lin(n), dim3(256), 0, qs, n, stopsum,
                           o->tol, it, active, info);
        SPX_HIP(hipGetLastError());
      }
    }
    hipLaunchKernelGGL(finish_it_kernel<T>, lin(n), dim3(256), 0, qs, n, active, it, info);
    SPX_HIP(hipGetLastError());
    return evaluate(1, it);
  };

  if (!(o->graph && !bt)) return enqueue(st);

  // ---- fixed step: the whole loop as one graph, captured once per key
  struct Key {
    int elem, m, p, n, max_it, regul, pos, ista, it0, prec, gram;
    const void *X, *W0, *W, *info, *D, *Dt, *ws;
    double L0, lambda1, lambda2, tol;
  } key;
  std::memset(&key, 0, sizeof(key));
  key.elem = sizeof(T);
  key.m = m;
  key.p = p;
  key.n = n;
  key.max_it = o->max_it;
  key.regul = o->regul;
  key.pos = o->pos;
  key.ista = ista;
  key.it0 = it0;
  key.prec = prec;
  key.gram = gram;
  key.X = X;
  key.W0 = W0;
  key.W = W;
  key.info = info;
  key.D = D;
  key.Dt = Dt;
  key.ws = ws;
  key.L0 = o->L0;
  key.lambda1 = o->lambda1;
  key.lambda2 = o->lambda2;
  key.tol = o->tol;
  const char* kb = reinterpret_cast<const char*>(&key);
  hipGraphExec_t exec = nullptr;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (size_t i = 0; i < ctx->graphs.size(); ++i) {
      auto& g = ctx->graphs[i];
      if (g.key.size() == sizeof(key) && std::memcmp(g.key.data(), kb, sizeof(key)) == 0) {
        exec = g.exec;
        std::rotate(ctx->graphs.begin() + i, ctx->graphs.begin() + i + 1, ctx->graphs.end());
        break;
      }
    }
  }
  if (!exec) {
    if (!ctx->capture_stream)
      SPX_HIP(hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
    hipStream_t cs = ctx->capture_stream;
    SPX_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    rc = enqueue(cs);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &graph);
    if (rc || ec != hipSuccess || !graph) {
      if (graph) (void)hipGraphDestroy(graph);
      if (!rc) set_error("fistaFlat: graph capture failed: %s", hipGetErrorString(ec));
      return rc ? rc : SPX_ERR_HIP;
    }
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      set_error("fistaFlat: graph instantiation failed: %s", hipGetErrorString(ei));
      return SPX_ERR_HIP;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->graphs.size() >= 8) {
      (void)hipGraphExecDestroy(ctx->graphs.front().exec);
      ctx->graphs.erase(ctx->graphs.begin());
    }
    Ctx::CachedGraph cg;
    cg.key.assign(kb, kb + sizeof(key));
    cg.exec = exec;
    ctx->graphs.push_back(std::move(cg));
  }
  SPX_HIP(hipGraphLaunch(exec, st));
  return SPX_OK;
}

template int fista_run<float>(Ctx*, const Dict*, const float*, int, const float*, float*,
                              const spx_fista_opts*, float*);
template int fista_run<double>(Ctx*, const Dict*, const double*, int, const double*, double*,
                               const spx_fista_opts*, double*);

}  // namespace spx
