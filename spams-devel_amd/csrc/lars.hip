// LARS-Lasso solver: launch logic of the batched and plain forms over the
// kernels of lars_impl.h (plain penalty here; the weighted-penalty and
// masked-signal kernels are compiled in lars_w.hip and lars_m.hip).
#include "lars_impl.h"

namespace spx {

namespace {

template <typename T>
int lars_pass_w(Ctx* ctx, int w, bool fast, const LarsArgs<T>& a, bool persistent,
                int* counter, long max_blocks) {
  if (a.weights) return lars_weighted_pass<T>(ctx, w, fast, a, persistent, counter, max_blocks);
  if (a.mobs) return lars_masked_pass<T>(ctx, w, fast, a, persistent, counter, max_blocks);
  return lars_pass_wt<T, 0>(ctx, w, fast, a, persistent, counter, max_blocks);
}

// Batched form: fast pass (capacity FastCap) over every signal in work order,
// then the full-capacity pass over the signals that outgrew it, with at least
// 4 wavefronts per signal (their active sets are the large ones).
template <typename T>
int lars_batched(Ctx* ctx, const LarsArgs<T>& a0, int w) {
  int* cnt = nullptr;
  int rc = ctx->scratch_get(1, 4 * sizeof(int), reinterpret_cast<void**>(&cnt));
  if (rc) return rc;
  int* ovf = nullptr;
  rc = ctx->scratch_get(8, sizeof(int) * (size_t)(a0.n > 0 ? a0.n : 1),
                        reinterpret_cast<void**>(&ovf));
  if (rc) return rc;
  rc = zero_ints(ctx->stream, cnt, 4);
  if (rc) return rc;
  LarsArgs<T> a = a0;
  if (a.chunk <= 0) a.chunk = 1;
  a.ovf_list = ovf;
  a.ovf_count = cnt + 1;
  if (a0.L > FastCap<T>::C && a0.n > 1) {
    T* key = nullptr;
    rc = ctx->scratch_get(14, sizeof(T) * (size_t)a0.n, reinterpret_cast<void**>(&key));
    if (rc) return rc;
    a.ovf_key = key;
  }
  a.list = nullptr;
  a.list_count = nullptr;
  if (a0.L > FastCap<T>::C) {
    // one hand-over slot per signal (only overflowing signals write theirs),
    // so whether a signal resumes never depends on the batch around it
    const size_t sb = LarsState<T>::bytes(a0.ldg);
    void* stp = nullptr;
    rc = ctx->scratch_get(13, sb * (size_t)a0.n, &stp);
    if (rc) return rc;
    a.state = stp;
    a.state_bytes = sb;
  }
  if (a0.nlam > 1 && !a.order) {
    // longest paths first: the homotopy is longest for the smallest lambda
    int* order = nullptr;
    rc = ctx->scratch_get(10, sizeof(int) * (size_t)a0.n, reinterpret_cast<void**>(&order));
    if (rc) return rc;
    ScopedTimer tm(ctx, SPX_TIMER_LARS_ORDER);
    rc = sort_order<T>(ctx, a0.lambda, a0.n, order);
    if (rc) return rc;
    a.order = order;
  }
  {
    ScopedTimer tm(ctx, SPX_TIMER_LARS_FAST);
    rc = lars_pass_w<T>(ctx, w, true, a, true, cnt, a0.n);
    if (rc) return rc;
  }
  if (a0.L <= FastCap<T>::C) return SPX_OK;  // every active set fitted
  a.list = ovf;
  a.list_count = cnt + 1;
  if (a.ovf_key) {
    // longest remaining path first (stable sort of the listed signals by
    // their key; unused entries sort last)
    ScopedTimer tm(ctx, SPX_TIMER_LARS_LIST);
    int* sorted = nullptr;
    rc = ctx->scratch_get(15, sizeof(int) * (size_t)a0.n, reinterpret_cast<void**>(&sorted));
    if (rc) return rc;
    rc = sort_list<T>(ctx, a.ovf_key, ovf, cnt + 1, a0.n, sorted);
    if (rc) return rc;
    a.list = sorted;
  }
  int w2 = w < 4 ? 4 : w;
  ScopedTimer tm(ctx, SPX_TIMER_LARS_FULL);
  rc = lars_pass_w<T>(ctx, w2, false, a, true, cnt + 2, a0.n);
  if (rc < 0) rc = lars_pass_w<T>(ctx, 16, false, a, true, cnt + 2, a0.n);
  return rc;
}

// Whether group width W covers p for variant var (0 plain, 1 weighted, 2
// masked): its row chunks reach p and the LDS pass's layout (the group's
// correlation rows, the weighted variant's weights, the capacity-FastCap
// inverse) fits the 160 KiB of LDS; past that no active-set size does.
template <typename T, int W>
bool lars_fits_w(int p, int var) {
  constexpr int C = FastCap<T>::C;
  constexpr size_t kLds = 160 * 1024;
  switch (lars_nq<T, W>(p)) {
    case 1: return LarsLds<T, W, 1, C>::bytes(C, 0, var) <= kLds;
    case 2: return LarsLds<T, W, 2, C>::bytes(C, 0, var) <= kLds;
    case 4: return LarsLds<T, W, 4, C>::bytes(C, 0, var) <= kLds;
    case 8: return LarsLds<T, W, 8, C>::bytes(C, 0, var) <= kLds;
    case 16: return LarsLds<T, W, 16, C>::bytes(C, 0, var) <= kLds;
    default: return false;
  }
}

template <typename T>
bool lars_fits_var(int w, int p, int var) {
  switch (w) {
    case 1: return lars_fits_w<T, 1>(p, var);
    case 2: return lars_fits_w<T, 2>(p, var);
    case 4: return lars_fits_w<T, 4>(p, var);
    case 16: return lars_fits_w<T, 16>(p, var);
    default: return false;
  }
}

}  // namespace

// Group width `waves` as resolved by the caller (explicit, the dictionary's
// autotuned width or the size rule; 4 for the plain form), widened until the
// group covers p.
template <typename T>
int lars_solve(Ctx* ctx, const LarsArgs<T>& a, int form, int waves) {
  if (waves <= 0) waves = 4;
  static const int widths[] = {1, 2, 4, 16};
  for (int w : widths) {
    if (w < waves) continue;
    int rc;
    if (form == SPX_FORM_BATCHED) {
      if (!lars_fits_var<T>(w, a.p, a.weights ? 1 : a.mobs ? 2 : 0)) continue;
      rc = lars_batched<T>(ctx, a, w);
    } else {
      rc = lars_pass_w<T>(ctx, w, false, a, false, nullptr, 0);
    }
    if (rc >= 0) return rc;
  }
  set_error("lasso: p=%d exceeds the largest supported dictionary", a.p);
  return SPX_ERR_UNSUP;
}

template <typename T>
bool lars_fits(int w, int p) {
  return lars_fits_var<T>(w, p, 0);
}

template bool lars_fits<float>(int, int);
template bool lars_fits<double>(int, int);
template int lars_solve<float>(Ctx*, const LarsArgs<float>&, int, int);
template int lars_solve<double>(Ctx*, const LarsArgs<double>&, int, int);

}  // namespace spx
