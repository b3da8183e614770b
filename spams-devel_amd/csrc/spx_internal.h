// Internal declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <vector>

#include "spx.h"

struct spx_ctx;
struct spx_dict;

namespace spx {

typedef ::spx_ctx Ctx;
typedef ::spx_dict Dict;

// Kernel arguments of the LARS solver (coreLARS2 semantics).
template <typename T>
struct LarsArgs {
  const T* G;       // p x p Gram with the ridge on its diagonal, leading dim ldg
  int ldg;
  int gsym;         // G is exactly symmetric: G(c, k) is read from column k
  T diag;           // the ridge max(lambda2, 1e-10), which G carries on its
                    // diagonal; added to the masked Gram columns formed over
                    // the observed rows (0 when G is used as given)
  const T* DtX;     // p x n, leading dim ldd
  int ldd;
  const T* norms;   // ||x_i||^2, n
  const T* lambda;  // nlam (1 or n)
  int nlam;
  int p, n, L, mode, pos, length_path;
  T* outV;          // L x n
  int* outR;        // L x n
  int* outCount;    // n
  int* stats;       // 2 x n or null
  T* scratch;       // per workgroup 2 * L * L (full-capacity pass)
  int* queue;       // work counter (batched form) or null (plain form)
  // Batched form, two passes: the fast pass (active set capped at a small
  // LDS capacity) appends the signals that outgrow it to ovf_list /
  // ovf_count; the full-capacity pass then solves exactly those, reading the
  // list (list / list_count) instead of 0..n-1.
  int* ovf_list;
  int* ovf_count;
  const int* list;
  const int* list_count;
  int chunk;        // signals taken per atomic grab of the work counter
  const int* order; // work order (signal index per work item) or null
  int refit;        // FP64 refit of the final active set (PENALTY mode)
  // Per-signal state slots (state_bytes each) through which the LDS pass
  // hands its overflowing signals to the full-capacity pass, which resumes
  // them; null = the full pass solves every listed signal from the start.
  void* state;
  size_t state_bytes;
  T* ovf_key;       // per list position: work key of the overflowing signal
  int lds_cap;      // full-capacity pass: LDS capacity of the inverse (0 = HBM only)
  T* path;          // signal 0's regularisation path (p x path_cols) or null
  int path_cols;
  int ols;          // least-squares refit of the final active set
  // Weighted penalty (coreLARS2W): per-signal atom weights, p x n with
  // leading dimension ldw, or null for the plain penalty.
  const T* weights;
  int ldw;
  T tol;            // stop tolerance of |step|, normX and the mode margins
  // Cholesky variant (coreLARS, decomp/decomp.h:844-1093): no residual cap
  // on the L2ERROR step, its stop tested after admissions only, and an exit
  // on a residual increase
  int chol;
  // Masked signals (lasso_mask): per-signal count of observed rows (0: empty
  // column; m: the plain path), per signal the rows its Gram columns visit
  // (m x n, mask_gram.h) and the dictionary (D m x p, Dt p x m) from which a
  // partly observed signal's Gram columns over its observed rows are formed
  // into its workgroup's slot of `ga` (ga_stride elements per workgroup,
  // columns of ldg).
  const int* mobs;
  const int* mrows;
  const T* Dd;
  const T* Dt;
  int m;
  T* ga;
  size_t ga_stride;
};

// Kernel arguments of the ORMP solver (coreORMP semantics).
template <typename T>
struct OmpArgs {
  const T* G;       // p x p Gram (no diagonal shift), leading dim ldg
  int ldg;
  int gsym;         // G is exactly symmetric: G(c, k) is read from column k
  const T* DtX;     // p x n
  int ldd;
  const T* norms;   // n
  const int* Ls;    // per-signal or scalar L (nL entries; 0 = not given)
  int nL;
  const T* eps;
  int nEps;
  const T* lambda;
  int nLambda;
  int scalarL;      // L used when not given (min(m, p))
  int p, n, L;      // L: leading dimension of the outputs (max L, clamped)
  T* outV;
  int* outR;
  int* outCount;
  T* scratch;       // per workgroup: L*L (Un) + p*L (Undn)
  int* queue;
  int chunk;        // signals taken per atomic grab of the work counter
  T* path;          // signal 0's path (p x path_cols) or null
  int path_cols;
  // Masked signals (omp_mask): as in LarsArgs; a partly observed signal's
  // Gram columns (ridge 1e-10) fill its workgroup's slot of ga in selection
  // order.
  const int* mobs;
  const int* mrows;
  const T* Dd;
  const T* Dt;
  int m;
  T* ga;
  size_t ga_stride;
};

template <typename T>
int gemm_tn(hipStream_t st, int M, int N, int K, const T* A, int lda,
            const T* B, int ldb, T* C, int ldc, T alpha, T beta, const T* E,
            int lde, int precision);
// C (M x N, leading dimension M) = (accumulate ? C : 0) + A'B, with the
// inner dimension split into about S slices computed in one launch (grid z)
// into `part` (S x M x N elements) and summed in slice order.
template <typename T>
int gemm_tn_split(hipStream_t st, int M, int N, int K, const T* A, int lda, const T* B, int ldb,
                  T* C, int accumulate, T* part, int S, int precision);
// gemm_tn's C = alpha A'B + beta E (C leading dimension ldc), split over the
// inner dimension into S slices when S > 1 (`part`: S x M x N elements), the
// slices summed in order; S <= 1 is gemm_tn itself.
template <typename T>
int gemm_tn_product(hipStream_t st, int M, int N, int K, const T* A, int lda, const T* B,
                    int ldb, T* C, int ldc, T alpha, T beta, const T* E, int lde, T* part,
                    int S, int precision);
// Slices that fill the chip for an M x N product with inner dimension K.
int gemm_split_count(int M, int N, int K);
// gemm_split_count for products with fewer than 128 output tiles, else 1.
int gemm_small_split(int M, int N, int K);
template <typename T>
int col_sqnorms(hipStream_t st, int m, int n, const T* X, int ldx, T* out);
template <typename T>
int lars_solve(Ctx* ctx, const LarsArgs<T>& a, int form, int waves);
// One LARS pass of the weighted-penalty kernels (group width w; fast = LDS
// capacity pass); -1 when the width does not cover p.
template <typename T>
int lars_weighted_pass(Ctx* ctx, int w, bool fast, const LarsArgs<T>& a, bool persistent,
                       int* counter, long max_blocks);
// The same for the masked-signal kernels (a.mobs set).
template <typename T>
int lars_masked_pass(Ctx* ctx, int w, bool fast, const LarsArgs<T>& a, bool persistent,
                     int* counter, long max_blocks);
template <typename T>
int omp_solve(Ctx* ctx, const OmpArgs<T>& a, int form, int waves);
template <typename T>
int lists_to_csc(Ctx* ctx, const T* v, const int* r, const int* count, int L,
                 int n, int* pB, int* rOut, T* vOut, long long* nnz);
template <typename T>
int csc_count_scan(Ctx* ctx, const int* count, int n, int* pB, long long* nnz);
template <typename T>
int csc_fill(Ctx* ctx, const T* v, const int* r, int L, int n, const int* pB,
             int* rOut, T* vOut);
template <typename T>
int fista_run(Ctx* ctx, const Dict* dict, const T* X, int n, const T* W0,
              T* W, const spx_fista_opts* o, T* info);

// order[i] = indices 0..n-1 sorted by key ascending (stable radix sort on
// the 16 leading bits of the keys rounded to FP32, NaN last: a work order,
// never a result); scratch slot 9.
template <typename T>
int sort_order(Ctx* ctx, const T* key, int n, int* order);

// out[0:count) = list[0:count) stably sorted by key[0:count) ascending (as
// sort_order), where count = *count_dev (device value <= n); scratch slot 9.
template <typename T>
int sort_list(Ctx* ctx, const T* key, const int* list, const int* count_dev, int n, int* out);

// Scratch bytes the sorts of n entries take (slot 9).
size_t sorted_scratch_bytes(int n);

// Masked-signal set-up (one wavefront per signal): Xm (m x n) = X with the
// unobserved rows (B == 0) zeroed, obs[i] = signal i's count of observed
// rows, rows (m x n) = per signal the rows its masked Gram columns visit
// (mask_gram.h), and s_out[i] = s_in[i or 0] (0 when ns == 0), times
// obs[i] / m when `scale` and 0 < obs[i] < m.
template <typename T>
int mask_prep(hipStream_t st, const T* X, int ldx, int m, int n, const unsigned char* B, int ldb,
              T* Xm, int* obs, int* rows, const T* s_in, int ns, T* s_out, int scale);

// Clears n ints with a kernel (capture-safe replacement for a memset node).
int zero_ints(hipStream_t st, int* p, int n);

// Wavefronts per signal for a batched solve (solver 0 = LARS, 1 = OMP): the
// width the autotuner stored on the dictionary, else a size rule (one
// wavefront while a wavefront covers p with at most 16 chunks per lane).
int tuned_waves(const Dict* d, int solver);

// True when a workgroup of w wavefronts covers p rows (the solver's register
// and LDS layout), i.e. the width can run this dictionary.
template <typename T>
bool lars_fits(int w, int p);
template <typename T>
bool ormp_fits(int w, int p);

// Records the duration of the enclosed launches on the context's stream
// under timer `id` (SPX_TIMER_*) while timing is switched on; no-op else,
// and for a negative id.
class ScopedTimer {
 public:
  ScopedTimer(Ctx* ctx, int id);
  ~ScopedTimer();

 private:
  Ctx* ctx_;
  int id_;
  hipEvent_t a_ = nullptr;
};

}  // namespace spx

// Context: one device, one stream, grow-only scratch slots.
struct spx_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  struct Slot {
    void* ptr = nullptr;
    size_t bytes = 0;
  };
  std::vector<Slot> slots;
  std::mutex mu;
  // Allocation-free calls: while `arena` is set, scratch_get carves every
  // slot from it (bump order, 256-byte aligned, a slot re-requested with no
  // more bytes keeps its region) and fails past arena_bytes instead of
  // allocating; launch timers record nothing.
  char* arena = nullptr;
  size_t arena_bytes = 0, arena_used = 0;
  std::vector<Slot> arena_slots;
  // bytes of the arena still free (SIZE_MAX when no arena is set)
  size_t scratch_room() const { return arena ? arena_bytes - arena_used : ~size_t(0); }
  // launch timing (spx_ctx_set_timing): pending event pairs per timer id,
  // folded into the totals when read
  bool timing = false;
  struct Pending {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;
  double timer_ms[SPX_TIMER_COUNT] = {};
  long long timer_n[SPX_TIMER_COUNT] = {};
  hipEvent_t event_get();
  int timers_fold();
  // captured FISTA iteration loops (spx_fista_opts.graph): key bytes ->
  // instantiated graph, most recent last; capture stream created on demand
  struct CachedGraph {
    std::vector<char> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<CachedGraph> graphs;
  hipStream_t capture_stream = nullptr;
  // Device scratch slot `id` of at least `bytes` bytes (contents undefined).
  int scratch_get(int id, size_t bytes, void** out);
  ~spx_ctx();
};

// Resident dictionary: D (m x p), D' (p x m) and G = D'D (p x p, leading
// dimension ldg = pad_ld(p), zero padding rows) in device memory.
struct spx_dict {
  spx_ctx* ctx = nullptr;  // creating context (not used after creation)
  int device = 0;
  int m = 0, p = 0, ldg = 0;
  int elem = 4;         // 4 = float, 8 = double
  int precision = 0;    // precision G was formed with
  void* D = nullptr;    // m x p
  void* Dt = nullptr;   // p x m
  void* G = nullptr;    // ldg x p
  // G + 1e-10 I: the Gram the LARS solver uses for lambda2 <= 1e-10 (the
  // ridge max(lambda2, 1e-10) of decomp/decomp.h:1566); null for a Gram-only
  // dictionary (used as given)
  void* Gr = nullptr;
  double lipschitz = 0; // largest eigenvalue of G (power iteration), lazily
  int waves[2] = {0, 0};     // autotuned wavefronts per signal (LARS, OMP)
  int gram_only = 0;         // created from a Gram matrix (no D; no ridge)
  float tune_ms[2][4] = {};  // autotuner trial times for 1, 2, 4, 16 waves
};
