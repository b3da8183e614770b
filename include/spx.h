/* spx -- MI355X-native sparse coding engine: C ABI.
 *
 * This is the drop-in boundary of the engine.  Every entry point takes plain
 * pointers and sizes (column-major, `int` dimensions, CSC with `int` indices),
 * returns an int status (0 on success) and leaves a message for
 * spx_last_error() on failure.  The C++ header spams_amd.h re-creates the
 * reference's spams.h templates (_lassoD, _lassoQq, _lassoMask,
 * _lassoWeighted, _omp, _ompMask, _fistaFlat) on top of it.
 *
 * Reference interfaces replaced (paths relative to the SPAMS tree):
 *   spx_lasso_*       _lassoD      cpp_library/spams.h:107-143
 *                     lasso2       decomp/decomp.h:1563-1646 (+coreLARS2 :1706)
 *   spx_lasso_qq_*    _lassoQq     cpp_library/spams.h:145-188
 *   spx_lasso_mask_*  _lassoMask   cpp_library/spams.h:190-215
 *                     lasso_mask   decomp/decomp.h:1466-1561
 *   spx_lasso_weighted_*  _lassoWeighted  cpp_library/spams.h:217-249
 *                     lassoWeight  decomp/decomp.h:1315-1391 (+coreLARS2W :1917)
 *   spx_omp_*         _omp         cpp_library/spams.h:250-293
 *                     omp          decomp/decomp.h:322-404 (+coreORMP :538)
 *   spx_omp_mask_*    _ompMask     cpp_library/spams.h:295-343
 *                     omp_mask     decomp/decomp.h:406-519
 *   spx_fista_flat_*  _fistaFlat   cpp_library/spams.h:460-608
 *                     FISTA::solver prox/fista.h:3618-3752
 *   spx_csc           SpMatrix<T>  linalg/linalg.h:775-934 (pB/pE/r/v CSC)
 *
 * Two families of entry points:
 *   - host entry points (spx_lasso_f32, ...) take host buffers and return a
 *     host CSC matrix allocated by the library (free with spx_csc_free), like
 *     the reference's heap-allocated SpMatrix;
 *   - device entry points (suffix _dev) take device pointers that are already
 *     resident in HBM and run asynchronously on the context's stream.
 */
#ifndef SPX_H
#define SPX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SPX_OK 0
#define SPX_ERR_ARG 1      /* invalid argument (dimension mismatch, ...) */
#define SPX_ERR_HIP 2      /* HIP runtime failure */
#define SPX_ERR_NOMEM 3    /* device allocation / admission budget exceeded */
#define SPX_ERR_UNSUP 4    /* option combination not supported */

/* constraint_type (decomp/decomp.h:95) */
#define SPX_L1COEFFS 0
#define SPX_L2ERROR 1
#define SPX_PENALTY 2

/* arithmetic of the D'X / Gram / FISTA products (f32 entry points) */
#define SPX_PREC_FP32 0    /* FP32 MFMA (v_mfma_f32_16x16x4_f32)           */
#define SPX_PREC_BF16X3 1  /* BF16 hi/lo split, 3 MFMAs, FP32 accumulation  */
/* flag, combinable with the above (lasso, PENALTY mode): once the homotopy
 * reaches lambda, refine the coefficients on the final active set to the
 * FP64 solution of (G_AA + d I) a = (D'x)_A - lambda sign(a) by two steps of
 * mixed-precision iterative refinement (FP64 residual from the resident Gram,
 * the solver's own inverse as preconditioner). */
#define SPX_PREC_REFIT64 2

/* solver launch form */
#define SPX_FORM_BATCHED 0 /* persistent kernel + device work queue (default) */
#define SPX_FORM_PLAIN 1   /* a grid of resident workgroups, one signal each at
                            * a time, assigned grid-stride (no queue, no order) */

typedef struct spx_ctx spx_ctx;
typedef struct spx_dict spx_dict;

/* CSC matrix, rows sorted ascending within each column (the output format of
 * SpMatrix<T>::convert, linalg/linalg.h:4840-4862).  pB has n+1 entries;
 * pE of the reference is pB+1.  `v` is float* or double* by entry point. */
typedef struct {
  int m;          /* rows (= p, atoms)   */
  int n;          /* columns (= signals) */
  long long nnz;
  int* pB;
  int* r;
  void* v;
} spx_csc;

/* Options of spx_lasso_*.  Field meaning follows _lassoD's arguments
 * (cpp_library/spams.h:107-112); -1 selects the same defaults. */
typedef struct {
  int L;               /* max active set, -1 -> p (then clamped as _lassoD) */
  int mode;            /* SPX_L1COEFFS / SPX_L2ERROR / SPX_PENALTY           */
  int pos;             /* positivity constraint                              */
  int max_length_path; /* homotopy step cap, -1 -> 4*L (before clamping)     */
  int precision;       /* SPX_PREC_* (f32 entry points only)                 */
  int form;            /* SPX_FORM_*                                         */
  int waves_per_signal;/* 0 = autotuned; else 1, 2, 4 or 16                  */
  int ols;             /* 1: least-squares fit on the final active set (the
                        * facade's ols, decomp/decomp.h:1087-1091); implies
                        * cholesky, whose path then tracks the least-squares
                        * residual of the active set (the L2ERROR stop reads
                        * it, decomp/decomp.h:1044-1050) and has no exit on a
                        * residual increase */
  int cholesky;        /* 1: the path of the Cholesky variant coreLARS
                        * (decomp/decomp.h:844-1093, _lassoD's cholesky): in
                        * L2ERROR mode the last step is not cut at the
                        * residual target; a step that would increase the
                        * residual ends the path */
} spx_lasso_opts;

/* Options of spx_omp_* (the _omp facade's vector arguments are passed
 * separately: NULL/0 = not given, size 1 = scalar, size n = per signal). */
typedef struct {
  int precision;
  int form;
  int waves_per_signal;
} spx_omp_opts;

/* Options of spx_fista_flat_*.  Field meaning follows _fistaFlat
 * (cpp_library/spams.h:460-499) restricted to the square loss and the vector
 * regularisers "l1" and "elastic-net". */
typedef struct {
  int max_it;
  double L0;
  int fixed_step;
  double gamma;
  double lambda1;
  double lambda2;
  double tol;
  int it0;
  int max_iter_backtracking;
  int regul;           /* 0 = "l1", 1 = "elastic-net" */
  int pos;
  int ista;
  int precision;       /* SPX_PREC_* for the two products per iteration */
  int graph;           /* fixed step: replay the whole iteration loop as one
                        * hipGraph, captured once per (buffers, sizes,
                        * options) and cached on the context */
  int compute_gram;    /* gradient G alpha - D'x with G = D'D and D'X formed
                        * once, backtracking test (y-z)'G(y-z)
                        * (prox/fista.h:364-413); 0 = D'(D alpha - x) */
} spx_fista_opts;

/* ---------------------------------------------------------------- context */

/* Creates a context on HIP device `device` (the calling thread's current
 * device is left unchanged).  The context owns a stream and a pool of device
 * scratch buffers. */
int spx_ctx_create(spx_ctx** out, int device);
void spx_ctx_destroy(spx_ctx* ctx);
/* Run the context's work on a caller-owned hipStream_t (NULL = the HIP null
 * stream).  spx_ctx_own_stream() returns to a stream owned by the context. */
int spx_ctx_set_stream(spx_ctx* ctx, void* hip_stream);
int spx_ctx_own_stream(spx_ctx* ctx);
int spx_ctx_synchronize(spx_ctx* ctx);
/* Device scratch the context holds (bytes), and its release (waits for the
 * context's stream, then frees every scratch buffer; later calls allocate
 * again).  Engine controls with no reference counterpart: a caller that
 * budgets device memory (the served form) counts and trims them. */
size_t spx_ctx_scratch_bytes(const spx_ctx* ctx);
int spx_ctx_release_scratch(spx_ctx* ctx);

/* Launch timing.  While on, the batched solvers, FISTA and the CSC assembly
 * record HIP events on the context's stream around their launches, under
 * these timers: */
enum {
  SPX_TIMER_LARS_ORDER = 0, /* work-order sort by lambda (longest path first) */
  SPX_TIMER_LARS_FAST = 1,  /* LARS pass with the active set in LDS */
  SPX_TIMER_LARS_FULL = 2,  /* LARS pass over the signals that outgrew LDS */
  SPX_TIMER_ORMP = 3,       /* ORMP solver */
  SPX_TIMER_LARS_LIST = 4,  /* ordering of the overflow list (longest first) */
  SPX_TIMER_FISTA_PROX = 5, /* FISTA fused prox + momentum pass (fixed step,
                             * eager launches; a replayed graph is not timed) */
  SPX_TIMER_FISTA_PRODUCTS = 6, /* FISTA's two products per iteration (same) */
  SPX_TIMER_CSC = 7,        /* device CSC assembly: count scan and column fill */
  SPX_TIMER_COUNT = 8
};
int spx_ctx_set_timing(spx_ctx* ctx, int on);
/* Totals since the last reset: ms[i] and launches[i] for i < count (either
 * may be NULL).  Waits for the recorded events. */
int spx_ctx_read_timing(spx_ctx* ctx, double* ms, long long* launches, int count);
int spx_ctx_reset_timing(spx_ctx* ctx);
/* Thread-local message of the last failing call on this thread. */
const char* spx_last_error(void);
/* Library version string. */
const char* spx_version(void);

/* ------------------------------------------------------------ dictionaries */

/* A resident dictionary: D (m x p, column-major) and its Gram matrix
 * G = D'D (computed once on the device with the chosen precision) kept in
 * HBM.  `D` is a host pointer when on_device == 0, else a device pointer
 * (copied).  Replaces the materialised ProdMatrix G (linalg/linalg.h:
 * 5090-5142, decomp/decomp.h:1566 and :342). */
int spx_dict_create_f32(spx_ctx* ctx, const float* D, int m, int p,
                        int on_device, int precision, spx_dict** out);
int spx_dict_create_f64(spx_ctx* ctx, const double* D, int m, int p,
                        int on_device, spx_dict** out);
/* A resident Gram matrix Q (p x p, column-major) without a dictionary, for
 * the Gram-given LARS (spx_lasso_dtx_dev_* with correlations the caller
 * computed): Q is used as given, no ridge is added; m is the signal length.
 * Entry points that need D (products, FISTA, OMP) reject it. */
int spx_dict_create_gram_f32(spx_ctx* ctx, const float* Q, int m, int p,
                             int on_device, spx_dict** out);
int spx_dict_create_gram_f64(spx_ctx* ctx, const double* Q, int m, int p,
                             int on_device, spx_dict** out);
void spx_dict_destroy(spx_dict* dict);
int spx_dict_dims(const spx_dict* dict, int* m, int* p);
/* Device bytes held by the dictionary (D, D', G). */
size_t spx_dict_bytes(const spx_dict* dict);

/* Autotuner (no counterpart in the reference, whose OpenMP loop has one
 * thread per signal): times the batched solver at 1, 2, 4 and 16 wavefronts
 * per signal on the first min(n, 65536) signals of a sample batch and stores
 * the fastest width on the dictionary; later batched calls on the dictionary
 * with waves_per_signal <= 0 use it (results depend on the width only through
 * FP32 summation order).  `ms` (optional, 4 entries) receives each width's
 * time, 0 for widths that cannot cover p.  Synchronous. */
int spx_dict_autotune_lasso_f32(spx_ctx* ctx, spx_dict* dict, const float* X,
                                int ldx, int n, const float* lambda,
                                int n_lambda, float lambda2,
                                const spx_lasso_opts* opts, int* waves,
                                float* ms);
int spx_dict_autotune_lasso_f64(spx_ctx* ctx, spx_dict* dict, const double* X,
                                int ldx, int n, const double* lambda,
                                int n_lambda, double lambda2,
                                const spx_lasso_opts* opts, int* waves,
                                float* ms);
int spx_dict_autotune_omp_f32(spx_ctx* ctx, spx_dict* dict, const float* X,
                              int ldx, int n, const int* L, int nL,
                              const float* eps, int n_eps, const float* lambda,
                              int n_lambda, int* waves, float* ms);
int spx_dict_autotune_omp_f64(spx_ctx* ctx, spx_dict* dict, const double* X,
                              int ldx, int n, const int* L, int nL,
                              const double* eps, int n_eps,
                              const double* lambda, int n_lambda, int* waves,
                              float* ms);
/* Width used by batched calls on the dictionary (solver 0 = lasso, 1 = omp)
 * when waves_per_signal <= 0; spx_dict_set_waves pins it (0 = size rule). */
int spx_dict_get_waves(const spx_dict* dict, int solver, int* waves);
int spx_dict_set_waves(spx_dict* dict, int solver, int waves);

/* ------------------------------------------------------- host entry points */

/* LARS-Lasso: replaces _lassoD (cpp_library/spams.h:107-143) with
 * cholesky=false (lasso2 + coreLARS2).  `lambda` holds n_lambda constraints:
 * 1 (the reference's scalar) or n (one per signal).  lambda2 adds
 * max(lambda2, 1e-10) to the Gram diagonal (decomp/decomp.h:1569). */
int spx_lasso_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D,
                  int p, const float* lambda, int n_lambda, float lambda2,
                  const spx_lasso_opts* opts, spx_csc* out);
/* Gram-given LARS: replaces _lassoQq (cpp_library/spams.h:145-188), i.e.
 * lasso2(X, Q, q, ...) (decomp/decomp.h:1574-1646).  X (m x n) only gives
 * ||x_i||^2; Q (p x p) is the Gram matrix used as given (no ridge; lambda2
 * only enters the facade's L clamp); q (p x n) holds the correlations.
 * `path` as in spx_lasso_path_* (may be NULL). */
int spx_lasso_qq_f32(spx_ctx* ctx, const float* X, int m, int n, const float* Q,
                     const float* q, int p, const float* lambda, int n_lambda,
                     float lambda2, const spx_lasso_opts* opts, spx_csc* out,
                     float* path);
int spx_lasso_qq_f64(spx_ctx* ctx, const double* X, int m, int n,
                     const double* Q, const double* q, int p,
                     const double* lambda, int n_lambda, double lambda2,
                     const spx_lasso_opts* opts, spx_csc* out, double* path);
/* As spx_lasso_*, with _lassoD's return_reg_path: `path` (host, p x
 * max_length_path after the facade's normalisation, column-major) receives
 * the first signal's regularisation path. */
int spx_lasso_path_f32(spx_ctx* ctx, const float* X, int m, int n,
                       const float* D, int p, const float* lambda,
                       int n_lambda, float lambda2, const spx_lasso_opts* opts,
                       spx_csc* out, float* path);
int spx_lasso_path_f64(spx_ctx* ctx, const double* X, int m, int n,
                       const double* D, int p, const double* lambda,
                       int n_lambda, double lambda2,
                       const spx_lasso_opts* opts, spx_csc* out, double* path);
int spx_lasso_f64(spx_ctx* ctx, const double* X, int m, int n,
                  const double* D, int p, const double* lambda, int n_lambda,
                  double lambda2, const spx_lasso_opts* opts, spx_csc* out);
/* LARS over partly observed signals: replaces _lassoMask
 * (cpp_library/spams.h:190-215), i.e. lasso_mask (decomp/decomp.h:1466-1561).
 * B: host m x n bytes, nonzero = row observed.  A signal with every row
 * observed runs the plain path; with none, its column is empty; otherwise
 * the path over its observed rows (Gram columns of those rows, ridge
 * max(lambda2, 1e-10)) with the constraint scaled by the observed fraction
 * (PENALTY, L2ERROR).  Step cap 4 min(L, p); max_length_path, ols and
 * SPX_PREC_REFIT64 do not apply (the last two are rejected). */
int spx_lasso_mask_f32(spx_ctx* ctx, const float* X, int m, int n,
                       const float* D, int p, const unsigned char* B,
                       const float* lambda, int n_lambda, float lambda2,
                       const spx_lasso_opts* opts, spx_csc* out);
int spx_lasso_mask_f64(spx_ctx* ctx, const double* X, int m, int n,
                       const double* D, int p, const unsigned char* B,
                       const double* lambda, int n_lambda, double lambda2,
                       const spx_lasso_opts* opts, spx_csc* out);
/* Weighted LARS: replaces _lassoWeighted (cpp_library/spams.h:217-249), i.e.
 * lassoWeight + coreLARS2W (decomp/decomp.h:1315-1391, :1917-2107): signal i
 * minimises 0.5 ||x_i - D a||^2 + lambda sum_j W(j, i) |a_j| (PENALTY; the
 * other modes bound the weighted l1 norm / the residual).  W: host p x n,
 * positive.  Gram ridge 1e-10, stop tolerance 1e-10, at most 4 min(L, p) + 1
 * homotopy steps; opts->max_length_path and opts->ols do not apply (ols and
 * SPX_PREC_REFIT64 are rejected). */
int spx_lasso_weighted_f32(spx_ctx* ctx, const float* X, int m, int n,
                           const float* D, int p, const float* W,
                           const float* lambda, int n_lambda,
                           const spx_lasso_opts* opts, spx_csc* out);
int spx_lasso_weighted_f64(spx_ctx* ctx, const double* X, int m, int n,
                           const double* D, int p, const double* W,
                           const double* lambda, int n_lambda,
                           const spx_lasso_opts* opts, spx_csc* out);

/* ORMP: replaces _omp (cpp_library/spams.h:250-293).  L/eps/lambda follow the
 * facade: pointer NULL or size 0 = not given, size 1 = scalar, size n = per
 * signal.  At least one must be given. */
int spx_omp_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D,
                int p, const int* L, int nL, const float* eps, int n_eps,
                const float* lambda, int n_lambda, const spx_omp_opts* opts,
                spx_csc* out);
int spx_omp_f64(spx_ctx* ctx, const double* X, int m, int n, const double* D,
                int p, const int* L, int nL, const double* eps, int n_eps,
                const double* lambda, int n_lambda, const spx_omp_opts* opts,
                spx_csc* out);
/* As spx_omp_*, with _omp's return_reg_path: `path` (host, p x min(m, p),
 * column-major) receives the first signal's path -- column j holds its
 * coefficients after j+1 selections, the last column written its final
 * coefficients (decomp/decomp.h:605-613, :634-639). */
int spx_omp_path_f32(spx_ctx* ctx, const float* X, int m, int n, const float* D,
                     int p, const int* L, int nL, const float* eps, int n_eps,
                     const float* lambda, int n_lambda, const spx_omp_opts* opts,
                     spx_csc* out, float* path);
int spx_omp_path_f64(spx_ctx* ctx, const double* X, int m, int n,
                     const double* D, int p, const int* L, int nL,
                     const double* eps, int n_eps, const double* lambda,
                     int n_lambda, const spx_omp_opts* opts, spx_csc* out,
                     double* path);

/* ORMP over partly observed signals: replaces _ompMask
 * (cpp_library/spams.h:295-343), i.e. omp_mask (decomp/decomp.h:406-519).
 * B: host m x n bytes, nonzero = row observed.  A signal with every row
 * observed runs the plain path; with none, its column is empty; otherwise
 * ORMP over its observed rows (their Gram columns plus 1e-10 on the
 * diagonal) with eps scaled by the observed fraction.  `path` as in
 * spx_omp_path_* (may be NULL). */
int spx_omp_mask_f32(spx_ctx* ctx, const float* X, int m, int n,
                     const float* D, int p, const unsigned char* B,
                     const int* L, int nL, const float* eps, int n_eps,
                     const float* lambda, int n_lambda,
                     const spx_omp_opts* opts, spx_csc* out, float* path);
int spx_omp_mask_f64(spx_ctx* ctx, const double* X, int m, int n,
                     const double* D, int p, const unsigned char* B,
                     const int* L, int nL, const double* eps, int n_eps,
                     const double* lambda, int n_lambda,
                     const spx_omp_opts* opts, spx_csc* out, double* path);

/* FISTA-flat, square loss + l1 / elastic-net: replaces _fistaFlat
 * (cpp_library/spams.h:460-608).  W (p x n) is written; optim_info is 4 x n
 * (loss, best dual, relative gap, iterations; prox/fista.h:2994-3006). */
int spx_fista_flat_f32(spx_ctx* ctx, const float* X, int m, int n,
                       const float* D, int p, const float* W0, float* W,
                       const spx_fista_opts* opts, float* optim_info);
int spx_fista_flat_f64(spx_ctx* ctx, const double* X, int m, int n,
                       const double* D, int p, const double* W0, double* W,
                       const spx_fista_opts* opts, double* optim_info);

/* Options of spx_train_dl_*: the trainDL arguments of _alltrainDL
 * (cpp_library/spams.h:1153-1190), with the meaning and defaults of
 * swig/python/spams.py:485-500, for modeD = L2 without whitening. */
typedef struct {
  int batchsize;    /* signals per iteration; <= 0 -> 256 * (num_threads + 1) */
  int num_threads;  /* only the batchsize default reads it; <= 0 -> host CPUs  */
  double lambda1;   /* LARS constraint, SPARSITY's atom count, L2ERROR2's eps  */
  double lambda2;   /* ridge max(lambda2, 1e-10) on the Gram diagonal          */
  int iter;         /* iterations (> 0), or seconds of training when < 0      */
  double t0;        /* online: initial model A = |t0| I, B = |t0| D            */
  int mode;         /* SPX_L1COEFFS / SPX_L2ERROR / SPX_PENALTY (LARS),
                     * 3 SPARSITY, 4 L2ERROR2, 5 PENALTY2 (ORMP)              */
  int posAlpha;     /* nonnegative codes (LARS modes)                          */
  int posD;         /* nonnegative dictionary                                  */
  int clean;        /* replace duplicated / vanishing atoms                    */
  double rho;       /* modeParam 1 and 2                                       */
  int iter_updateD; /* dictionary-update sweeps per iteration (1 online, 5 batch
                     * in spams.py)                                            */
  int modeParam;    /* 0 AUTO, 1 PARAM1, 2 PARAM2 (the online weights)         */
  int batch;        /* batch mode: every iteration codes all n signals         */
  int precision;    /* SPX_PREC_BF16X3: f32 products on BF16x3 MFMA            */
  int waves_per_signal; /* solver width (0 = size rule)                        */
  int eager_update; /* 1: launch the dictionary-update kernels one by one
                     * instead of replaying them as a captured hipGraph        */
} spx_train_dl_opts;

/* Dictionary learning: replaces _trainDL / _alltrainDL (cpp_library/spams.h:
 * 1153-1340, Trainer::train dictLearn/dicts.h:346-759).  X: m x n signals.
 * D: m x K, the learned dictionary (out); read as the initial dictionary when
 * D_given (else K signals drawn like the reference's srandom(0)/random()).
 * A (K x K), B (m x K) and *itercount: the model (return_model); a model is
 * given when *itercount > 0 (then A and B are read), and the three are
 * written on return (A and B may be NULL when no model is wanted).  A model
 * needs its dictionary (D_given), as in the facade. */
int spx_train_dl_f32(spx_ctx* ctx, const float* X, int m, int n, float* D,
                     int K, int D_given, float* A, float* B, int* itercount,
                     const spx_train_dl_opts* opts);
int spx_train_dl_f64(spx_ctx* ctx, const double* X, int m, int n, double* D,
                     int K, int D_given, double* A, double* B, int* itercount,
                     const spx_train_dl_opts* opts);

void spx_csc_free(spx_csc* csc);

/* ----------------------------------------------------- device entry points */

/* Per-signal solver outputs on the device, the L x n (value, index) lists of
 * the reference (vM/rM in lasso2/omp, decomp/decomp.h:1578-1583, :337-339):
 * column i holds signal i's active set, index -1 terminated, in active-set
 * order.  `count` (n ints) receives each signal's list length. */
typedef struct {
  void* v;     /* T[L x n] */
  int* r;      /* int[L x n] */
  int* count;  /* int[n] */
  int L;       /* leading dimension of v/r (see spx_lasso_L / spx_omp_L) */
  int* stats;  /* optional int[2 x n]: homotopy steps, atom admissions */
  void* path;  /* optional T[p x path_cols]: signal 0's regularisation path,
                * column `iter` = its coefficients after homotopy step iter
                * (decomp/decomp.h:1868-1871); entries never written are 0 */
  int path_cols;
} spx_lists_dev;

/* Effective L after the facade's clamping (cpp_library/spams.h:117-131 and
 * decomp/decomp.h:333-335). */
int spx_lasso_L(int m, int p, int L, int mode, double lambda1, int pos,
                double lambda2);
int spx_omp_L(int m, int p, const int* L, int nL, int n);

/* Device LARS over a resident dictionary.  X: device m x n (ldx >= m);
 * lambda: device array of n_lambda (1 or n).  Asynchronous on ctx's stream. */
int spx_lasso_dev_f32(spx_ctx* ctx, const spx_dict* dict, const float* X,
                      int ldx, int n, const float* lambda, int n_lambda,
                      float lambda2, const spx_lasso_opts* opts,
                      spx_lists_dev* out);
int spx_lasso_dev_f64(spx_ctx* ctx, const spx_dict* dict, const double* X,
                      int ldx, int n, const double* lambda, int n_lambda,
                      double lambda2, const spx_lasso_opts* opts,
                      spx_lists_dev* out);

/* Device weighted LARS (spx_lasso_weighted_* on a resident dictionary): W is
 * a device p x n weight matrix with leading dimension ldw >= p. */
int spx_lasso_weighted_dev_f32(spx_ctx* ctx, const spx_dict* dict,
                               const float* X, int ldx, int n, const float* W,
                               int ldw, const float* lambda, int n_lambda,
                               const spx_lasso_opts* opts, spx_lists_dev* out);
int spx_lasso_weighted_dev_f64(spx_ctx* ctx, const spx_dict* dict,
                               const double* X, int ldx, int n,
                               const double* W, int ldw, const double* lambda,
                               int n_lambda, const spx_lasso_opts* opts,
                               spx_lists_dev* out);

/* Device masked LARS (spx_lasso_mask_* on a resident dictionary): B is a
 * device m x n byte mask with leading dimension ldb >= m. */
int spx_lasso_mask_dev_f32(spx_ctx* ctx, const spx_dict* dict, const float* X,
                           int ldx, int n, const unsigned char* B, int ldb,
                           const float* lambda, int n_lambda, float lambda2,
                           const spx_lasso_opts* opts, spx_lists_dev* out);
int spx_lasso_mask_dev_f64(spx_ctx* ctx, const spx_dict* dict,
                           const double* X, int ldx, int n,
                           const unsigned char* B, int ldb,
                           const double* lambda, int n_lambda, double lambda2,
                           const spx_lasso_opts* opts, spx_lists_dev* out);

/* Device LARS from precomputed correlations (the Gram-given form of the
 * reference, lasso2(X, G, DtX, ...) decomp/decomp.h:1574-1646 and _lassoQq
 * cpp_library/spams.h:145-188): G is the dictionary's resident Gram, DtX a
 * device p x n matrix (leading dimension ldd >= p, rows p..ldd-1 ignored) and
 * norms the device ||x_i||^2 (n). */
int spx_lasso_dtx_dev_f32(spx_ctx* ctx, const spx_dict* dict, const float* DtX,
                          int ldd, const float* norms, int n,
                          const float* lambda, int n_lambda, float lambda2,
                          const spx_lasso_opts* opts, spx_lists_dev* out);
int spx_lasso_dtx_dev_f64(spx_ctx* ctx, const spx_dict* dict,
                          const double* DtX, int ldd, const double* norms,
                          int n, const double* lambda, int n_lambda,
                          double lambda2, const spx_lasso_opts* opts,
                          spx_lists_dev* out);
/* Allocation-free spx_lasso_dtx_dev_f32.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as spx_lasso_dtx_dev_f32, bit for bit; the
 * caller passes the inputs and outputs, and every piece of solver scratch
 * (work order, overflow list, hand-over states, inverse slots, a ridge Gram
 * for lambda2 > 1e-10) comes from a static 64 MiB device array, one per
 * device, instead of the context's allocations.  A call whose scratch
 * exceeds 64 MiB fails with SPX_ERR_NOMEM and a message naming the need
 * (about 6.6 KB per signal at p = 1024 when L > 24, plus 512 KB per
 * workgroup of the full-capacity pass at L = 256; the pass runs as many
 * workgroups as fit).  No hipMalloc, hipFree, event creation or memset.
 * Every call on a device shares that device's array, so calls that overlap
 * on one device (from any context or thread) race; calls in sequence are
 * safe, and each clears what must start at zero. */
int spx_lasso_dtx_dev_noalloc_f32(spx_ctx* ctx, const spx_dict* dict,
                                  const float* DtX, int ldd,
                                  const float* norms, int n,
                                  const float* lambda, int n_lambda,
                                  float lambda2, const spx_lasso_opts* opts,
                                  spx_lists_dev* out);
/* ||x_i||^2 of the n columns of X (device, leading dimension ldx). */
int spx_sqnorms_dev_f32(spx_ctx* ctx, const float* X, int m, int ldx, int n,
                        float* norms);
int spx_sqnorms_dev_f64(spx_ctx* ctx, const double* X, int m, int ldx, int n,
                        double* norms);

/* Device masked ORMP (spx_omp_mask_* on a resident dictionary): B is a
 * device m x n byte mask with leading dimension ldb >= m; L/eps/lambda as in
 * spx_omp_dev_*. */
int spx_omp_mask_dev_f32(spx_ctx* ctx, const spx_dict* dict, const float* X,
                         int ldx, int n, const unsigned char* B, int ldb,
                         const int* L, int nL, const float* eps, int n_eps,
                         const float* lambda, int n_lambda,
                         const spx_omp_opts* opts, spx_lists_dev* out);
int spx_omp_mask_dev_f64(spx_ctx* ctx, const spx_dict* dict, const double* X,
                         int ldx, int n, const unsigned char* B, int ldb,
                         const int* L, int nL, const double* eps, int n_eps,
                         const double* lambda, int n_lambda,
                         const spx_omp_opts* opts, spx_lists_dev* out);

/* Device ORMP over a resident dictionary.  L/eps/lambda are device arrays
 * (NULL = not given; size 1 or n). */
int spx_omp_dev_f32(spx_ctx* ctx, const spx_dict* dict, const float* X,
                    int ldx, int n, const int* L, int nL, const float* eps,
                    int n_eps, const float* lambda, int n_lambda,
                    const spx_omp_opts* opts, spx_lists_dev* out);
int spx_omp_dev_f64(spx_ctx* ctx, const spx_dict* dict, const double* X,
                    int ldx, int n, const int* L, int nL, const double* eps,
                    int n_eps, const double* lambda, int n_lambda,
                    const spx_omp_opts* opts, spx_lists_dev* out);

/* Device FISTA-flat over a resident dictionary: X m x n, W0/W p x n
 * (device, leading dimension p), optim_info 4 x n (device). */
int spx_fista_flat_dev_f32(spx_ctx* ctx, const spx_dict* dict, const float* X,
                           int n, const float* W0, float* W,
                           const spx_fista_opts* opts, float* optim_info);
int spx_fista_flat_dev_f64(spx_ctx* ctx, const spx_dict* dict,
                           const double* X, int n, const double* W0, double* W,
                           const spx_fista_opts* opts, double* optim_info);

/* Device CSC assembly from solver lists (SpMatrix::convert + per-column sort,
 * linalg/linalg.h:4840-4862): pB (n+1), r and v (capacity >= total count)
 * are device arrays; *nnz receives the total (host value, synchronous). */
int spx_lists_to_csc_dev_f32(spx_ctx* ctx, const spx_lists_dev* lists, int n,
                             int* pB, int* r, float* v, long long* nnz);
int spx_lists_to_csc_dev_f64(spx_ctx* ctx, const spx_lists_dev* lists, int n,
                             int* pB, int* r, double* v, long long* nnz);

/* Products on a resident dictionary (exposed for tests and for callers that
 * need them): DtX = D'X (p x n, leading dimension ldc), norms[i] = ||x_i||^2. */
int spx_dtx_dev_f32(spx_ctx* ctx, const spx_dict* dict, const float* X,
                    int ldx, int n, float* DtX, int ldc, int precision);
int spx_dtx_dev_f64(spx_ctx* ctx, const spx_dict* dict, const double* X,
                    int ldx, int n, double* DtX, int ldc);
/* Largest eigenvalue of the resident Gram matrix (the Lipschitz constant of
 * the square loss's gradient, the fixed FISTA step 1/L) by `iters` steps of
 * the power iteration from the all-ones vector.  Synchronous. */
int spx_dict_lipschitz(spx_ctx* ctx, const spx_dict* dict, int iters, double* L);

/* Copies the resident Gram matrix (p x p, leading dimension p) to `G`
 * (device pointer). */
int spx_dict_gram_f32(spx_ctx* ctx, const spx_dict* dict, float* G);
int spx_dict_gram_f64(spx_ctx* ctx, const spx_dict* dict, double* G);

/* Device dictionary learning: X (device, m x n, leading dimension ldx >= m),
 * D (m x K), A (K x K), B (m x K) device buffers, all read/written as in
 * spx_train_dl_* (A and B required).  Synchronous. */
int spx_train_dl_dev_f32(spx_ctx* ctx, const float* X, int ldx, int m, int n,
                         float* D, int K, int D_given, float* A, float* B,
                         int* itercount, const spx_train_dl_opts* opts);
int spx_train_dl_dev_f64(spx_ctx* ctx, const double* X, int ldx, int m, int n,
                         double* D, int K, int D_given, double* A, double* B,
                         int* itercount, const spx_train_dl_opts* opts);

#ifdef __cplusplus
}
#endif

#endif /* SPX_H */
