/* C restatement of the SPAMS batch decomposition path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and the CPU-baseline
 * leg of bench.py load this library, and only as the checker or as the timed
 * CPU baseline; the engine never links it.
 *
 * Restated reference routines (one per function, file:line in each comment):
 *   orc_gram / orc_dtx     ProdMatrix G and D'X   linalg/linalg.h:1751,1967
 *   orc_corelars2          coreLARS2              decomp/decomp.h:1706-1915
 *   orc_lasso              _lassoD + lasso2       cpp_library/spams.h:107-143,
 *                                                 decomp/decomp.h:1563-1646
 *                          (+ signal 0's path, decomp/decomp.h:1868-1871)
 *   orc_lasso_qq           _lassoQq + lasso2      cpp_library/spams.h:145-188,
 *                                                 decomp/decomp.h:1574-1646
 *   orc_coreormp           coreORMP               decomp/decomp.h:538-640
 *   orc_omp                _omp + omp             cpp_library/spams.h:250-293,
 *                                                 decomp/decomp.h:322-404
 *   orc_fista_flat         FISTA::solver etc.     prox/fista.h:2905-3007,
 *                                                 3530-3573, 3618-3752
 *   orc_convert            SpMatrix::convert      linalg/linalg.h:4840-4862
 *   orc_train_dl           _trainDL + Trainer::train  cpp_library/spams.h:1153-1340,
 *                                                 dictLearn/dicts.h:296-334, 346-759
 *
 * The f32/f64 twins are generated from spx_oracle_body.inc.
 */
#define _GNU_SOURCE
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* constraint_type (decomp/decomp.h:95) */
#define SPX_L1COEFFS 0
#define SPX_L2ERROR 1
#define SPX_PENALTY 2
/* regulariser ids used by this oracle */
#define SPX_REG_L1 0
#define SPX_REG_ELASTICNET 1

#define MAX_D(a, b) ((a) > (b) ? (a) : (b))

/* oracle BLAS (blas_ref.c) */
#define BLAS_DECL(P, T)                                                        \
  T P##dot_(int*, T*, int*, T*, int*);                                         \
  T P##asum_(int*, T*, int*);                                                  \
  void P##axpy_(int*, T*, T*, int*, T*, int*);                                 \
  void P##scal_(int*, T*, T*, int*);                                           \
  void P##gemv_(char*, int*, int*, T*, T*, int*, T*, int*, T*, T*, int*);      \
  void P##symv_(char*, int*, T*, T*, int*, T*, int*, T*, T*, int*);            \
  void P##syr_(char*, int*, T*, T*, int*, T*, int*);                           \
  void P##trmv_(char*, char*, char*, int*, T*, int*, T*, int*);                \
  T P##nrm2_(int*, T*, int*);                                                  \
  void P##syrk_(char*, char*, int*, int*, T*, T*, int*, T*, T*, int*);         \
  int i##P##amax_(int*, T*, int*);
BLAS_DECL(s, float)
BLAS_DECL(d, double)

/* The reference draws from the C library's random() after srandom(0)
 * (dictLearn/dicts.h:400, :459).  The same additive-feedback generator with a
 * private 128-byte state (the default state size) gives the same stream. */
typedef struct {
  struct random_data d;
  char state[128];
} orc_rng;

static void orc_rng_seed(orc_rng* g) {
  memset(g, 0, sizeof *g);
  initstate_r(0, g->state, sizeof g->state, &g->d);
}

static long orc_rng_next(orc_rng* g) {
  int32_t r;
  random_r(&g->d, &r);
  return (long)r;
}

#define T float
#define FN(x) x##_f32
#define BL(x) s##x
#define IAMAX isamax_
#define ABS_T(x) fabsf(x)
#define SQRT_T(x) sqrtf(x)
#define MIN_T(a, b) (((a) > (b)) ? (b) : (a))
#include "spx_oracle_body.inc"
#undef T
#undef FN
#undef BL
#undef IAMAX
#undef ABS_T
#undef SQRT_T

#define T double
#define FN(x) x##_f64
#define BL(x) d##x
#define IAMAX idamax_
#define ABS_T(x) fabs(x)
#define SQRT_T(x) sqrt(x)
#include "spx_oracle_body.inc"
#undef T
#undef FN
#undef BL
#undef IAMAX
#undef ABS_T
#undef SQRT_T
