// C shim over the REFERENCE implementation, compiled in place from the
// reference's own headers (include paths point into the reference tree; no
// reference source is copied here).  Built by oracle/Makefile into
// oracle/_ref/libspams_ref.so.
//
// TEST INFRASTRUCTURE ONLY: used by tests (as the checker), by the golden
// fixture generator and by bench.py's CPU-baseline leg.  Never linked into the
// engine.
//
// Entry points wrap the spams.h facade exactly as a C++ user would call it:
//   _lassoD     cpp_library/spams.h:107-143
//   _omp        cpp_library/spams.h:250-293
//   _fistaFlat  cpp_library/spams.h:460-608
//   _lassoQq    cpp_library/spams.h:145-188
//   _lassoMask / _lassoWeighted   cpp_library/spams.h:190-249
//   _ompMask    cpp_library/spams.h:295-343
//   _alltrainDL cpp_library/spams.h:1153-1340 (trainDL, dictLearn/dicts.h:346-759)
// plus one per-signal-lambda LARS loop (ref_lasso_pervec_*) that mirrors the
// OpenMP loop of lasso2 (decomp/decomp.h:1574-1646) with the constraint taken
// per signal, calling the reference coreLARS2 (decomp/decomp.h:1706-1915).
#include "spams.h"

#include <cstring>
#include <omp.h>

namespace {

template <typename T>
struct SpOut {
  SpMatrix<T>* sp;
  Matrix<T>* path;
};

template <typename T>
void* lasso_impl(const T* X, int m, int n, const T* D, int p, int L, T lambda1,
                 T lambda2, int mode, int pos, int ols, int cholesky,
                 int num_threads, int max_length_path, int return_path) {
  Matrix<T> Xm(const_cast<T*>(X), m, n);
  Matrix<T> Dm(const_cast<T*>(D), m, p);
  Matrix<T>* path = nullptr;
  SpMatrix<T>* sp = nullptr;
  try {
    sp = _lassoD<T>(&Xm, &Dm, &path, return_path != 0, L, lambda1, lambda2,
                    static_cast<constraint_type>(mode), pos != 0, ols != 0,
                    num_threads, max_length_path, false, cholesky != 0);
  } catch (const char* msg) {
    fprintf(stderr, "reference lasso: %s\n", msg);
    return nullptr;
  }
  return new SpOut<T>{sp, path};
}

template <typename T>
void* lasso_qq_impl(const T* X, int m, int n, const T* Q, const T* q, int p, int L,
                    T lambda1, T lambda2, int mode, int pos, int ols, int cholesky,
                    int num_threads, int max_length_path, int return_path) {
  Matrix<T> Xm(const_cast<T*>(X), m, n);
  Matrix<T> Qm(const_cast<T*>(Q), p, p);
  Matrix<T> qm(const_cast<T*>(q), p, n);
  Matrix<T>* path = nullptr;
  SpMatrix<T>* sp = nullptr;
  try {
    sp = _lassoQq<T>(&Xm, &Qm, &qm, &path, return_path != 0, L, lambda1, lambda2,
                     static_cast<constraint_type>(mode), pos != 0, ols != 0, num_threads,
                     max_length_path, false, cholesky != 0);
  } catch (const char* msg) {
    fprintf(stderr, "reference lassoQq: %s\n", msg);
    return nullptr;
  }
  return new SpOut<T>{sp, path};
}

template <typename T>
void* lasso_weighted_impl(const T* X, int m, int n, const T* D, int p, const T* W, int L,
                          T lambda1, int mode, int pos, int num_threads) {
  Matrix<T> Xm(const_cast<T*>(X), m, n);
  Matrix<T> Dm(const_cast<T*>(D), m, p);
  Matrix<T> Wm(const_cast<T*>(W), p, n);
  SpMatrix<T>* sp = nullptr;
  try {
    sp = _lassoWeighted<T>(&Xm, &Dm, &Wm, L, lambda1, static_cast<constraint_type>(mode),
                           pos != 0, num_threads, false);
  } catch (const char* msg) {
    fprintf(stderr, "reference lassoWeighted: %s\n", msg);
    return nullptr;
  }
  return new SpOut<T>{sp, nullptr};
}

// masks arrive as bytes (0 / nonzero), m x n column-major
struct BoolMat {
  bool* b;
  BoolMat(const unsigned char* B, size_t cnt) : b(new bool[cnt ? cnt : 1]) {
    for (size_t k = 0; k < cnt; ++k) b[k] = B[k] != 0;
  }
  ~BoolMat() { delete[] b; }
};

template <typename T>
void* lasso_mask_impl(const T* X, int m, int n, const T* D, int p, const unsigned char* B,
                      int L, T lambda1, T lambda2, int mode, int pos, int num_threads) {
  Matrix<T> Xm(const_cast<T*>(X), m, n);
  Matrix<T> Dm(const_cast<T*>(D), m, p);
  BoolMat bm(B, (size_t)m * n);
  Matrix<bool> Bm(bm.b, m, n);
  SpMatrix<T>* sp = nullptr;
  try {
    sp = _lassoMask<T>(&Xm, &Dm, &Bm, L, lambda1, lambda2, static_cast<constraint_type>(mode),
                       pos != 0, num_threads, false);
  } catch (const char* msg) {
    fprintf(stderr, "reference lassoMask: %s\n", msg);
    return nullptr;
  }
  return new SpOut<T>{sp, nullptr};
}

template <typename T>
void* omp_mask_impl(const T* X, int m, int n, const T* D, int p, const unsigned char* B,
                    const int* L, int nL, const T* eps, int nEps, const T* lambda, int nLambda,
                    int num_threads, int return_path) {
  Matrix<T> Xm(const_cast<T*>(X), m, n);
  Matrix<T> Dm(const_cast<T*>(D), m, p);
  BoolMat bm(B, (size_t)m * n);
  Matrix<bool> Bm(bm.b, m, n);
  Vector<int> vL(const_cast<int*>(L), nL > 0 ? nL : 1);
  Vector<T> vE(const_cast<T*>(eps), nEps > 0 ? nEps : 1);
  Vector<T> vLa(const_cast<T*>(lambda), nLambda > 0 ? nLambda : 1);
  Matrix<T>* path = nullptr;
  SpMatrix<T>* sp = nullptr;
  try {
    sp = _ompMask<T>(&Xm, &Dm, &Bm, &path, return_path != 0, nL > 0, &vL, nEps > 0, &vE,
                     nLambda > 0, &vLa, num_threads);
  } catch (const char* msg) {
    fprintf(stderr, "reference ompMask: %s\n", msg);
    return nullptr;
  }
  return new SpOut<T>{sp, path};
}

template <typename T>
void* omp_impl(const T* X, int m, int n, const T* D, int p, const int* L,
               int nL, const T* eps, int nEps, const T* lambda, int nLambda,
               int num_threads, int return_path) {
  Matrix<T> Xm(const_cast<T*>(X), m, n);
  Matrix<T> Dm(const_cast<T*>(D), m, p);
  // The facade reads the vectors' sizes to choose scalar vs per-signal mode.
  Vector<int> vL(const_cast<int*>(L), nL > 0 ? nL : 1);
  Vector<T> vE(const_cast<T*>(eps), nEps > 0 ? nEps : 1);
  Vector<T> vLa(const_cast<T*>(lambda), nLambda > 0 ? nLambda : 1);
  Matrix<T>* path = nullptr;
  SpMatrix<T>* sp = nullptr;
  try {
    sp = _omp<T>(&Xm, &Dm, &path, return_path != 0, nL > 0, &vL, nEps > 0,
                 &vE, nLambda > 0, &vLa, num_threads);
  } catch (const char* msg) {
    fprintf(stderr, "reference omp: %s\n", msg);
    return nullptr;
  }
  return new SpOut<T>{sp, path};
}

template <typename T>
void* lasso_pervec_impl(const T* X, int m, int n, const T* D, int p, int L,
                        const T* lambdas, T lambda2, int mode, int pos,
                        int num_threads, int max_length_path) {
  // Same argument normalisation as _lassoD (cpp_library/spams.h:117-131).
  if (L < 0) L = p;
  if (max_length_path < 0) max_length_path = 4 * L;
  if (L > m) L = m;
  if (L > p) L = p;
  Matrix<T> Xm(const_cast<T*>(X), m, n);
  Matrix<T> Dm(const_cast<T*>(D), m, p);
  // decomp/decomp.h:1566-1569
  ProdMatrix<T> G(Dm, n > 10 && p < 50000);
  ProdMatrix<T> DtX(Dm, Xm, false);
  G.addDiag(MAX(lambda2, T(1e-10)));
  Matrix<T> vM(L, n);
  Matrix<INTM> rM(L, n);
  const int NT = init_omp(num_threads);
  Vector<T> norms;
  Xm.norm_2sq_cols(norms);
#pragma omp parallel num_threads(NT)
  {
    Vector<T> DtR(p), u(p);
    u.setZeros();
    Matrix<T> Gs(L, L), invGs(L, L), Ga(p, L), work(p, 3);
    work.setZeros();
#pragma omp for schedule(static)
    for (int i = 0; i < n; ++i) {
      T normX = norms[i];
      Vector<INTM> ind;
      rM.refCol(i, ind);
      Vector<T> coeffs;
      vM.refCol(i, coeffs);
      DtX.copyCol(i, DtR);
      coreLARS2(DtR, G, Gs, Ga, invGs, u, coeffs, ind, work, normX,
                static_cast<constraint_type>(mode), lambdas[i], pos != 0,
                static_cast<T*>(nullptr), max_length_path);
    }
  }
  SpMatrix<T>* sp = new SpMatrix<T>();
  sp->convert(vM, rM, p);
  return new SpOut<T>{sp, nullptr};
}

template <typename T>
int sp_nnz(void* h) {
  return static_cast<int>(static_cast<SpOut<T>*>(h)->sp->nzmax());
}

template <typename T>
void sp_copy(void* h, int* pB, int* r, T* v, T* path) {
  SpOut<T>* o = static_cast<SpOut<T>*>(h);
  SpMatrix<T>* sp = o->sp;
  const int n = sp->n();
  if (sp->pB() == nullptr) {  // cleared output (L <= 0): no columns stored
    pB[0] = 0;
    return;
  }
  for (int j = 0; j <= n; ++j) pB[j] = static_cast<int>(sp->pB()[j]);
  const int nnz = static_cast<int>(sp->pB()[n]);
  for (int k = 0; k < nnz; ++k) {
    r[k] = static_cast<int>(sp->r()[k]);
    v[k] = sp->v()[k];
  }
  if (path && o->path)
    std::memcpy(path, o->path->rawX(),
                sizeof(T) * o->path->m() * o->path->n());
}

template <typename T>
void sp_free(void* h) {
  SpOut<T>* o = static_cast<SpOut<T>*>(h);
  delete o->sp;
  delete o->path;
  delete o;
}

template <typename T>
int fista_impl(const T* X, int m, int n, const T* D, int p, const T* W0, T* W,
               T* optim_info, int num_threads, int max_it, T L0,
               int fixed_step, T gamma, T lambda1, T lambda2, T tol, int it0,
               int max_iter_backtracking, int compute_gram, const char* regul,
               const char* loss, int pos, int intercept, int ista) {
  Matrix<T> Xm(const_cast<T*>(X), m, n);
  Matrix<T> Dm(const_cast<T*>(D), m, p);
  Matrix<T> A0(const_cast<T*>(W0), p, n);
  Matrix<T> A(W, p, n);
  Vector<int> groups;  // empty: not given
  Vector<T> inner;     // unused
  char name_regul[64], name_loss[64], log_name[4] = "";
  std::strncpy(name_regul, regul, 63);
  name_regul[63] = 0;
  std::strncpy(name_loss, loss, 63);
  name_loss[63] = 0;
  Matrix<T>* info = nullptr;
  try {
    // Python defaults for the arguments the decomposition path does not use
    // (swig/python/spams.py:260-288).
    info = _fistaFlat<T>(&Xm, &Dm, &A0, &A, &groups, num_threads, max_it, L0,
                         fixed_step != 0, gamma, lambda1, T(1.0), lambda2,
                         T(0.0), T(1.0), T(0.0), T(1.0), tol, it0,
                         max_iter_backtracking, compute_gram != 0, false,
                         false, intercept != 0, false, name_regul, name_loss,
                         false, pos != 0, true, false, ista != 0, false,
                         log_name, false, &inner, 1, true, false, 0);
  } catch (const char* msg) {
    fprintf(stderr, "reference fistaFlat: %s\n", msg);
    return -1;
  }
  std::memcpy(optim_info, info->rawX(), sizeof(T) * 4 * n);
  delete info;
  return 0;
}

// _alltrainDL (cpp_library/spams.h:1153-1340) with the trainDL arguments of
// spams.py (swig/python/spams.py:485-500): no structured regulariser, no log.
// D: m x K in/out (initial dictionary when D_given); A, B, itercount: the
// model in/out (given when itercount_io > 0, like m_A/m_B/m_iter).
template <typename T>
int train_dl_impl(const T* X, int m, int M, T* D, int K, int D_given, T* A, T* B,
                  int* itercount_io, int num_threads, int batchsize, T lambda1,
                  T lambda2, int iter, T t0, int mode, int posAlpha, int posD,
                  int clean, T rho, int iter_updateD, int modeParam, int batch) {
  Matrix<T> Xm(const_cast<T*>(X), m, M);
  Matrix<T> D1;
  if (D_given) D1.setData(D, m, K);
  Matrix<T> mA, mB;
  if (*itercount_io > 0) {
    mA.setData(A, K, K);
    mB.setData(B, m, K);
  }
  Vector<T> eta_g;
  SpMatrix<bool> groups, groups_var;
  Vector<int> own_variables, N_own_variables;
  Matrix<T>* omA = nullptr;
  Matrix<T>* omB = nullptr;
  Vector<int>* omiter = nullptr;
  char regul[] = "none";
  char logname[] = "";
  Matrix<T>* Dout = nullptr;
  try {
    Dout = _alltrainDL<T>(&Xm, false, &omA, &omB, &omiter, true, &mA, &mB,
                          *itercount_io, &D1, &eta_g, &groups, &groups_var,
                          &own_variables, &N_own_variables, num_threads, T(1e-6),
                          true, false, batchsize, K, lambda1, lambda2, 0.0, iter, t0,
                          static_cast<constraint_type>(mode), regul, posAlpha != 0,
                          posD != 0, false, L2, false, clean != 0, false, 0.0, 0.0, rho,
                          iter_updateD, false, modeParam, batch != 0, false, logname);
  } catch (const char* msg) {
    fprintf(stderr, "reference trainDL: %s\n", msg);
    return -1;
  }
  std::memcpy(D, Dout->rawX(), sizeof(T) * (size_t)m * K);
  std::memcpy(A, omA->rawX(), sizeof(T) * (size_t)K * K);
  std::memcpy(B, omB->rawX(), sizeof(T) * (size_t)m * K);
  *itercount_io = (*omiter)[0];
  delete Dout;
  delete omA;
  delete omB;
  delete omiter;
  return 0;
}

}  // namespace

extern "C" {

#define SPX_REF_DEFS(SUF, T)                                                   \
  void* ref_lasso_##SUF(const T* X, int m, int n, const T* D, int p, int L,    \
                        T lambda1, T lambda2, int mode, int pos, int ols,      \
                        int cholesky, int num_threads, int max_length_path,    \
                        int return_path) {                                     \
    return lasso_impl<T>(X, m, n, D, p, L, lambda1, lambda2, mode, pos, ols,   \
                         cholesky, num_threads, max_length_path, return_path); \
  }                                                                            \
  void* ref_lasso_qq_##SUF(const T* X, int m, int n, const T* Q, const T* q,  \
                           int p, int L, T lambda1, T lambda2, int mode,       \
                           int pos, int ols, int cholesky, int num_threads,    \
                           int max_length_path, int return_path) {             \
    return lasso_qq_impl<T>(X, m, n, Q, q, p, L, lambda1, lambda2, mode, pos,  \
                            ols, cholesky, num_threads, max_length_path,       \
                            return_path);                                      \
  }                                                                            \
  void* ref_omp_##SUF(const T* X, int m, int n, const T* D, int p,             \
                      const int* L, int nL, const T* eps, int nEps,            \
                      const T* lambda, int nLambda, int num_threads,           \
                      int return_path) {                                       \
    return omp_impl<T>(X, m, n, D, p, L, nL, eps, nEps, lambda, nLambda,       \
                       num_threads, return_path);                              \
  }                                                                            \
  void* ref_lasso_pervec_##SUF(const T* X, int m, int n, const T* D, int p,    \
                               int L, const T* lambdas, T lambda2, int mode,   \
                               int pos, int num_threads,                       \
                               int max_length_path) {                          \
    return lasso_pervec_impl<T>(X, m, n, D, p, L, lambdas, lambda2, mode, pos, \
                                num_threads, max_length_path);                 \
  }                                                                            \
  void* ref_lasso_weighted_##SUF(const T* X, int m, int n, const T* D, int p, \
                                 const T* W, int L, T lambda1, int mode,       \
                                 int pos, int num_threads) {                   \
    return lasso_weighted_impl<T>(X, m, n, D, p, W, L, lambda1, mode, pos,     \
                                  num_threads);                                \
  }                                                                            \
  void* ref_lasso_mask_##SUF(const T* X, int m, int n, const T* D, int p,      \
                             const unsigned char* B, int L, T lambda1,         \
                             T lambda2, int mode, int pos, int num_threads) {  \
    return lasso_mask_impl<T>(X, m, n, D, p, B, L, lambda1, lambda2, mode,     \
                              pos, num_threads);                               \
  }                                                                            \
  void* ref_omp_mask_##SUF(const T* X, int m, int n, const T* D, int p,        \
                           const unsigned char* B, const int* L, int nL,       \
                           const T* eps, int nEps, const T* lambda,            \
                           int nLambda, int num_threads, int return_path) {    \
    return omp_mask_impl<T>(X, m, n, D, p, B, L, nL, eps, nEps, lambda,        \
                            nLambda, num_threads, return_path);                \
  }                                                                            \
  int ref_train_dl_##SUF(const T* X, int m, int M, T* D, int K, int D_given,  \
                         T* A, T* B, int* itercount_io, int num_threads,       \
                         int batchsize, T lambda1, T lambda2, int iter, T t0,  \
                         int mode, int posAlpha, int posD, int clean, T rho,   \
                         int iter_updateD, int modeParam, int batch) {         \
    return train_dl_impl<T>(X, m, M, D, K, D_given, A, B, itercount_io,        \
                            num_threads, batchsize, lambda1, lambda2, iter,    \
                            t0, mode, posAlpha, posD, clean, rho,              \
                            iter_updateD, modeParam, batch);                   \
  }                                                                            \
  int ref_sp_nnz_##SUF(void* h) { return sp_nnz<T>(h); }                       \
  int ref_sp_n_##SUF(void* h) {                                                \
    return static_cast<int>(static_cast<SpOut<T>*>(h)->sp->n());               \
  }                                                                            \
  void ref_sp_copy_##SUF(void* h, int* pB, int* r, T* v, T* path) {            \
    sp_copy<T>(h, pB, r, v, path);                                             \
  }                                                                            \
  void ref_sp_free_##SUF(void* h) { sp_free<T>(h); }                           \
  int ref_fista_flat_##SUF(const T* X, int m, int n, const T* D, int p,        \
                           const T* W0, T* W, T* optim_info, int num_threads,  \
                           int max_it, T L0, int fixed_step, T gamma,          \
                           T lambda1, T lambda2, T tol, int it0,               \
                           int max_iter_backtracking, int compute_gram,        \
                           const char* regul, const char* loss, int pos,       \
                           int intercept, int ista) {                          \
    return fista_impl<T>(X, m, n, D, p, W0, W, optim_info, num_threads,        \
                         max_it, L0, fixed_step, gamma, lambda1, lambda2, tol, \
                         it0, max_iter_backtracking, compute_gram, regul,      \
                         loss, pos, intercept, ista);                          \
  }

SPX_REF_DEFS(f32, float)
SPX_REF_DEFS(f64, double)

int ref_num_procs(void) { return omp_get_num_procs(); }

}  // extern "C"
