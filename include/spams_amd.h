// spams_amd.h -- spams.h-compatible C++ entry points over the spx C ABI.
//
// Drop-in for the decomposition part of SPAMS's C++ facade: the templates
// below keep the reference signatures, argument normalisation, ownership and
// error conventions, and run on the GPU through include/spx.h.
//
//   _lassoD     cpp_library/spams.h:107-143   (lasso2 / coreLARS2 homotopy)
//   _lassoQq    cpp_library/spams.h:145-188   (Gram matrix given)
//   _lassoWeighted  cpp_library/spams.h:217-249  (lassoWeight / coreLARS2W)
//   _lassoMask  cpp_library/spams.h:190-215   (lasso_mask)
//   _ompMask    cpp_library/spams.h:295-343   (omp_mask)
//   _omp        cpp_library/spams.h:250-293   (omp / coreORMP)
//   _fistaFlat  cpp_library/spams.h:460-608   (square loss; l1, elastic-net)
//   lasso2      decomp/decomp.h:126-135        omp  decomp/decomp.h:53-57
//   lasso       decomp/decomp.h:111-123 (the Cholesky variant; D or G/DtX given)
//   cppLasso / cppOMP  cpp_library/cppspams.h:45-51, :106-113
//   _alltrainDL cpp_library/spams.h:1153-1340 (trainDL: Trainer::train,
//               dictLearn/dicts.h:346-759; modeD = L2, no whitening)
//
// Types mirror the reference's layouts (linalg/linalg.h): Matrix<T> is
// column-major with int dimensions and may wrap external memory; Vector<T>
// likewise; SpMatrix<T> is CSC with pB/pE/r/v and rows sorted per column.
// Errors are thrown as `const char*` like the reference facade (its dynamic
// exception specifications are omitted: they are ill-formed since C++17).
#ifndef SPAMS_AMD_H
#define SPAMS_AMD_H

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "spx.h"

// constraint_type (decomp/decomp.h:95)
enum constraint_type { L1COEFFS, L2ERROR, PENALTY, SPARSITY, L2ERROR2, PENALTY2, FISTAMODE };
// constraint_type_D (dictLearn/dicts.h:39)
enum constraint_type_D { L2, L1L2, L1L2FL, L1L2MU };

template <typename T>
class Vector {
 public:
  Vector() {}
  explicit Vector(int n) { resize(n); }
  Vector(T* X, int n) : _X(X), _n(n), _extern(true) {}
  ~Vector() { clear(); }
  Vector(const Vector&) = delete;
  Vector& operator=(const Vector&) = delete;
  void resize(int n) {
    clear();
    _n = n;
    _X = n > 0 ? static_cast<T*>(std::calloc((size_t)n, sizeof(T))) : nullptr;
    _extern = false;
  }
  void clear() {
    if (!_extern) std::free(_X);
    _X = nullptr;
    _n = 0;
  }
  int n() const { return _n; }
  T* rawX() const { return _X; }
  T& operator[](int i) { return _X[i]; }
  T operator[](int i) const { return _X[i]; }

 private:
  T* _X = nullptr;
  int _n = 0;
  bool _extern = false;
};

// AbstractMatrixB (linalg/linalg.h:138): the common base of the dense and
// sparse matrices that _fistaFlat takes as its dictionary.  Here it gives
// the sizes and a dense column-major copy (the engine's products are dense
// MFMA GEMMs, so a sparse dictionary is expanded once per call).
template <typename T>
class AbstractMatrixB {
 public:
  virtual ~AbstractMatrixB() {}
  virtual int m() const = 0;
  virtual int n() const = 0;
  // writes the m() x n() matrix, column-major, to out
  virtual void copyDense(T* out) const = 0;
};

template <typename T>
class Matrix : public AbstractMatrixB<T> {
 public:
  Matrix() {}
  Matrix(int m, int n) { resize(m, n); }
  Matrix(T* X, int m, int n) : _X(X), _m(m), _n(n), _extern(true) {}
  ~Matrix() { clear(); }
  Matrix(const Matrix&) = delete;
  Matrix& operator=(const Matrix&) = delete;
  void resize(int m, int n) {
    clear();
    _m = m;
    _n = n;
    const size_t sz = (size_t)m * n;
    _X = sz ? static_cast<T*>(std::calloc(sz, sizeof(T))) : nullptr;
    _extern = false;
  }
  void clear() {
    if (!_extern) std::free(_X);
    _X = nullptr;
    _m = _n = 0;
  }
  void setZeros() {
    if (_X) std::memset(_X, 0, sizeof(T) * (size_t)_m * _n);
  }
  int m() const override { return _m; }
  int n() const override { return _n; }
  void copyDense(T* out) const override {
    if (_X) std::memcpy(out, _X, sizeof(T) * (size_t)_m * _n);
  }
  T* rawX() const { return _X; }
  T& operator()(int i, int j) { return _X[i + (size_t)j * _m]; }
  T operator()(int i, int j) const { return _X[i + (size_t)j * _m]; }

 private:
  T* _X = nullptr;
  int _m = 0, _n = 0;
  bool _extern = false;
};

// Data<T> stand-in for trainDL's signal argument: dense column-major signals.
template <typename T>
using Data = Matrix<T>;
// AbstractMatrix<T> stand-in for the Gram-given forms' G and DtX: dense.
template <typename T>
using AbstractMatrix = Matrix<T>;

template <typename T>
class SpMatrix : public AbstractMatrixB<T> {
 public:
  SpMatrix() {}
  // wraps caller-owned CSC arrays (linalg/linalg.h:4119-4122): column j has
  // rows r[pB[j] .. pE[j]-1] with values v[...]
  SpMatrix(T* v, int* r, int* pB, int* pE, int m, int n, int nzmax)
      : _pB(pB), _pEx(pE), _r(r), _v(v), _m(m), _n(n), _nzmax(nzmax), _extern(true) {}
  ~SpMatrix() { clear(); }
  SpMatrix(const SpMatrix&) = delete;
  SpMatrix& operator=(const SpMatrix&) = delete;
  void clear() {
    if (!_extern) {
      std::free(_pB);
      std::free(_r);
      std::free(_v);
    }
    _pB = _pEx = _r = nullptr;
    _v = nullptr;
    _m = _n = 0;
    _nzmax = 0;
    _extern = false;
  }
  int m() const override { return _m; }
  int n() const override { return _n; }
  void copyDense(T* out) const override {
    std::memset(out, 0, sizeof(T) * (size_t)_m * _n);
    for (int j = 0; j < _n; ++j)
      for (int k = _pB[j]; k < pE()[j]; ++k) out[_r[k] + (size_t)j * _m] += _v[k];
  }
  int nzmax() const { return (int)_nzmax; }
  int* pB() const { return _pB; }
  int* pE() const { return _pEx ? _pEx : (_pB ? _pB + 1 : nullptr); }
  int* r() const { return _r; }
  T* v() const { return _v; }
  // takes ownership of a library CSC result
  void adopt(spx_csc& c) {
    clear();
    _extern = false;
    _m = c.m;
    _n = c.n;
    _nzmax = c.nnz;
    _pB = c.pB;
    _r = c.r;
    _v = static_cast<T*>(c.v);
    c.pB = c.r = nullptr;
    c.v = nullptr;
  }

 private:
  int* _pB = nullptr;
  int* _pEx = nullptr;  // column ends of wrapped arrays (null: pB + 1)
  int* _r = nullptr;
  T* _v = nullptr;
  int _m = 0, _n = 0;
  long long _nzmax = 0;
  bool _extern = false;
};

namespace spams_amd_detail {

inline spx_ctx* context() {
  static spx_ctx* ctx = nullptr;
  if (!ctx) {
    if (spx_ctx_create(&ctx, 0) != SPX_OK) throw spx_last_error();
  }
  return ctx;
}

inline void check(int rc) {
  if (rc != SPX_OK) throw spx_last_error();
}

inline int lasso_call(const float* X, int m, int n, const float* D, int p, const float* lam,
                      int nlam, float l2, const spx_lasso_opts* o, spx_csc* out,
                      float* path) {
  return spx_lasso_path_f32(context(), X, m, n, D, p, lam, nlam, l2, o, out, path);
}
inline int lasso_call(const double* X, int m, int n, const double* D, int p,
                      const double* lam, int nlam, double l2, const spx_lasso_opts* o,
                      spx_csc* out, double* path) {
  return spx_lasso_path_f64(context(), X, m, n, D, p, lam, nlam, l2, o, out, path);
}
inline int lasso_qq_call(const float* X, int m, int n, const float* Q, const float* q, int p,
                         const float* lam, float l2, const spx_lasso_opts* o, spx_csc* out,
                         float* path) {
  return spx_lasso_qq_f32(context(), X, m, n, Q, q, p, lam, 1, l2, o, out, path);
}
inline int lasso_qq_call(const double* X, int m, int n, const double* Q, const double* q,
                         int p, const double* lam, double l2, const spx_lasso_opts* o,
                         spx_csc* out, double* path) {
  return spx_lasso_qq_f64(context(), X, m, n, Q, q, p, lam, 1, l2, o, out, path);
}
inline int lasso_weighted_call(const float* X, int m, int n, const float* D, int p,
                               const float* W, const float* lam, const spx_lasso_opts* o,
                               spx_csc* out) {
  return spx_lasso_weighted_f32(context(), X, m, n, D, p, W, lam, 1, o, out);
}
inline int lasso_weighted_call(const double* X, int m, int n, const double* D, int p,
                               const double* W, const double* lam, const spx_lasso_opts* o,
                               spx_csc* out) {
  return spx_lasso_weighted_f64(context(), X, m, n, D, p, W, lam, 1, o, out);
}
inline int lasso_mask_call(const float* X, int m, int n, const float* D, int p,
                           const unsigned char* B, const float* lam, float l2,
                           const spx_lasso_opts* o, spx_csc* out) {
  return spx_lasso_mask_f32(context(), X, m, n, D, p, B, lam, 1, l2, o, out);
}
inline int lasso_mask_call(const double* X, int m, int n, const double* D, int p,
                           const unsigned char* B, const double* lam, double l2,
                           const spx_lasso_opts* o, spx_csc* out) {
  return spx_lasso_mask_f64(context(), X, m, n, D, p, B, lam, 1, l2, o, out);
}
inline int omp_mask_call(const float* X, int m, int n, const float* D, int p,
                         const unsigned char* B, const int* L, int nL, const float* e, int ne,
                         const float* la, int nla, spx_csc* out, float* path) {
  return spx_omp_mask_f32(context(), X, m, n, D, p, B, L, nL, e, ne, la, nla, nullptr, out,
                          path);
}
inline int omp_mask_call(const double* X, int m, int n, const double* D, int p,
                         const unsigned char* B, const int* L, int nL, const double* e, int ne,
                         const double* la, int nla, spx_csc* out, double* path) {
  return spx_omp_mask_f64(context(), X, m, n, D, p, B, L, nL, e, ne, la, nla, nullptr, out,
                          path);
}
// Matrix<bool> masks are passed as their bytes (0 / 1)
static_assert(sizeof(bool) == 1, "masks are read as one byte per entry");
inline const unsigned char* mask_bytes(const Matrix<bool>* B) {
  return reinterpret_cast<const unsigned char*>(B->rawX());
}
inline int omp_call(const float* X, int m, int n, const float* D, int p, const int* L, int nL,
                    const float* e, int ne, const float* la, int nla, spx_csc* out,
                    float* path) {
  return spx_omp_path_f32(context(), X, m, n, D, p, L, nL, e, ne, la, nla, nullptr, out, path);
}
inline int omp_call(const double* X, int m, int n, const double* D, int p, const int* L,
                    int nL, const double* e, int ne, const double* la, int nla, spx_csc* out,
                    double* path) {
  return spx_omp_path_f64(context(), X, m, n, D, p, L, nL, e, ne, la, nla, nullptr, out, path);
}

// The path matrix the host entry points fill: p x length, length = the
// facade's normalisation of max_length_path.
template <typename T>
T* path_buffer(Matrix<T>* path, int p, int L, int length_path) {
  if (!path) return nullptr;
  const int cols = length_path < 0 ? 4 * L : length_path;
  if (path->m() != p || path->n() < cols) throw "lasso : the path matrix must be p x max_length_path";
  return path->rawX();
}
inline int train_dl_call(const float* X, int m, int n, float* D, int K, int D_given, float* A,
                         float* B, int* it, const spx_train_dl_opts* o) {
  return spx_train_dl_f32(context(), X, m, n, D, K, D_given, A, B, it, o);
}
inline int train_dl_call(const double* X, int m, int n, double* D, int K, int D_given,
                         double* A, double* B, int* it, const spx_train_dl_opts* o) {
  return spx_train_dl_f64(context(), X, m, n, D, K, D_given, A, B, it, o);
}
inline int fista_call(const float* X, int m, int n, const float* D, int p, const float* W0,
                      float* W, const spx_fista_opts* o, float* info) {
  return spx_fista_flat_f32(context(), X, m, n, D, p, W0, W, o, info);
}
inline int fista_call(const double* X, int m, int n, const double* D, int p,
                      const double* W0, double* W, const spx_fista_opts* o, double* info) {
  return spx_fista_flat_f64(context(), X, m, n, D, p, W0, W, o, info);
}

}  // namespace spams_amd_detail

// lasso2 (decomp/decomp.h:126-135): fills spalpha, L already normalised.
template <typename T>
void lasso2(const Matrix<T>& X, const Matrix<T>& D, SpMatrix<T>& spalpha, int L,
            const T constraint, const T lambda2 = 0, constraint_type mode = PENALTY,
            const bool pos = false, const int numThreads = -1, Matrix<T>* path = nullptr,
            const int length_path = -1) {
  (void)numThreads;
  spx_lasso_opts o{L, (int)mode, pos ? 1 : 0, length_path, SPX_PREC_FP32, SPX_FORM_BATCHED, 0};
  spx_csc out{};
  T* pp = spams_amd_detail::path_buffer(path, D.n(), L, length_path);
  spams_amd_detail::check(spams_amd_detail::lasso_call(X.rawX(), X.m(), X.n(), D.rawX(), D.n(),
                                                       &constraint, 1, lambda2, &o, &out, pp));
  spalpha.adopt(out);
}

// lasso (decomp/decomp.h:111-116): the Cholesky-based variant of the
// reference.  spx runs its inverse-Gram homotopy with coreLARS's step and
// stop rules (decomp/decomp.h:844-1093) and, with ols, refits each final
// active set by least squares.  L already normalised.
template <typename T>
void lasso(const Matrix<T>& X, const Matrix<T>& D, SpMatrix<T>& spalpha, int L,
           const T constraint, const T lambda2 = 0, constraint_type mode = PENALTY,
           const bool pos = false, const bool ols = false, const int numThreads = -1,
           Matrix<T>* path = nullptr, const int length_path = -1) {
  (void)numThreads;
  spx_lasso_opts o{L, (int)mode, pos ? 1 : 0, length_path, SPX_PREC_FP32, SPX_FORM_BATCHED, 0,
                   ols ? 1 : 0, 1};
  spx_csc out{};
  T* pp = spams_amd_detail::path_buffer(path, D.n(), L, length_path);
  spams_amd_detail::check(spams_amd_detail::lasso_call(X.rawX(), X.m(), X.n(), D.rawX(), D.n(),
                                                       &constraint, 1, lambda2, &o, &out, pp));
  spalpha.adopt(out);
}

// lasso2 / lasso with the Gram matrix G (p x p, used as given) and the
// correlations DtX (p x n) given (decomp/decomp.h:118-123, :131-135); X only
// gives the signals' norms.  L already normalised.
template <typename T>
void lasso2(const Data<T>& X, const AbstractMatrix<T>& G, const AbstractMatrix<T>& DtX,
            SpMatrix<T>& spalpha, int L, const T constraint, constraint_type mode = PENALTY,
            const bool pos = false, const int numThreads = -1, Matrix<T>* path = nullptr,
            const int length_path = -1) {
  (void)numThreads;
  if (G.m() != G.n() || DtX.m() != G.n() || DtX.n() != X.n())
    throw "lasso2 : incompatible matrix dimensions";
  spx_lasso_opts o{L, (int)mode, pos ? 1 : 0, length_path, SPX_PREC_FP32, SPX_FORM_BATCHED, 0};
  spx_csc out{};
  T* pp = spams_amd_detail::path_buffer(path, G.n(), L, length_path);
  spams_amd_detail::check(spams_amd_detail::lasso_qq_call(X.rawX(), X.m(), X.n(), G.rawX(),
                                                          DtX.rawX(), G.n(), &constraint, T(0),
                                                          &o, &out, pp));
  spalpha.adopt(out);
}

template <typename T>
void lasso(const Data<T>& X, const AbstractMatrix<T>& G, const AbstractMatrix<T>& DtX,
           SpMatrix<T>& spalpha, int L, const T constraint, constraint_type mode = PENALTY,
           const bool pos = false, const bool ols = false, const int numThreads = -1,
           Matrix<T>* path = nullptr, const int length_path = -1) {
  (void)numThreads;
  if (G.m() != G.n() || DtX.m() != G.n() || DtX.n() != X.n())
    throw "lasso : incompatible matrix dimensions";
  spx_lasso_opts o{L, (int)mode, pos ? 1 : 0, length_path, SPX_PREC_FP32, SPX_FORM_BATCHED, 0,
                   ols ? 1 : 0, 1};
  spx_csc out{};
  T* pp = spams_amd_detail::path_buffer(path, G.n(), L, length_path);
  spams_amd_detail::check(spams_amd_detail::lasso_qq_call(X.rawX(), X.m(), X.n(), G.rawX(),
                                                          DtX.rawX(), G.n(), &constraint, T(0),
                                                          &o, &out, pp));
  spalpha.adopt(out);
}

// _lassoD (cpp_library/spams.h:107-143).  cholevsky selects the step and
// stop rules of the Cholesky variant coreLARS (they differ from coreLARS2's in
// L2ERROR mode); spx runs its single inverse-Gram homotopy for either.  ols
// implies cholevsky (cpp_library/spams.h:136), tracks the least-squares
// residual of the active set and returns the least-squares fit on each final
// active set.
template <typename T>
SpMatrix<T>* _lassoD(Matrix<T>* X, Matrix<T>* D, Matrix<T>** path, bool return_reg_path, int L,
                     const T constraint, const T lambda2, constraint_type mode, const bool pos,
                     const bool ols, const int numThreads, int max_length_path,
                     const bool verbose, bool cholevsky) {
  (void)verbose;
  (void)numThreads;
  if (X->m() != D->m()) throw "lasso : incompatible matrix dimensions";
  if (L < 0) L = D->n();
  if (max_length_path < 0) max_length_path = 4 * L;
  *path = return_reg_path ? new Matrix<T>(D->n(), max_length_path) : nullptr;
  SpMatrix<T>* alpha = new SpMatrix<T>();
  try {
    spx_lasso_opts o{L, (int)mode, pos ? 1 : 0, max_length_path, SPX_PREC_FP32,
                     SPX_FORM_BATCHED, 0, ols ? 1 : 0, cholevsky ? 1 : 0};
    spx_csc out{};
    spams_amd_detail::check(spams_amd_detail::lasso_call(
        X->rawX(), X->m(), X->n(), D->rawX(), D->n(), &constraint, 1, lambda2, &o, &out,
        *path ? (*path)->rawX() : (T*)nullptr));
    alpha->adopt(out);
  } catch (...) {
    delete alpha;
    delete *path;
    *path = nullptr;
    throw;
  }
  return alpha;
}

// _lassoQq (cpp_library/spams.h:145-188): Gram matrix Q (p x p) and
// correlations q (p x n) given; X only gives the signals' norms.
template <typename T>
SpMatrix<T>* _lassoQq(Matrix<T>* X, Matrix<T>* Q, Matrix<T>* q, Matrix<T>** path,
                      bool return_reg_path, int L, const T constraint, const T lambda2,
                      constraint_type mode, const bool pos, const bool ols,
                      const int numThreads, int max_length_path, const bool verbose,
                      bool cholevsky) {
  (void)verbose;
  (void)numThreads;
  if (Q->m() != Q->n()) throw "lasso : Q must be square";
  const int K = Q->m();
  if (q->m() != K || q->n() != X->n()) throw "lasso : incompatible matrix dimensions";
  if (L < 0) L = K;
  if (max_length_path < 0) max_length_path = 4 * L;
  *path = return_reg_path ? new Matrix<T>(K, max_length_path) : nullptr;
  SpMatrix<T>* alpha = new SpMatrix<T>();
  try {
    spx_lasso_opts o{L, (int)mode, pos ? 1 : 0, max_length_path, SPX_PREC_FP32,
                     SPX_FORM_BATCHED, 0, ols ? 1 : 0, cholevsky ? 1 : 0};
    spx_csc out{};
    spams_amd_detail::check(spams_amd_detail::lasso_qq_call(
        X->rawX(), X->m(), X->n(), Q->rawX(), q->rawX(), K, &constraint, lambda2, &o, &out,
        *path ? (*path)->rawX() : (T*)nullptr));
    alpha->adopt(out);
  } catch (...) {
    delete alpha;
    delete *path;
    *path = nullptr;
    throw;
  }
  return alpha;
}

// _lassoWeighted (cpp_library/spams.h:217-249): signal i's penalty is
// constraint * sum_j W(j, i) |alpha_j|, W p x n.
template <typename T>
SpMatrix<T>* _lassoWeighted(Matrix<T>* X, Matrix<T>* D, Matrix<T>* W, int L, const T constraint,
                            constraint_type mode, const bool pos, const int numThreads,
                            bool verbose) {
  (void)verbose;
  (void)numThreads;
  if (X->m() != D->m()) throw "lassoWeighted : incompatible matrix dimensions";
  if (W->m() != D->n() || W->n() != X->n())
    throw "lassoWeighted : inconsistent dimensions of matrix W";
  SpMatrix<T>* alpha = new SpMatrix<T>();
  try {
    spx_lasso_opts o{L, (int)mode, pos ? 1 : 0, -1, SPX_PREC_FP32, SPX_FORM_BATCHED, 0, 0};
    spx_csc out{};
    spams_amd_detail::check(spams_amd_detail::lasso_weighted_call(
        X->rawX(), X->m(), X->n(), D->rawX(), D->n(), W->rawX(), &constraint, &o, &out));
    alpha->adopt(out);
  } catch (...) {
    delete alpha;
    throw;
  }
  return alpha;
}

// _lassoMask (cpp_library/spams.h:190-215): B (m x n) marks the observed
// rows of each signal.  A column with no observed row comes back empty.
template <typename T>
SpMatrix<T>* _lassoMask(Matrix<T>* X, Matrix<T>* D, Matrix<bool>* B, int L, const T constraint,
                        const T lambda2, constraint_type mode, const bool pos,
                        const int numThreads, bool verbose) {
  (void)verbose;
  (void)numThreads;
  if (X->m() != D->m()) throw "lassoMask : incompatible matrix dimensions";
  if (B->m() != X->m() || B->n() != X->n()) throw "lassoMask : Mask has non acceptable dimensions";
  SpMatrix<T>* alpha = new SpMatrix<T>();
  try {
    spx_lasso_opts o{L, (int)mode, pos ? 1 : 0, -1, SPX_PREC_FP32, SPX_FORM_BATCHED, 0, 0};
    spx_csc out{};
    spams_amd_detail::check(spams_amd_detail::lasso_mask_call(
        X->rawX(), X->m(), X->n(), D->rawX(), D->n(), spams_amd_detail::mask_bytes(B),
        &constraint, lambda2, &o, &out));
    alpha->adopt(out);
  } catch (...) {
    delete alpha;
    throw;
  }
  return alpha;
}

// _ompMask (cpp_library/spams.h:295-343)
template <typename T>
SpMatrix<T>* _ompMask(Matrix<T>* X, Matrix<T>* D, Matrix<bool>* B, Matrix<T>** path,
                      bool return_reg_path, bool given_L, Vector<int>* L, bool given_eps,
                      Vector<T>* eps, bool given_Lambda, Vector<T>* Lambda,
                      const int numThreads) {
  (void)numThreads;
  if (X->m() != D->m()) throw "ompMask : incompatible matrix dimensions";
  if (B->m() != X->m() || B->n() != X->n()) throw "ompMask : Mash has non acceptable dimensions";
  if (!given_L && !given_eps && !given_Lambda)
    throw "omp : You should either provide L, eps or lambda";
  const int scalar_L = X->m() < D->n() ? X->m() : D->n();
  *path = return_reg_path ? new Matrix<T>(D->n(), scalar_L) : nullptr;
  SpMatrix<T>* alpha = new SpMatrix<T>();
  try {
    const int n = X->n();
    spx_csc out{};
    spams_amd_detail::check(spams_amd_detail::omp_mask_call(
        X->rawX(), X->m(), n, D->rawX(), D->n(), spams_amd_detail::mask_bytes(B),
        given_L ? L->rawX() : nullptr, given_L ? (L->n() > 1 ? n : 1) : 0,
        given_eps ? eps->rawX() : nullptr, given_eps ? (eps->n() > 1 ? n : 1) : 0,
        given_Lambda ? Lambda->rawX() : nullptr, given_Lambda ? (Lambda->n() > 1 ? n : 1) : 0,
        &out, *path ? (*path)->rawX() : (T*)nullptr));
    alpha->adopt(out);
  } catch (...) {
    delete alpha;
    delete *path;
    *path = nullptr;
    throw;
  }
  return alpha;
}

// omp (decomp/decomp.h:53-57)
template <typename T>
void omp(const Matrix<T>& X, const Matrix<T>& D, SpMatrix<T>& spalpha, const int* L,
         const T* eps, const T* lambda, const bool vecL = false, const bool vecEps = false,
         const bool Lambda = false, const int numThreads = -1, Matrix<T>* path = nullptr) {
  (void)numThreads;
  const int n = X.n();
  const int cols = X.m() < D.n() ? X.m() : D.n();
  if (path && (path->m() != D.n() || path->n() < cols))
    throw "omp : the path matrix must be p x min(m, p)";
  spx_csc out{};
  spams_amd_detail::check(spams_amd_detail::omp_call(
      X.rawX(), X.m(), n, D.rawX(), D.n(), L, L ? (vecL ? n : 1) : 0, eps,
      eps ? (vecEps ? n : 1) : 0, lambda, lambda ? (Lambda ? n : 1) : 0, &out,
      path ? path->rawX() : (T*)nullptr));
  spalpha.adopt(out);
}

// _omp (cpp_library/spams.h:250-293)
template <typename T>
SpMatrix<T>* _omp(Matrix<T>* X, Matrix<T>* D, Matrix<T>** path, bool return_reg_path,
                  bool given_L, Vector<int>* L, bool given_eps, Vector<T>* eps,
                  bool given_Lambda, Vector<T>* Lambda, const int numThreads) {
  if (X->m() != D->m()) throw "omp : incompatible matrix dimensions";
  if (!given_L && !given_eps && !given_Lambda)
    throw "omp : You should either provide L, eps or lambda";
  const int scalar_L = X->m() < D->n() ? X->m() : D->n();
  *path = return_reg_path ? new Matrix<T>(D->n(), scalar_L) : nullptr;
  SpMatrix<T>* alpha = new SpMatrix<T>();
  try {
    omp(*X, *D, *alpha, given_L ? L->rawX() : nullptr, given_eps ? eps->rawX() : nullptr,
        given_Lambda ? Lambda->rawX() : nullptr, given_L && L->n() > 1,
        given_eps && eps->n() > 1, given_Lambda && Lambda->n() > 1, numThreads, *path);
  } catch (...) {
    delete alpha;
    delete *path;
    *path = nullptr;
    throw;
  }
  return alpha;
}

// _fistaFlat (cpp_library/spams.h:460-608), square loss with "l1" or
// "elastic-net".  alpha (p x n) is written in place; returns optim_info 4 x n.
template <typename T>
Matrix<T>* _fistaFlat(Matrix<T>* X, AbstractMatrixB<T>* D, Matrix<T>* alpha0, Matrix<T>* alpha,
                      Vector<int>* groups, int num_threads, int max_it, T L0, bool fixed_step,
                      T gamma, T _lambda, T delta, T lambda2, T lambda3, T a, T b, T c, T tol,
                      int it0, int max_iter_backtracking, bool compute_gram, bool lin_admm,
                      bool admm, bool intercept, bool resetflow, char* name_regul,
                      char* name_loss, bool verbose, bool pos, bool clever, bool log,
                      bool ista, bool subgrad, char* logName, bool is_inner_weights,
                      Vector<T>* inner_weights, int size_group, bool sqrt_step,
                      bool transpose, int linesearch_mode) {
  (void)num_threads; (void)delta; (void)lambda3; (void)a; (void)b; (void)c;
  (void)resetflow; (void)verbose; (void)clever; (void)log;
  (void)logName; (void)inner_weights; (void)size_group; (void)sqrt_step;
  const std::string loss = name_loss ? name_loss : "";
  const std::string reg = name_regul ? name_regul : "";
  if (loss != "square") throw "fistaFlat: Unknown loss";
  int regul;
  if (reg == "l1") regul = 0;
  else if (reg == "elastic-net") regul = 1;
  else throw "fistaFlat: Unknown regularization.";
  if (lin_admm || admm || intercept || subgrad || transpose || is_inner_weights ||
      linesearch_mode != 0 || (groups && groups->n() > 0))
    throw "fistaFlat: option not provided by spx";
  const int m = D->m(), p = D->n(), n = X->n();
  if (alpha0->m() != p || alpha0->n() != n || X->m() != m)
    throw "fistaFlat : Argument sizes are not consistent";
  if (alpha->m() != p || alpha->n() != n) alpha->resize(p, n);
  spx_fista_opts o{max_it, (double)L0, fixed_step ? 1 : 0, (double)gamma, (double)_lambda,
                   (double)lambda2, (double)tol, it0, max_iter_backtracking, regul,
                   pos ? 1 : 0, ista ? 1 : 0, SPX_PREC_FP32, 0, compute_gram ? 1 : 0};
  // a dense dictionary is passed as it is; any other (SpMatrix) is expanded
  const Matrix<T>* Dm = dynamic_cast<const Matrix<T>*>(D);
  std::vector<T> Dd;
  if (!Dm) {
    Dd.resize((size_t)m * p);
    D->copyDense(Dd.data());
  }
  Matrix<T>* info = new Matrix<T>(4, n);
  const int rc = spams_amd_detail::fista_call(X->rawX(), m, n, Dm ? Dm->rawX() : Dd.data(), p,
                                              alpha0->rawX(), alpha->rawX(), &o, info->rawX());
  if (rc != SPX_OK) {
    delete info;
    throw spx_last_error();
  }
  return info;
}

// _alltrainDL (cpp_library/spams.h:1153-1340) for trainDL: online (or batch)
// dictionary learning with LARS (L1COEFFS, L2ERROR, PENALTY) or ORMP
// (SPARSITY, L2ERROR2, PENALTY2) codes, modeD = L2.  Returns the learned
// dictionary (m x K); with return_model, *omA (K x K), *omB (m x K) and
// *omiter (1) receive the model, else NULL.  A model (m_A non-empty) is
// continued from iteration m_iter when D1 is given.  The structured
// (FISTAMODE), in-memory, sparse-dictionary, whitening, stochastic and log
// options are not provided and throw.
template <typename T>
Matrix<T>* _alltrainDL(Data<T>* X, bool in_memory, Matrix<T>** omA, Matrix<T>** omB,
                       Vector<int>** omiter, bool return_model, Matrix<T>* m_A, Matrix<T>* m_B,
                       int m_iter, Matrix<T>* D1, Vector<T>* eta_g, SpMatrix<bool>* groups,
                       SpMatrix<bool>* groups_var, Vector<int>* own_variables,
                       Vector<int>* N_own_variables, int num_threads, T tol, bool fixed_step,
                       bool ista, int batch_size, int K, double lambda1, double lambda2,
                       double lambda3, int iter, double t0, constraint_type mode,
                       char* name_regul, bool posAlpha, bool posD, bool expand,
                       constraint_type_D modeD, bool whiten, bool clean, bool verbose,
                       double gamma1, double gamma2, T rho, int iter_updateD, bool stochastic,
                       int modeParam, bool batch, bool log, char* logName) {
  (void)eta_g; (void)groups; (void)groups_var; (void)own_variables; (void)N_own_variables;
  (void)tol; (void)fixed_step; (void)ista; (void)lambda3; (void)name_regul; (void)verbose;
  (void)gamma1; (void)gamma2; (void)logName;
  if (in_memory || mode == FISTAMODE || modeD != L2 || whiten || expand || stochastic || log)
    throw "trainDL: option not provided by spx (in_memory, structured, modeD, whiten, "
          "expand, stochastic or log)";
  const int n = X->m();
  const int M = X->n();
  const bool D_given = D1 && D1->n() > 0;
  if (!D_given) {
    if (K < 0) throw "trainDL : bad parameter K\n";
  } else {
    K = D1->n();
    if (n != D1->m()) throw "trainDL : sizes of D are not consistent\n";
  }
  const bool model = D_given && m_A && m_A->n() > 0;
  if (model && (m_A->m() != K || m_A->n() != K || !m_B || m_B->m() != n || m_B->n() != K))
    throw "trainDL : sizes of the model are not consistent";
  Matrix<T>* D = new Matrix<T>(n, K);
  if (D_given) std::memcpy(D->rawX(), D1->rawX(), sizeof(T) * (size_t)n * K);
  Matrix<T>* A = new Matrix<T>(K, K);
  Matrix<T>* B = new Matrix<T>(n, K);
  int it = 0;
  if (model) {
    std::memcpy(A->rawX(), m_A->rawX(), sizeof(T) * (size_t)K * K);
    std::memcpy(B->rawX(), m_B->rawX(), sizeof(T) * (size_t)n * K);
    it = m_iter;
  }
  spx_train_dl_opts o{batch_size, num_threads, lambda1, lambda2, iter, t0, (int)mode,
                      posAlpha ? 1 : 0, posD ? 1 : 0, clean ? 1 : 0, (double)rho,
                      iter_updateD, modeParam, batch ? 1 : 0, SPX_PREC_FP32, 0, 0};
  const int rc = spams_amd_detail::train_dl_call(X->rawX(), n, M, D->rawX(), K, D_given ? 1 : 0,
                                                 A->rawX(), B->rawX(), &it, &o);
  if (rc != SPX_OK) {
    delete D;
    delete A;
    delete B;
    throw spx_last_error();
  }
  if (return_model) {
    *omA = A;
    *omB = B;
    *omiter = new Vector<int>(1);
    (**omiter)[0] = it;
  } else {
    delete A;
    delete B;
    if (omA) *omA = nullptr;
    if (omB) *omB = nullptr;
    if (omiter) *omiter = nullptr;
  }
  return D;
}

// cppLasso / cppOMP (cpp_library/cppspams.h:45-51, :106-113).
template <typename T>
SpMatrix<T>* cppLasso(Matrix<T>& X, Matrix<T>& D, Matrix<T>** path, bool return_reg_path,
                      int L, const T constraint, const T lambda2 = 0.,
                      constraint_type mode = PENALTY, const bool pos = false,
                      const bool ols = false, const int numThreads = -1,
                      int max_length_path = -1, const bool verbose = false,
                      bool cholevsky = false) {
  return _lassoD(&X, &D, path, return_reg_path, L, constraint, lambda2, mode, pos, ols,
                 numThreads, max_length_path, verbose, cholevsky);
}

template <typename T>
void cppOMP(const Matrix<T>& X, const Matrix<T>& D, SpMatrix<T>& spalpha, const int* L,
            const T* eps, const T* lambda, const bool vecL = false, const bool vecEps = false,
            const bool vecLambda = false, const int numThreads = -1,
            Matrix<T>* path = nullptr) {
  omp(X, D, spalpha, L, eps, lambda, vecL, vecEps, vecLambda, numThreads, path);
}

#endif  // SPAMS_AMD_H
