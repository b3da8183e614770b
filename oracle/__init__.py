"""ctypes front for the oracle libraries (TEST INFRASTRUCTURE ONLY).

Two CPU implementations of the SPAMS decomposition path live here:

* ``liboracle.so`` -- the C restatement (spx_oracle.c), always built;
* ``_ref/libspams_ref.so`` -- the reference's own headers compiled in place
  (ref_shim.cpp), present when the reference tree was available at build time.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
import this package, and only as the checker / the timed CPU baseline.  The
engine never imports or links it.

Every result is returned as ``scipy.sparse.csc_matrix`` (rows sorted within a
column, like SpMatrix::convert, linalg/linalg.h:4840-4862) or dense arrays, in
the dtype of the inputs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libspams_ref.so")

# constraint_type (decomp/decomp.h:95)
L1COEFFS, L2ERROR, PENALTY = 0, 1, 2
REG_L1, REG_ELASTICNET = 0, 1

_P = ctypes.c_void_p
_I = ctypes.c_int
_lib = None
_ref = None


def build(with_ref: bool = True) -> None:
    """Compile the oracle (and the in-place reference build when present)."""
    target = "all" if with_ref else "liboracle.so"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def _sfx(dtype) -> str:
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _ct(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def has_ref() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not has_ref():
            raise FileNotFoundError(REF_PATH + " (reference build absent)")
        _ref = ctypes.CDLL(REF_PATH)
    return _ref


def _f(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _csc_from_lists(vM, rM, L, n, p, dtype):
    """SpMatrix::convert of L x n (value, index) lists via the oracle."""
    f = getattr(lib(), "orc_convert_" + _sfx(dtype))
    f.restype = ctypes.c_long
    f.argtypes = [_P, _P, _I, _I, _P, _P, _P, ctypes.c_long]
    cap = max(1, L * n)
    pB = np.zeros(n + 1, np.int32)
    r = np.zeros(cap, np.int32)
    v = np.zeros(cap, dtype)
    nnz = f(_ptr(vM), _ptr(rM), L, n, _ptr(pB), _ptr(r), _ptr(v), cap)
    assert nnz >= 0
    return sp.csc_matrix((v[:nnz], r[:nnz], pB), shape=(p, n))


# ------------------------------------------------------------------ restatement

def lasso(X, D, lambda1, lambda2=0.0, L=-1, mode=PENALTY, pos=False,
          max_length_path=-1, num_threads=-1, return_stats=False, return_reg_path=False,
          cholesky=False):
    """_lassoD + lasso2 + coreLARS2 (restated).  ``lambda1`` may be a scalar
    or a length-n vector (one constraint per signal).  ``return_reg_path``:
    also the first signal's path (p x max_length_path), NaN where coreLARS2
    never writes (the reference leaves those entries uninitialised).
    ``cholesky``: the path of the Cholesky variant coreLARS (its L2ERROR step
    and stop, its exit on a residual increase)."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    m, n = X.shape
    p = D.shape[1]
    lam = np.atleast_1d(np.asarray(lambda1, dtype=dtype))
    sfx = _sfx(dtype)
    ct = _ct(dtype)
    fL = getattr(lib(), "orc_lasso_L_" + sfx)
    fL.restype = _I
    fL.argtypes = [_I, _I, _I, _I, ct, _I, ct]
    Le = fL(m, p, L, mode, lam[0], int(pos), lambda2)
    if Le <= 0:
        return sp.csc_matrix((p, n), dtype=dtype)
    vM = np.zeros((Le, n), dtype, order="F")
    rM = np.full((Le, n), -1, np.int32, order="F")
    stats = np.zeros((2, n), np.int32, order="F")
    path = None
    if return_reg_path:
        mlp = max_length_path if max_length_path >= 0 else 4 * (p if L < 0 else L)
        path = np.full((p, mlp), np.nan, dtype, order="F")
    f = getattr(lib(), "orc_lasso_" + sfx)
    f.restype = _I
    f.argtypes = [_P, _I, _I, _P, _I, _I, _P, _I, ct, _I, _I, _I, _I, _P, _P, _P, _P, _I]
    f(_ptr(X), m, n, _ptr(D), p, L, _ptr(lam), lam.size, lambda2, mode,
      int(pos), max_length_path, num_threads, _ptr(vM), _ptr(rM), _ptr(stats), _ptr(path),
      int(cholesky))
    out = _csc_from_lists(vM, rM, Le, n, p, dtype)
    res = (out,)
    if return_stats:
        res += (stats,)
    if return_reg_path:
        res += (path,)
    return res[0] if len(res) == 1 else res


def lasso_qq(X, Q, q, lambda1, lambda2=0.0, L=-1, mode=PENALTY, pos=False,
             max_length_path=-1, num_threads=-1, cholesky=False):
    """_lassoQq + lasso2 + coreLARS2 (restated): Gram Q used as given."""
    dtype = X.dtype
    X = _f(X, dtype)
    Q = _f(Q, dtype)
    q = _f(q, dtype)
    m, n = X.shape
    p = Q.shape[1]
    lam = np.atleast_1d(np.asarray(lambda1, dtype=dtype))
    sfx = _sfx(dtype)
    ct = _ct(dtype)
    fL = getattr(lib(), "orc_lasso_L_" + sfx)
    fL.restype = _I
    fL.argtypes = [_I, _I, _I, _I, ct, _I, ct]
    Le = fL(m, p, L, mode, lam[0], int(pos), lambda2)
    if Le <= 0:
        return sp.csc_matrix((p, n), dtype=dtype)
    vM = np.zeros((Le, n), dtype, order="F")
    rM = np.full((Le, n), -1, np.int32, order="F")
    f = getattr(lib(), "orc_lasso_qq_" + sfx)
    f.restype = _I
    f.argtypes = [_P, _I, _I, _P, _P, _I, _I, _P, _I, ct, _I, _I, _I, _I, _P, _P, _P, _I]
    f(_ptr(X), m, n, _ptr(Q), _ptr(q), p, L, _ptr(lam), lam.size, lambda2, mode,
      int(pos), max_length_path, num_threads, _ptr(vM), _ptr(rM), None, int(cholesky))
    return _csc_from_lists(vM, rM, Le, n, p, dtype)


def lasso_weighted(X, D, W, lambda1, L=-1, mode=PENALTY, pos=False, num_threads=-1):
    """_lassoWeighted + lassoWeight + coreLARS2W (restated).  W: p x n
    weights; ``lambda1`` scalar or length-n vector."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    W = _f(W, dtype)
    m, n = X.shape
    p = D.shape[1]
    assert W.shape == (p, n)
    lam = np.atleast_1d(np.asarray(lambda1, dtype=dtype))
    sfx = _sfx(dtype)
    ct = _ct(dtype)
    fL = getattr(lib(), "orc_lasso_L_" + sfx)
    fL.restype = _I
    fL.argtypes = [_I, _I, _I, _I, ct, _I, ct]
    Le = fL(m, p, L, mode, lam[0], int(pos), 0.0)
    if Le <= 0:
        return sp.csc_matrix((p, n), dtype=dtype)
    vM = np.zeros((Le, n), dtype, order="F")
    rM = np.full((Le, n), -1, np.int32, order="F")
    f = getattr(lib(), "orc_lasso_weighted_" + sfx)
    f.restype = _I
    f.argtypes = [_P, _I, _I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _P, _P]
    f(_ptr(X), m, n, _ptr(D), p, _ptr(W), L, _ptr(lam), lam.size, mode, int(pos),
      num_threads, _ptr(vM), _ptr(rM))
    return _csc_from_lists(vM, rM, Le, n, p, dtype)


def lasso_mask(X, D, B, lambda1, lambda2=0.0, L=-1, mode=PENALTY, pos=False,
               num_threads=-1):
    """_lassoMask + lasso_mask + coreLARS2 (restated).  B: m x n mask
    (nonzero = observed).  A column with no observed row is empty."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    Bm = np.asfortranarray(B != 0, dtype=np.uint8)
    m, n = X.shape
    p = D.shape[1]
    sfx = _sfx(dtype)
    ct = _ct(dtype)
    fL = getattr(lib(), "orc_lasso_L_" + sfx)
    fL.restype = _I
    fL.argtypes = [_I, _I, _I, _I, ct, _I, ct]
    Le = fL(m, p, L, mode, lambda1, int(pos), lambda2)
    if Le <= 0:
        return sp.csc_matrix((p, n), dtype=dtype)
    vM = np.zeros((Le, n), dtype, order="F")
    rM = np.full((Le, n), -1, np.int32, order="F")
    f = getattr(lib(), "orc_lasso_mask_" + sfx)
    f.restype = _I
    f.argtypes = [_P, _I, _I, _P, _I, _P, _I, ct, ct, _I, _I, _I, _P, _P]
    f(_ptr(X), m, n, _ptr(D), p, _ptr(Bm), L, lambda1, lambda2, mode, int(pos), num_threads,
      _ptr(vM), _ptr(rM))
    return _csc_from_lists(vM, rM, Le, n, p, dtype)


def omp_mask(X, D, B, L=None, eps=None, lambda1=None, num_threads=-1):
    """_ompMask + omp_mask + coreORMP (restated)."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    Bm = np.asfortranarray(B != 0, dtype=np.uint8)
    m, n = X.shape
    p = D.shape[1]
    sfx = _sfx(dtype)

    def vec(a, dt):
        if a is None:
            return np.zeros(1, dt), 0
        a = np.atleast_1d(np.asarray(a, dtype=dt))
        return a, a.size

    Ls, nL = vec(L, np.int32)
    E, nE = vec(eps, dtype)
    La, nLa = vec(lambda1, dtype)
    fL = getattr(lib(), "orc_omp_L_" + sfx)
    fL.restype = _I
    fL.argtypes = [_I, _I, _P, _I, _I]
    Le = fL(m, p, _ptr(Ls), nL, n)
    if Le <= 0:
        return sp.csc_matrix((p, n), dtype=dtype)
    vM = np.zeros((Le, n), dtype, order="F")
    rM = np.full((Le, n), -1, np.int32, order="F")
    f = getattr(lib(), "orc_omp_mask_" + sfx)
    f.restype = _I
    f.argtypes = [_P, _I, _I, _P, _I, _P, _P, _I, _P, _I, _P, _I, _I, _P, _P, _P]
    f(_ptr(X), m, n, _ptr(D), p, _ptr(Bm), _ptr(Ls), nL, _ptr(E), nE, _ptr(La), nLa,
      num_threads, _ptr(vM), _ptr(rM), None)
    return _csc_from_lists(vM, rM, Le, n, p, dtype)


def omp(X, D, L=None, eps=None, lambda1=None, num_threads=-1):
    """_omp + omp + coreORMP (restated).  L/eps/lambda1: None (not given),
    scalar, or length-n vector."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    m, n = X.shape
    p = D.shape[1]
    sfx = _sfx(dtype)

    def vec(a, dt):
        if a is None:
            return np.zeros(1, dt), 0
        a = np.atleast_1d(np.asarray(a, dtype=dt))
        return a, a.size

    Ls, nL = vec(L, np.int32)
    E, nE = vec(eps, dtype)
    La, nLa = vec(lambda1, dtype)
    fL = getattr(lib(), "orc_omp_L_" + sfx)
    fL.restype = _I
    fL.argtypes = [_I, _I, _P, _I, _I]
    Le = fL(m, p, _ptr(Ls), nL, n)
    if Le <= 0:
        return sp.csc_matrix((p, n), dtype=dtype)
    vM = np.zeros((Le, n), dtype, order="F")
    rM = np.full((Le, n), -1, np.int32, order="F")
    f = getattr(lib(), "orc_omp_" + sfx)
    f.restype = _I
    f.argtypes = [_P, _I, _I, _P, _I, _P, _I, _P, _I, _P, _I, _I, _P, _P]
    f(_ptr(X), m, n, _ptr(D), p, _ptr(Ls), nL, _ptr(E), nE, _ptr(La), nLa,
      num_threads, _ptr(vM), _ptr(rM))
    return _csc_from_lists(vM, rM, Le, n, p, dtype)


def fista_flat(X, D, W0, lambda1, regul="l1", lambda2=0.0, max_it=1000,
               L0=1.0, fixed_step=False, gamma=1.5, tol=1e-6, it0=100,
               max_iter_backtracking=1000, compute_gram=False, pos=False,
               ista=False, num_threads=-1):
    """FISTA_Generic for square loss + l1 / elastic-net (restated).
    Returns (W, optim_info 4 x n)."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    W0 = _f(W0, dtype)
    m, n = X.shape
    p = D.shape[1]
    W = np.zeros((p, n), dtype, order="F")
    info = np.zeros((4, n), dtype, order="F")
    reg = {"l1": REG_L1, "elastic-net": REG_ELASTICNET}[regul]
    ct = _ct(dtype)
    f = getattr(lib(), "orc_fista_flat_" + _sfx(dtype))
    f.restype = _I
    f.argtypes = [_P, _I, _I, _P, _I, _P, _P, _P, _I, _I, ct, _I, ct, ct, ct,
                  ct, _I, _I, _I, _I, _I, _I]
    f(_ptr(X), m, n, _ptr(D), p, _ptr(W0), _ptr(W), _ptr(info), num_threads,
      max_it, L0, int(fixed_step), gamma, lambda1, lambda2, tol, it0,
      max_iter_backtracking, int(compute_gram), reg, int(pos), int(ista))
    return W, info


def gram(D, diag=0.0):
    dtype = D.dtype
    D = _f(D, dtype)
    m, p = D.shape
    G = np.zeros((p, p), dtype, order="F")
    f = getattr(lib(), "orc_gram_" + _sfx(dtype))
    f.argtypes = [_P, _I, _I, _ct(dtype), _P]
    f(_ptr(D), m, p, diag, _ptr(G))
    return G


# ------------------------------------------------------------- reference build

def _ref_sp(h, n, p, dtype, path_shape=None):
    r_ = ref()
    sfx = _sfx(dtype)
    if not h:
        raise RuntimeError("reference call failed")
    fn = getattr(r_, "ref_sp_nnz_" + sfx)
    fn.argtypes = [_P]
    fn.restype = _I
    fnn = getattr(r_, "ref_sp_n_" + sfx)
    fnn.argtypes = [_P]
    fnn.restype = _I
    fc = getattr(r_, "ref_sp_copy_" + sfx)
    fc.argtypes = [_P] * 5
    ff = getattr(r_, "ref_sp_free_" + sfx)
    ff.argtypes = [_P]
    nnz = fn(h)
    nsp = fnn(h)
    pB = np.zeros(max(nsp, n) + 1, np.int32)
    r = np.zeros(max(nnz, 1), np.int32)
    v = np.zeros(max(nnz, 1), dtype)
    path = None
    if path_shape is not None:
        path = np.zeros(path_shape, dtype, order="F")
    fc(h, _ptr(pB), _ptr(r), _ptr(v), _ptr(path))
    ff(h)
    if nsp == 0:
        pB[:] = 0
    nz = pB[n]
    out = sp.csc_matrix((v[:nz], r[:nz], pB[: n + 1]), shape=(p, n))
    return (out, path) if path_shape is not None else out


def ref_lasso(X, D, lambda1, lambda2=0.0, L=-1, mode=PENALTY, pos=False,
              ols=False, cholesky=False, max_length_path=-1, num_threads=-1,
              return_reg_path=False):
    """The reference _lassoD (cpp_library/spams.h:107-143), compiled in place."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    m, n = X.shape
    p = D.shape[1]
    ct = _ct(dtype)
    f = getattr(ref(), "ref_lasso_" + _sfx(dtype))
    f.restype = _P
    f.argtypes = [_P, _I, _I, _P, _I, _I, ct, ct] + [_I] * 7
    h = f(_ptr(X), m, n, _ptr(D), p, L, lambda1, lambda2, mode, int(pos),
          int(ols), int(cholesky), num_threads, max_length_path,
          int(return_reg_path))
    if return_reg_path:
        Lp = p if L < 0 else L
        mlp = 4 * Lp if max_length_path < 0 else max_length_path
        return _ref_sp(h, n, p, dtype, path_shape=(p, mlp))
    return _ref_sp(h, n, p, dtype)


def ref_lasso_qq(X, Q, q, lambda1, lambda2=0.0, L=-1, mode=PENALTY, pos=False,
                 ols=False, cholesky=False, max_length_path=-1, num_threads=-1,
                 return_reg_path=False):
    """The reference _lassoQq (cpp_library/spams.h:145-188), compiled in place."""
    dtype = X.dtype
    X = _f(X, dtype)
    Q = _f(Q, dtype)
    q = _f(q, dtype)
    m, n = X.shape
    p = Q.shape[1]
    ct = _ct(dtype)
    f = getattr(ref(), "ref_lasso_qq_" + _sfx(dtype))
    f.restype = _P
    f.argtypes = [_P, _I, _I, _P, _P, _I, _I, ct, ct] + [_I] * 7
    h = f(_ptr(X), m, n, _ptr(Q), _ptr(q), p, L, lambda1, lambda2, mode, int(pos),
          int(ols), int(cholesky), num_threads, max_length_path, int(return_reg_path))
    if return_reg_path:
        Lp = p if L < 0 else L
        mlp = 4 * Lp if max_length_path < 0 else max_length_path
        return _ref_sp(h, n, p, dtype, path_shape=(p, mlp))
    return _ref_sp(h, n, p, dtype)


def ref_lasso_pervec(X, D, lambdas, lambda2=0.0, L=-1, mode=PENALTY, pos=False,
                     max_length_path=-1, num_threads=-1):
    """lasso2's loop with one constraint per signal over the reference
    coreLARS2 (decomp/decomp.h:1611-1635 with lambdas[i])."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    m, n = X.shape
    p = D.shape[1]
    lam = np.ascontiguousarray(lambdas, dtype=dtype)
    assert lam.size == n
    ct = _ct(dtype)
    f = getattr(ref(), "ref_lasso_pervec_" + _sfx(dtype))
    f.restype = _P
    f.argtypes = [_P, _I, _I, _P, _I, _I, _P, ct, _I, _I, _I, _I]
    h = f(_ptr(X), m, n, _ptr(D), p, L, _ptr(lam), lambda2, mode, int(pos),
          num_threads, max_length_path)
    return _ref_sp(h, n, p, dtype)


def ref_lasso_weighted(X, D, W, lambda1, L=-1, mode=PENALTY, pos=False, num_threads=-1):
    """The reference _lassoWeighted (cpp_library/spams.h:217-249), compiled in
    place."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    W = _f(W, dtype)
    m, n = X.shape
    p = D.shape[1]
    ct = _ct(dtype)
    f = getattr(ref(), "ref_lasso_weighted_" + _sfx(dtype))
    f.restype = _P
    f.argtypes = [_P, _I, _I, _P, _I, _P, _I, ct, _I, _I, _I]
    h = f(_ptr(X), m, n, _ptr(D), p, _ptr(W), L, lambda1, mode, int(pos), num_threads)
    return _ref_sp(h, n, p, dtype)


def ref_lasso_mask(X, D, B, lambda1, lambda2=0.0, L=-1, mode=PENALTY, pos=False,
                   num_threads=-1):
    """The reference _lassoMask (cpp_library/spams.h:190-215), compiled in
    place.  B: m x n mask (nonzero = observed)."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    Bm = np.asfortranarray(B != 0, dtype=np.uint8)
    m, n = X.shape
    p = D.shape[1]
    ct = _ct(dtype)
    f = getattr(ref(), "ref_lasso_mask_" + _sfx(dtype))
    f.restype = _P
    f.argtypes = [_P, _I, _I, _P, _I, _P, _I, ct, ct, _I, _I, _I]
    h = f(_ptr(X), m, n, _ptr(D), p, _ptr(Bm), L, lambda1, lambda2, mode, int(pos),
          num_threads)
    return _ref_sp(h, n, p, dtype)


def ref_omp_mask(X, D, B, L=None, eps=None, lambda1=None, num_threads=-1,
                 return_reg_path=False):
    """The reference _ompMask (cpp_library/spams.h:295-343), compiled in place."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    Bm = np.asfortranarray(B != 0, dtype=np.uint8)
    m, n = X.shape
    p = D.shape[1]

    def vec(a, dt):
        if a is None:
            return np.zeros(1, dt), 0
        a = np.atleast_1d(np.asarray(a, dtype=dt))
        return a, a.size

    Ls, nL = vec(L, np.int32)
    E, nE = vec(eps, dtype)
    La, nLa = vec(lambda1, dtype)
    f = getattr(ref(), "ref_omp_mask_" + _sfx(dtype))
    f.restype = _P
    f.argtypes = [_P, _I, _I, _P, _I, _P, _P, _I, _P, _I, _P, _I, _I, _I]
    h = f(_ptr(X), m, n, _ptr(D), p, _ptr(Bm), _ptr(Ls), nL, _ptr(E), nE, _ptr(La),
          nLa, num_threads, int(return_reg_path))
    if return_reg_path:
        return _ref_sp(h, n, p, dtype, path_shape=(p, min(m, p)))
    return _ref_sp(h, n, p, dtype)


def ref_omp(X, D, L=None, eps=None, lambda1=None, num_threads=-1,
            return_reg_path=False):
    """The reference _omp (cpp_library/spams.h:250-293), compiled in place."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    m, n = X.shape
    p = D.shape[1]

    def vec(a, dt):
        if a is None:
            return np.zeros(1, dt), 0
        a = np.atleast_1d(np.asarray(a, dtype=dt))
        return a, a.size

    Ls, nL = vec(L, np.int32)
    E, nE = vec(eps, dtype)
    La, nLa = vec(lambda1, dtype)
    f = getattr(ref(), "ref_omp_" + _sfx(dtype))
    f.restype = _P
    f.argtypes = [_P, _I, _I, _P, _I, _P, _I, _P, _I, _P, _I, _I, _I]
    h = f(_ptr(X), m, n, _ptr(D), p, _ptr(Ls), nL, _ptr(E), nE, _ptr(La),
          nLa, num_threads, int(return_reg_path))
    if return_reg_path:
        return _ref_sp(h, n, p, dtype, path_shape=(p, min(m, p)))
    return _ref_sp(h, n, p, dtype)


def ref_fista_flat(X, D, W0, lambda1, regul="l1", lambda2=0.0, max_it=1000,
                   L0=1.0, fixed_step=False, gamma=1.5, tol=1e-6, it0=100,
                   max_iter_backtracking=1000, compute_gram=False, pos=False,
                   ista=False, num_threads=-1):
    """The reference _fistaFlat (cpp_library/spams.h:460-608), compiled in
    place.  Returns (W, optim_info 4 x n)."""
    dtype = X.dtype
    X = _f(X, dtype)
    D = _f(D, dtype)
    W0 = _f(W0, dtype)
    m, n = X.shape
    p = D.shape[1]
    W = np.zeros((p, n), dtype, order="F")
    info = np.zeros((4, n), dtype, order="F")
    ct = _ct(dtype)
    f = getattr(ref(), "ref_fista_flat_" + _sfx(dtype))
    f.restype = _I
    f.argtypes = [_P, _I, _I, _P, _I, _P, _P, _P, _I, _I, ct, _I, ct, ct, ct,
                  ct, _I, _I, _I, ctypes.c_char_p, ctypes.c_char_p, _I, _I, _I]
    rc = f(_ptr(X), m, n, _ptr(D), p, _ptr(W0), _ptr(W), _ptr(info),
           num_threads, max_it, L0, int(fixed_step), gamma, lambda1, lambda2,
           tol, it0, max_iter_backtracking, int(compute_gram),
           regul.encode(), b"square", int(pos), 0, int(ista))
    if rc != 0:
        raise RuntimeError("reference fistaFlat failed")
    return W, info


# ------------------------------------------------------- dictionary learning

# constraint_type of trainDL (decomp/decomp.h:95): the LARS modes above plus
# the ORMP modes of Trainer::train (dictLearn/dicts.h:584-591)
SPARSITY, L2ERROR2, PENALTY2 = 3, 4, 5


def _train_dl_call(fn_name, libh, X, D, K, A, B, itercount, batchsize, lambda1,
                   lambda2, iter, t0, mode, posAlpha, posD, clean, rho,
                   iter_updateD, modeParam, batch, num_threads=None):
    dtype = X.dtype
    X = _f(X, dtype)
    m, M = X.shape
    D_given = D is not None
    if D_given:
        K = D.shape[1]
    Dw = _f(D, dtype).copy(order="F") if D_given else np.zeros((m, K), dtype, order="F")
    Aw = _f(A, dtype).copy(order="F") if A is not None else np.zeros((K, K), dtype, order="F")
    Bw = _f(B, dtype).copy(order="F") if B is not None else np.zeros((m, K), dtype, order="F")
    it = ctypes.c_int(int(itercount) if A is not None else 0)
    ct = _ct(dtype)
    f = getattr(libh, fn_name + "_" + _sfx(dtype))
    f.restype = _I
    args = [_ptr(X), m, M, _ptr(Dw), K, int(D_given), _ptr(Aw), _ptr(Bw), ctypes.byref(it)]
    types = [_P, _I, _I, _P, _I, _I, _P, _P, ctypes.POINTER(_I)]
    if num_threads is not None:
        args.append(num_threads)
        types.append(_I)
    args += [batchsize, lambda1, lambda2, iter, t0, mode, int(posAlpha), int(posD),
             int(clean), rho, iter_updateD, modeParam, int(batch)]
    types += [_I, ct, ct, _I, ct, _I, _I, _I, _I, ct, _I, _I, _I]
    f.argtypes = types
    rc = f(*args)
    if rc != 0:
        raise RuntimeError(fn_name + " failed (unsupported options?)")
    return Dw, Aw, Bw, it.value


def train_dl(X, D=None, K=-1, A=None, B=None, itercount=0, batchsize=256,
             lambda1=0.1, lambda2=1e-9, iter=10, t0=1e-5, mode=PENALTY,
             posAlpha=False, posD=False, clean=True, rho=1.0, iter_updateD=1,
             modeParam=0, batch=False):
    """C restatement of trainDL (Trainer::train, dictLearn/dicts.h:346-759),
    one thread.  Returns (D, A, B, itercount)."""
    return _train_dl_call("orc_train_dl", lib(), X, D, K, A, B, itercount, batchsize,
                          lambda1, lambda2, iter, t0, mode, posAlpha, posD, clean, rho,
                          iter_updateD, modeParam, batch)


def ref_train_dl(X, D=None, K=-1, A=None, B=None, itercount=0, batchsize=256,
                 lambda1=0.1, lambda2=1e-9, iter=10, t0=1e-5, mode=PENALTY,
                 posAlpha=False, posD=False, clean=True, rho=1.0, iter_updateD=1,
                 modeParam=0, batch=False, num_threads=1):
    """The reference _alltrainDL (cpp_library/spams.h:1153-1340), compiled in
    place.  Returns (D, A, B, itercount)."""
    return _train_dl_call("ref_train_dl", ref(), X, D, K, A, B, itercount, batchsize,
                          lambda1, lambda2, iter, t0, mode, posAlpha, posD, clean, rho,
                          iter_updateD, modeParam, batch, num_threads=num_threads)
