"""Python host mirror of SPAMS's decomposition API over the spx C ABI.

``lasso``, ``omp`` and ``fistaFlat`` keep the argument names, defaults,
return types (scipy.sparse.csc_matrix / Fortran-ordered numpy arrays) and
error behaviour (ValueError / RuntimeError) of the reference's
``swig/python/spams.py`` (:109-142, :167-199, :260-293); the arithmetic runs
in the HIP kernels of ``libspx.so`` (include/spx.h).  There is no CPU
fallback: a missing library or a missing GPU raises.

Device-resident use (inputs already in HBM, e.g. torch tensors) goes through
``Context`` / ``Dictionary`` and the ``*_dev`` functions.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import scipy.sparse as ssp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspx.so")

# constraint_type (decomp/decomp.h:95)
L1COEFFS, L2ERROR, PENALTY, SPARSITY, L2ERROR2, PENALTY2 = 0, 1, 2, 3, 4, 5
# constraint_type_D (dictLearn/dicts.h:39)
L2 = 0
PREC = {"fp32": 0, "bf16x3": 1, "fp32+refit64": 2, "bf16x3+refit64": 3}
FORM = {"batched": 0, "plain": 1}
SOLVERS = {"lasso": 0, "omp": 1}
# launch timers of spx_ctx_read_timing, in SPX_TIMER_* order
TIMERS = ("lars_order", "lars_fast", "lars_full", "ormp", "lars_list", "fista_prox",
          "fista_products", "csc")

_P = ctypes.c_void_p
_I = ctypes.c_int


class spx_csc(ctypes.Structure):
    _fields_ = [("m", _I), ("n", _I), ("nnz", ctypes.c_longlong),
                ("pB", ctypes.POINTER(_I)), ("r", ctypes.POINTER(_I)),
                ("v", _P)]


class spx_lasso_opts(ctypes.Structure):
    _fields_ = [("L", _I), ("mode", _I), ("pos", _I), ("max_length_path", _I),
                ("precision", _I), ("form", _I), ("waves_per_signal", _I), ("ols", _I),
                ("cholesky", _I)]


class spx_omp_opts(ctypes.Structure):
    _fields_ = [("precision", _I), ("form", _I), ("waves_per_signal", _I)]


class spx_fista_opts(ctypes.Structure):
    _fields_ = [("max_it", _I), ("L0", ctypes.c_double), ("fixed_step", _I),
                ("gamma", ctypes.c_double), ("lambda1", ctypes.c_double),
                ("lambda2", ctypes.c_double), ("tol", ctypes.c_double),
                ("it0", _I), ("max_iter_backtracking", _I), ("regul", _I),
                ("pos", _I), ("ista", _I), ("precision", _I), ("graph", _I),
                ("compute_gram", _I)]


class spx_train_dl_opts(ctypes.Structure):
    _fields_ = [("batchsize", _I), ("num_threads", _I), ("lambda1", ctypes.c_double),
                ("lambda2", ctypes.c_double), ("iter", _I), ("t0", ctypes.c_double),
                ("mode", _I), ("posAlpha", _I), ("posD", _I), ("clean", _I),
                ("rho", ctypes.c_double), ("iter_updateD", _I), ("modeParam", _I),
                ("batch", _I), ("precision", _I), ("waves_per_signal", _I),
                ("eager_update", _I)]


class spx_lists_dev(ctypes.Structure):
    _fields_ = [("v", _P), ("r", _P), ("count", _P), ("L", _I), ("stats", _P),
                ("path", _P), ("path_cols", _I)]


_lib = None


def lib():
    """The loaded engine library (raises when it is not built)."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: torch bundles its own libamdhip64.so.7
        # (same soname as ROCm's).  Loading torch first lets the engine bind to
        # that runtime instead of a second copy, which would hide the GPU from
        # torch.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "spx: %s is missing; build it with spams-devel_amd/build.py" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.spx_last_error.restype = ctypes.c_char_p
        L.spx_version.restype = ctypes.c_char_p
        L.spx_dict_bytes.restype = ctypes.c_size_t
        L.spx_dict_bytes.argtypes = [_P]
        L.spx_dict_destroy.argtypes = [_P]
        L.spx_ctx_destroy.argtypes = [_P]
        L.spx_csc_free.argtypes = [ctypes.POINTER(spx_csc)]
        L.spx_lasso_L.argtypes = [_I, _I, _I, _I, ctypes.c_double, _I, ctypes.c_double]
        L.spx_omp_L.argtypes = [_I, _I, _P, _I, _I]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().spx_last_error().decode(errors="replace")
        if rc == 1:
            raise ValueError(msg)
        raise RuntimeError("spx error %d: %s" % (rc, msg))


def _sfx(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f32"
    if dt == np.float64:
        return "f64"
    raise ValueError("spx: only float32 and float64 are supported")


def _ct(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class Context:
    """An engine context on one HIP device (own stream, scratch pool)."""

    def __init__(self, device: int = 0):
        self.h = _P()
        _check(lib().spx_ctx_create(ctypes.byref(self.h), device))
        self.device = device

    def set_stream(self, stream_handle: int | None) -> None:
        """Run on a caller's hipStream_t (0 / None = the HIP null stream)."""
        lib().spx_ctx_set_stream.argtypes = [_P, _P]
        _check(lib().spx_ctx_set_stream(self.h, stream_handle))

    def synchronize(self) -> None:
        lib().spx_ctx_synchronize.argtypes = [_P]
        _check(lib().spx_ctx_synchronize(self.h))

    def set_timing(self, on: bool = True) -> None:
        """Record HIP events around the batched solvers' launches."""
        lib().spx_ctx_set_timing.argtypes = [_P, _I]
        _check(lib().spx_ctx_set_timing(self.h, int(bool(on))))

    def read_timing(self) -> dict:
        """{timer: (total ms, launches)} since the last reset (waits for
        the recorded events)."""
        ms = (ctypes.c_double * len(TIMERS))()
        cnt = (ctypes.c_longlong * len(TIMERS))()
        f = lib().spx_ctx_read_timing
        f.argtypes = [_P, _P, _P, _I]
        _check(f(self.h, ms, cnt, len(TIMERS)))
        return {name: (ms[i], int(cnt[i])) for i, name in enumerate(TIMERS)}

    def scratch_bytes(self) -> int:
        """Device scratch held by the context (bytes)."""
        f = lib().spx_ctx_scratch_bytes
        f.restype = ctypes.c_size_t
        f.argtypes = [_P]
        return int(f(self.h))

    def release_scratch(self) -> None:
        """Waits for the context's stream and frees its device scratch."""
        lib().spx_ctx_release_scratch.argtypes = [_P]
        _check(lib().spx_ctx_release_scratch(self.h))

    def reset_timing(self) -> None:
        lib().spx_ctx_reset_timing.argtypes = [_P]
        _check(lib().spx_ctx_reset_timing(self.h))

    def close(self) -> None:
        if self.h:
            lib().spx_ctx_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Dictionary:
    """A resident dictionary (D, D' and G = D'D in HBM).

    ``D`` is a numpy array (m x p, any order) or a torch CUDA tensor holding
    the column-major m x p matrix, i.e. shaped (p, m) and contiguous."""

    def __init__(self, ctx: Context, D, precision: str = "fp32"):
        self.ctx = ctx
        self.h = _P()
        if hasattr(D, "data_ptr"):
            p, m = D.shape
            dt = np.float32 if str(D.dtype) == "torch.float32" else np.float64
            ptr, on_dev = ctypes.c_void_p(D.data_ptr()), 1
            keep = D
        else:
            keep = np.asfortranarray(D)
            m, p = keep.shape
            dt = keep.dtype
            ptr, on_dev = _ptr(keep), 0
        self.m, self.p, self.dtype = m, p, np.dtype(dt)
        if self.dtype == np.float32:
            f = lib().spx_dict_create_f32
            f.argtypes = [_P, _P, _I, _I, _I, _I, _P]
            _check(f(ctx.h, ptr, m, p, on_dev, PREC[precision], ctypes.byref(self.h)))
        else:
            f = lib().spx_dict_create_f64
            f.argtypes = [_P, _P, _I, _I, _I, _P]
            _check(f(ctx.h, ptr, m, p, on_dev, ctypes.byref(self.h)))
        del keep

    def lipschitz(self, iters: int = 100) -> float:
        """Largest eigenvalue of D'D by the device power iteration."""
        out = ctypes.c_double()
        f = lib().spx_dict_lipschitz
        f.argtypes = [_P, _P, _I, ctypes.POINTER(ctypes.c_double)]
        _check(f(self.ctx.h, self.h, iters, ctypes.byref(out)))
        return out.value

    @property
    def nbytes(self) -> int:
        return int(lib().spx_dict_bytes(self.h))

    def waves(self, solver: str) -> int:
        """Wavefronts per signal that batched calls use ("lasso" | "omp")."""
        w = ctypes.c_int()
        f = lib().spx_dict_get_waves
        f.argtypes = [_P, _I, ctypes.POINTER(ctypes.c_int)]
        _check(f(self.h, SOLVERS[solver], ctypes.byref(w)))
        return w.value

    def set_waves(self, solver: str, waves: int) -> None:
        """Pins the width for a solver (0 = back to the size rule)."""
        f = lib().spx_dict_set_waves
        f.argtypes = [_P, _I, _I]
        _check(f(self.h, SOLVERS[solver], waves))

    def autotune_lasso(self, X, lam, lambda2=0.0, L=-1, mode=PENALTY, pos=False,
                       max_length_path=-1, precision="fp32"):
        """Times the batched LARS solver at 1, 2, 4 and 16 wavefronts per
        signal on a sample batch (X torch (n, m), lam torch (1,) or (n,)),
        keeps the fastest on the dictionary and returns
        (waves, {width: ms})."""
        o = spx_lasso_opts(L, mode, int(pos), max_length_path, PREC[precision],
                           FORM["batched"], 0)
        w = ctypes.c_int()
        ms = (ctypes.c_float * 4)()
        f = getattr(lib(), "spx_dict_autotune_lasso_" + _sfx(self.dtype))
        f.argtypes = [_P, _P, _P, _I, _I, _P, _I, _ct(self.dtype),
                      ctypes.POINTER(spx_lasso_opts), ctypes.POINTER(ctypes.c_int), _P]
        _check(f(self.ctx.h, self.h, _tptr(X), X.shape[1], X.shape[0], _tptr(lam),
                 lam.numel(), lambda2, ctypes.byref(o), ctypes.byref(w), ms))
        return w.value, {k: ms[i] for i, k in enumerate((1, 2, 4, 16)) if ms[i] > 0}

    def autotune_omp(self, X, L=None, eps=None, lambda1=None):
        """As autotune_lasso for the batched ORMP solver."""
        w = ctypes.c_int()
        ms = (ctypes.c_float * 4)()
        nn = lambda t: 0 if t is None else t.numel()
        f = getattr(lib(), "spx_dict_autotune_omp_" + _sfx(self.dtype))
        f.argtypes = [_P, _P, _P, _I, _I, _P, _I, _P, _I, _P, _I,
                      ctypes.POINTER(ctypes.c_int), _P]
        _check(f(self.ctx.h, self.h, _tptr(X), X.shape[1], X.shape[0], _tptr(L), nn(L),
                 _tptr(eps), nn(eps), _tptr(lambda1), nn(lambda1), ctypes.byref(w), ms))
        return w.value, {k: ms[i] for i, k in enumerate((1, 2, 4, 16)) if ms[i] > 0}

    def close(self) -> None:
        if self.h:
            lib().spx_dict_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _csc_to_scipy(c: spx_csc, dtype) -> ssp.csc_matrix:
    n, nnz = c.n, c.nnz
    pB = np.ctypeslib.as_array(c.pB, shape=(n + 1,)).copy()
    if nnz > 0:
        r = np.ctypeslib.as_array(c.r, shape=(nnz,)).copy()
        vp = ctypes.cast(c.v, ctypes.POINTER(_ct(dtype)))
        v = np.ctypeslib.as_array(vp, shape=(nnz,)).copy()
    else:
        r = np.zeros(0, np.int32)
        v = np.zeros(0, dtype)
    lib().spx_csc_free(ctypes.byref(c))
    return ssp.csc_matrix((v, r, pB), shape=(c.m, n))


# --------------------------------------------------------------- host API

def lasso(X, D=None, Q=None, q=None, return_reg_path=False, L=-1,
          lambda1=None, lambda2=0., mode=PENALTY, pos=False, ols=False,
          numThreads=-1, max_length_path=-1, verbose=False, cholesky=False,
          precision="fp32", form="batched", waves_per_signal=0, ctx=None):
    """LARS-Lasso (spams.lasso, swig/python/spams.py:109-142).

    ``lambda1`` is the reference's scalar constraint, or a length-n vector
    (one constraint per signal).  ``cholesky=True`` follows the path of the
    reference's Cholesky variant coreLARS (decomp/decomp.h:844-1093): in
    L2ERROR mode its last step is not cut at the residual target, and a step
    that would increase the residual ends the path (the engine keeps an
    explicit inverse of the active Gram block either way); ``ols=True``
    implies ``cholesky=True`` (as in the reference), tracks the
    least-squares residual of the active set (the L2ERROR stop reads it) and
    returns the least-squares fit on each final active set.  With ``Q``/``q`` (Gram matrix and correlations) the Gram is
    used as given.  ``return_reg_path`` also returns the first signal's path
    (p x max_length_path).  ``numThreads`` is accepted and ignored."""
    if Q is not None:
        if q is None:
            raise ValueError("lasso : q is needed when Q is given")
    elif D is None:
        raise ValueError("lasso : you must give D or Q and q")
    if lambda1 is None:
        raise ValueError("lasso : lambda1 must be defined")
    dtype = np.asarray(X).dtype
    sfx = _sfx(dtype)
    X = np.asfortranarray(X, dtype=dtype)
    m, n = X.shape
    if Q is not None:
        Q = np.asfortranarray(Q, dtype=dtype)
        q = np.asfortranarray(q, dtype=dtype)
        p = Q.shape[1]
        if Q.shape[0] != p:
            raise ValueError("lasso : Q must be square")
        if q.shape != (p, n):
            raise ValueError("lasso : incompatible matrix dimensions")
    else:
        D = np.asfortranarray(D, dtype=dtype)
        if D.shape[0] != m:
            raise ValueError("lasso : incompatible matrix dimensions")
        p = D.shape[1]
    lam = np.ascontiguousarray(np.atleast_1d(np.asarray(lambda1, dtype=dtype)))
    if lam.size not in (1, n):
        raise ValueError("lasso : lambda1 must be a scalar or have one entry per signal")
    o = spx_lasso_opts(L, mode, int(pos), max_length_path, PREC[precision],
                       FORM[form], waves_per_signal, int(bool(ols)),
                       int(bool(cholesky)))
    path = None
    if return_reg_path:
        mlp = max_length_path if max_length_path >= 0 else 4 * (p if L < 0 else L)
        path = np.zeros((p, max(mlp, 0)), dtype=dtype, order="F")
    out = spx_csc()
    ctx = ctx or default_context()
    if Q is not None:
        f = getattr(lib(), "spx_lasso_qq_" + sfx)
        f.argtypes = [_P, _P, _I, _I, _P, _P, _I, _P, _I, _ct(dtype),
                      ctypes.POINTER(spx_lasso_opts), ctypes.POINTER(spx_csc), _P]
        _check(f(ctx.h, _ptr(X), m, n, _ptr(Q), _ptr(q), p, _ptr(lam), lam.size, lambda2,
                 ctypes.byref(o), ctypes.byref(out), _ptr(path)))
    else:
        f = getattr(lib(), "spx_lasso_path_" + sfx)
        f.argtypes = [_P, _P, _I, _I, _P, _I, _P, _I, _ct(dtype),
                      ctypes.POINTER(spx_lasso_opts), ctypes.POINTER(spx_csc), _P]
        _check(f(ctx.h, _ptr(X), m, n, _ptr(D), p, _ptr(lam), lam.size, lambda2,
                 ctypes.byref(o), ctypes.byref(out), _ptr(path)))
    A = _csc_to_scipy(out, dtype)
    return (A, path) if return_reg_path else A


def lassoMask(X, D, B, L=-1, lambda1=None, lambda2=0., mode=PENALTY, pos=False,
              numThreads=-1, verbose=False, precision="fp32", form="batched",
              waves_per_signal=0, ctx=None):
    """LARS-Lasso over partly observed signals (spams.lassoMask,
    swig/python/spams.py:144-153).  B: m x n, nonzero (True) = row observed.
    A column with no observed row comes back empty.  ``numThreads`` is
    accepted and ignored."""
    if lambda1 is None:
        raise ValueError("lassoMask : lambda1 must be defined")
    dtype = np.asarray(X).dtype
    sfx = _sfx(dtype)
    X = np.asfortranarray(X, dtype=dtype)
    D = np.asfortranarray(D, dtype=dtype)
    m, n = X.shape
    if D.shape[0] != m:
        raise ValueError("lassoMask : incompatible matrix dimensions")
    B = np.asfortranarray(np.asarray(B) != 0, dtype=np.uint8)
    if B.shape != (m, n):
        raise ValueError("lassoMask : Mask has non acceptable dimensions")
    p = D.shape[1]
    lam = np.ascontiguousarray(np.atleast_1d(np.asarray(lambda1, dtype=dtype)))
    if lam.size not in (1, n):
        raise ValueError("lassoMask : lambda1 must be a scalar or have one entry per signal")
    o = spx_lasso_opts(L, mode, int(pos), -1, PREC[precision], FORM[form],
                       waves_per_signal, 0)
    out = spx_csc()
    ctx = ctx or default_context()
    f = getattr(lib(), "spx_lasso_mask_" + sfx)
    f.argtypes = [_P, _P, _I, _I, _P, _I, _P, _P, _I, _ct(dtype),
                  ctypes.POINTER(spx_lasso_opts), ctypes.POINTER(spx_csc)]
    _check(f(ctx.h, _ptr(X), m, n, _ptr(D), p, _ptr(B), _ptr(lam), lam.size, lambda2,
             ctypes.byref(o), ctypes.byref(out)))
    return _csc_to_scipy(out, dtype)


def lassoWeighted(X, D, W, L=-1, lambda1=None, mode=PENALTY, pos=False,
                  numThreads=-1, verbose=False, precision="fp32", form="batched",
                  waves_per_signal=0, ctx=None):
    """Weighted LARS-Lasso (spams.lassoWeighted, swig/python/spams.py:155-164):
    signal i's penalty is lambda1 * sum_j W[j, i] |a_j| (W: p x n, positive).
    ``lambda1`` may also be a length-n vector.  ``numThreads`` is accepted and
    ignored."""
    if lambda1 is None:
        raise ValueError("lassoWeighted : lambda1 must be defined")
    dtype = np.asarray(X).dtype
    sfx = _sfx(dtype)
    X = np.asfortranarray(X, dtype=dtype)
    D = np.asfortranarray(D, dtype=dtype)
    W = np.asfortranarray(W, dtype=dtype)
    m, n = X.shape
    if D.shape[0] != m:
        raise ValueError("lassoWeighted : incompatible matrix dimensions")
    p = D.shape[1]
    if W.shape != (p, n):
        raise ValueError("lassoWeighted : inconsistent dimensions of matrix W")
    lam = np.ascontiguousarray(np.atleast_1d(np.asarray(lambda1, dtype=dtype)))
    if lam.size not in (1, n):
        raise ValueError("lassoWeighted : lambda1 must be a scalar or have one entry per signal")
    o = spx_lasso_opts(L, mode, int(pos), -1, PREC[precision], FORM[form],
                       waves_per_signal, 0)
    out = spx_csc()
    ctx = ctx or default_context()
    f = getattr(lib(), "spx_lasso_weighted_" + sfx)
    f.argtypes = [_P, _P, _I, _I, _P, _I, _P, _P, _I,
                  ctypes.POINTER(spx_lasso_opts), ctypes.POINTER(spx_csc)]
    _check(f(ctx.h, _ptr(X), m, n, _ptr(D), p, _ptr(W), _ptr(lam), lam.size,
             ctypes.byref(o), ctypes.byref(out)))
    return _csc_to_scipy(out, dtype)


def omp(X, D, L=None, eps=None, lambda1=None, return_reg_path=False,
        numThreads=-1, precision="fp32", form="batched", waves_per_signal=0,
        ctx=None):
    """Cholesky-based ORMP (spams.omp, swig/python/spams.py:167-199).
    ``return_reg_path`` also returns the first signal's path
    (p x min(m, p))."""
    dtype = np.asarray(X).dtype
    sfx = _sfx(dtype)
    X = np.asfortranarray(X, dtype=dtype)
    D = np.asfortranarray(D, dtype=dtype)
    m, n = X.shape
    if D.shape[0] != m:
        raise ValueError("omp : incompatible matrix dimensions")
    p = D.shape[1]

    def vec(a, dt):
        if a is None:
            return None, 0
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(a, dtype=dt)))
        return a, a.size

    Ls, nL = vec(L, np.int32)
    E, nE = vec(eps, dtype)
    La, nLa = vec(lambda1, dtype)
    o = spx_omp_opts(PREC[precision], FORM[form], waves_per_signal)
    out = spx_csc()
    ctx = ctx or default_context()
    path = np.zeros((p, min(m, p)), dtype=dtype, order="F") if return_reg_path else None
    f = getattr(lib(), "spx_omp_path_" + sfx)
    f.argtypes = [_P, _P, _I, _I, _P, _I, _P, _I, _P, _I, _P, _I,
                  ctypes.POINTER(spx_omp_opts), ctypes.POINTER(spx_csc), _P]
    _check(f(ctx.h, _ptr(X), m, n, _ptr(D), p, _ptr(Ls), nL, _ptr(E), nE,
             _ptr(La), nLa, ctypes.byref(o), ctypes.byref(out), _ptr(path)))
    A = _csc_to_scipy(out, dtype)
    return (A, path) if return_reg_path else A


def ompMask(X, D, B, L=None, eps=None, lambda1=None, return_reg_path=False,
            numThreads=-1, precision="fp32", form="batched", waves_per_signal=0,
            ctx=None):
    """ORMP over partly observed signals (spams.ompMask,
    swig/python/spams.py:201-231).  B: m x n, nonzero (True) = row observed.
    ``return_reg_path`` also returns the first signal's path (p x min(m, p))."""
    dtype = np.asarray(X).dtype
    sfx = _sfx(dtype)
    X = np.asfortranarray(X, dtype=dtype)
    D = np.asfortranarray(D, dtype=dtype)
    m, n = X.shape
    if D.shape[0] != m:
        raise ValueError("ompMask : incompatible matrix dimensions")
    B = np.asfortranarray(np.asarray(B) != 0, dtype=np.uint8)
    if B.shape != (m, n):
        raise ValueError("ompMask : Mash has non acceptable dimensions")
    p = D.shape[1]

    def vec(a, dt):
        if a is None:
            return None, 0
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(a, dtype=dt)))
        return a, a.size

    Ls, nL = vec(L, np.int32)
    E, nE = vec(eps, dtype)
    La, nLa = vec(lambda1, dtype)
    o = spx_omp_opts(PREC[precision], FORM[form], waves_per_signal)
    out = spx_csc()
    ctx = ctx or default_context()
    path = np.zeros((p, min(m, p)), dtype=dtype, order="F") if return_reg_path else None
    f = getattr(lib(), "spx_omp_mask_" + sfx)
    f.argtypes = [_P, _P, _I, _I, _P, _I, _P, _P, _I, _P, _I, _P, _I,
                  ctypes.POINTER(spx_omp_opts), ctypes.POINTER(spx_csc), _P]
    _check(f(ctx.h, _ptr(X), m, n, _ptr(D), p, _ptr(B), _ptr(Ls), nL, _ptr(E), nE,
             _ptr(La), nLa, ctypes.byref(o), ctypes.byref(out), _ptr(path)))
    A = _csc_to_scipy(out, dtype)
    return (A, path) if return_reg_path else A


def fistaFlat(Y, X, W0, return_optim_info=False, numThreads=-1, max_it=1000,
              L0=1.0, fixed_step=False, gamma=1.5, lambda1=1.0, delta=1.0,
              lambda2=0., lambda3=0., a=1.0, b=0., c=1.0, tol=0.000001,
              it0=100, max_iter_backtracking=1000, compute_gram=False,
              lin_admm=False, admm=False, intercept=False, resetflow=False,
              regul="", loss="", verbose=False, pos=False, clever=False,
              log=False, ista=False, subgrad=False, logName="",
              is_inner_weights=False, inner_weights=None, size_group=1,
              groups=None, sqrt_step=True, transpose=False,
              linesearch_mode=0, precision="fp32", ctx=None):
    """FISTA-flat (spams.fistaFlat, swig/python/spams.py:260-293) for
    loss='square' with regul='l1' or 'elastic-net'.  As in the reference, Y
    holds the signals (m x n), X the dictionary (m x p) and W0 the start
    (p x n)."""
    if loss != "square":
        raise ValueError("fistaFlat: spx provides the square loss only (got %r)" % loss)
    regs = {"l1": 0, "elastic-net": 1}
    if regul not in regs:
        raise ValueError("fistaFlat: Unknown regularization %r (spx provides "
                         "'l1' and 'elastic-net')" % regul)
    if intercept or admm or lin_admm or subgrad or linesearch_mode != 0 or groups is not None:
        raise NotImplementedError("fistaFlat: option not provided by spx")
    dtype = np.asarray(Y).dtype
    sfx = _sfx(dtype)
    Yf = np.asfortranarray(Y, dtype=dtype)
    # the binding takes the dictionary dense or sparse (its AbstractMatrixB
    # typemap, swig/spams.i:187); the engine's products are dense
    if hasattr(X, "toarray"):
        X = X.toarray()
    Xf = np.asfortranarray(X, dtype=dtype)
    W0f = np.asfortranarray(W0, dtype=dtype)
    m, n = Yf.shape
    p = Xf.shape[1]
    if W0f.shape != (p, n) or Xf.shape[0] != m:
        raise ValueError("fistaFlat : Argument sizes are not consistent")
    W = np.zeros((p, n), dtype=dtype, order="F")
    info = np.zeros((4, n), dtype=dtype, order="F")
    o = spx_fista_opts(max_it, L0, int(fixed_step), gamma, lambda1, lambda2,
                       tol, it0, max_iter_backtracking, regs[regul], int(pos),
                       int(ista), PREC[precision], 0, int(bool(compute_gram)))
    ctx = ctx or default_context()
    f = getattr(lib(), "spx_fista_flat_" + sfx)
    f.argtypes = [_P, _P, _I, _I, _P, _I, _P, _P,
                  ctypes.POINTER(spx_fista_opts), _P]
    _check(f(ctx.h, _ptr(Yf), m, n, _ptr(Xf), p, _ptr(W0f), _ptr(W),
             ctypes.byref(o), _ptr(info)))
    if return_optim_info:
        return W, info
    return W


def trainDL(X, return_model=False, model=None, D=None, numThreads=-1, batchsize=-1,
            K=-1, lambda1=None, lambda2=10e-10, iter=-1, t0=1e-5, mode=PENALTY,
            posAlpha=False, posD=False, expand=False, modeD=L2, whiten=False,
            clean=True, verbose=True, gamma1=0., gamma2=0., rho=1.0,
            iter_updateD=None, stochastic_deprecated=False, modeParam=0,
            batch=False, log_deprecated=False, logName='', precision="fp32",
            waves_per_signal=0, eager_update=False, ctx=None):
    """Dictionary learning (spams.trainDL, swig/python/spams.py:485-500 and
    :412-484) for modeD=L2 without whitening.  Returns D (m x K, Fortran
    order), or (D, model) with model = {'A', 'B', 'iter'} when return_model.
    A model is used (warm restart) when D is given too, as in
    cpp_library/spams.h:1201-1210.  `iter` < 0 trains for -iter seconds."""
    if lambda1 is None:
        raise ValueError("trainDL : lambda1 must be defined")
    if modeD != L2 or whiten or expand or stochastic_deprecated or log_deprecated:
        raise NotImplementedError("trainDL: option not provided by spx (modeD=L2, no "
                                  "whitening / expand / stochastic / log)")
    if iter_updateD is None:
        iter_updateD = 5 if batch else 1
    dtype = np.asarray(X).dtype
    sfx = _sfx(dtype)
    Xf = np.asfortranarray(X, dtype=dtype)
    m, n = Xf.shape
    D_given = D is not None and np.asarray(D).size > 0
    if D_given:
        Dw = np.array(D, dtype=dtype, order="F", copy=True)
        if Dw.shape[0] != m:
            raise RuntimeError("trainDL : sizes of D are not consistent\n")
        K = Dw.shape[1]
    else:
        if K < 0:
            raise RuntimeError("trainDL : bad parameter K\n")
        Dw = np.zeros((m, K), dtype=dtype, order="F")
    it = ctypes.c_int(0)
    A = np.zeros((K, K), dtype=dtype, order="F")
    B = np.zeros((m, K), dtype=dtype, order="F")
    if model is not None and D_given:
        A = np.array(model['A'], dtype=dtype, order="F", copy=True)
        B = np.array(model['B'], dtype=dtype, order="F", copy=True)
        it = ctypes.c_int(int(model['iter']))
    o = spx_train_dl_opts(int(batchsize), int(numThreads), float(lambda1), float(lambda2),
                          int(iter), float(t0), int(mode), int(posAlpha), int(posD),
                          int(clean), float(rho), int(iter_updateD), int(modeParam),
                          int(batch), PREC[precision], int(waves_per_signal),
                          int(eager_update))
    ctx = ctx or default_context()
    f = getattr(lib(), "spx_train_dl_" + sfx)
    f.argtypes = [_P, _P, _I, _I, _P, _I, _I, _P, _P, ctypes.POINTER(_I),
                  ctypes.POINTER(spx_train_dl_opts)]
    _check(f(ctx.h, _ptr(Xf), m, n, _ptr(Dw), K, int(D_given), _ptr(A), _ptr(B),
             ctypes.byref(it), ctypes.byref(o)))
    if return_model:
        return Dw, {'A': A, 'B': B, 'iter': it.value}
    return Dw


def nmf(X, return_lasso=False, model=None, numThreads=-1, batchsize=-1, K=-1, iter=-1,
        t0=1e-5, clean=True, rho=1.0, modeParam=0, batch=False, ctx=None):
    """Non-negative matrix factorisation (spams.nmf, swig/python/spams.py:
    530-548): trainDL with lambda1 = 0 and nonnegative codes and dictionary;
    with return_lasso, also the nonnegative codes V = lasso(X, U, 0, pos)."""
    lambda1 = 0
    U = trainDL(X, model=model, numThreads=numThreads, batchsize=batchsize, K=K, iter=iter,
                t0=t0, clean=clean, rho=rho, verbose=False, modeParam=modeParam, batch=batch,
                lambda1=lambda1, mode=PENALTY, posAlpha=True, posD=True, whiten=False, ctx=ctx)
    if not return_lasso:
        return U
    if ssp.issparse(X):
        raise ValueError("sparse matrix for lasso not yet implemented")
    V = lasso(X, D=U, return_reg_path=False, numThreads=numThreads, lambda1=lambda1,
              mode=PENALTY, pos=True, ctx=ctx)
    return (U, V)


def nnsc(X, return_lasso=False, model=None, lambda1=None, numThreads=-1, batchsize=-1,
         K=-1, iter=-1, t0=1e-5, clean=True, rho=1.0, modeParam=0, batch=False, ctx=None):
    """Non-negative sparse coding (spams.nnsc, swig/python/spams.py:550-562):
    trainDL with nonnegative codes and dictionary; with return_lasso, also
    V = lasso(X, U, lambda1) (without the positivity constraint, as the
    reference calls it)."""
    if lambda1 is None:
        raise ValueError("nnsc : lambda1 must be defined")
    U = trainDL(X, model=model, numThreads=numThreads, batchsize=batchsize, K=K, iter=iter,
                t0=t0, clean=clean, rho=rho, verbose=False, modeParam=modeParam, batch=batch,
                lambda1=lambda1, mode=PENALTY, posAlpha=True, posD=True, whiten=False, ctx=ctx)
    if not return_lasso:
        return U
    V = lasso(X, D=U, return_reg_path=False, numThreads=numThreads, lambda1=lambda1,
              mode=PENALTY, ctx=ctx)
    return (U, V)


def train_dl_dev(ctx, X, D, A, B, itercount=0, D_given=True, **opts):
    """Device dictionary learning (spx_train_dl_dev_*) on torch tensors:
    X (m x n), D (m x K), A (K x K), B (m x K), column-major as the
    transposed views of row-major tensors (pass X.T.contiguous().T etc. or
    tensors created as (n, m) and viewed transposed).  Returns the new
    iteration count; D, A, B are updated in place."""
    dtype = np.float32 if X.element_size() == 4 else np.float64
    sfx = _sfx(dtype)
    m, n = X.shape
    K = D.shape[1]
    for name, t, shape, ld in (("D", D, (m, K), m), ("A", A, (K, K), K), ("B", B, (m, K), m)):
        if tuple(t.shape) != shape or t.stride(0) != 1 or t.stride(1) != ld or t.dtype != X.dtype:
            raise ValueError("train_dl_dev: %s must be a column-major %dx%d tensor of X's dtype"
                             % (name, shape[0], shape[1]))
    if X.stride(0) != 1 or X.stride(1) < m:
        raise ValueError("train_dl_dev: X must be column-major (a transposed row-major tensor)")
    o = spx_train_dl_opts(opts.get("batchsize", -1), opts.get("numThreads", -1),
                          float(opts["lambda1"]), float(opts.get("lambda2", 10e-10)),
                          int(opts.get("iter", 1)), float(opts.get("t0", 1e-5)),
                          int(opts.get("mode", PENALTY)), int(opts.get("posAlpha", False)),
                          int(opts.get("posD", False)), int(opts.get("clean", True)),
                          float(opts.get("rho", 1.0)), int(opts.get("iter_updateD", 1)),
                          int(opts.get("modeParam", 0)), int(opts.get("batch", False)),
                          PREC[opts.get("precision", "fp32")],
                          int(opts.get("waves_per_signal", 0)),
                          int(opts.get("eager_update", False)))
    it = ctypes.c_int(int(itercount))
    f = getattr(lib(), "spx_train_dl_dev_" + sfx)
    f.argtypes = [_P, _P, _I, _I, _I, _P, _I, _I, _P, _P, ctypes.POINTER(_I),
                  ctypes.POINTER(spx_train_dl_opts)]
    _check(f(ctx.h, _tptr(X), X.stride(1), m, n, _tptr(D), K, int(D_given), _tptr(A),
             _tptr(B), ctypes.byref(it), ctypes.byref(o)))
    return it.value


# ------------------------------------------------------------- device API

def lasso_L(m, p, L=-1, mode=PENALTY, lambda1=1.0, pos=False, lambda2=0.0):
    return lib().spx_lasso_L(m, p, L, mode, float(lambda1), int(pos), float(lambda2))


def omp_L(m, p, L=None, n=1):
    if L is None:
        return lib().spx_omp_L(m, p, None, 0, n)
    a = np.ascontiguousarray(np.atleast_1d(np.asarray(L, np.int32)))
    return lib().spx_omp_L(m, p, _ptr(a), a.size, n)


def _tptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def lasso_dev(ctx, d: Dictionary, X, lam, V, R, count, lambda2=0.0, L=-1,
              mode=PENALTY, pos=False, max_length_path=-1, precision="fp32",
              form="batched", waves_per_signal=0, stats=None):
    """Device LARS: X torch (n, m) (column-major m x n), lam torch (1,) or
    (n,), V/R torch (n, Lout) outputs, count torch (n,) int32."""
    n = X.shape[0]
    Lout = V.shape[1]
    sfx = "f32" if d.dtype == np.float32 else "f64"
    o = spx_lasso_opts(L, mode, int(pos), max_length_path, PREC[precision],
                       FORM[form], waves_per_signal)
    lists = spx_lists_dev(_tptr(V), _tptr(R), _tptr(count), Lout, _tptr(stats))
    f = getattr(lib(), "spx_lasso_dev_" + sfx)
    f.argtypes = [_P, _P, _P, _I, _I, _P, _I, _ct(d.dtype),
                  ctypes.POINTER(spx_lasso_opts), ctypes.POINTER(spx_lists_dev)]
    _check(f(ctx.h, d.h, _tptr(X), X.shape[1], n, _tptr(lam), lam.numel(),
             lambda2, ctypes.byref(o), ctypes.byref(lists)))


def lasso_weighted_dev(ctx, d: Dictionary, X, W, lam, V, R, count, L=-1,
                       mode=PENALTY, pos=False, precision="fp32", form="batched",
                       waves_per_signal=0, stats=None):
    """Device weighted LARS: X torch (n, m), W torch (n, p) (column-major
    p x n weights), other arguments as lasso_dev."""
    n = X.shape[0]
    Lout = V.shape[1]
    sfx = "f32" if d.dtype == np.float32 else "f64"
    o = spx_lasso_opts(L, mode, int(pos), -1, PREC[precision], FORM[form],
                       waves_per_signal)
    lists = spx_lists_dev(_tptr(V), _tptr(R), _tptr(count), Lout, _tptr(stats))
    f = getattr(lib(), "spx_lasso_weighted_dev_" + sfx)
    f.argtypes = [_P, _P, _P, _I, _I, _P, _I, _P, _I,
                  ctypes.POINTER(spx_lasso_opts), ctypes.POINTER(spx_lists_dev)]
    _check(f(ctx.h, d.h, _tptr(X), X.shape[1], n, _tptr(W), W.stride(0), _tptr(lam),
             lam.numel(), ctypes.byref(o), ctypes.byref(lists)))


def lasso_mask_dev(ctx, d: Dictionary, X, B, lam, V, R, count, lambda2=0.0, L=-1,
                   mode=PENALTY, pos=False, precision="fp32", form="batched",
                   waves_per_signal=0, stats=None):
    """Device masked LARS: X torch (n, m), B torch uint8 (n, m) (column-major
    m x n mask, nonzero = observed), other arguments as lasso_dev."""
    n = X.shape[0]
    Lout = V.shape[1]
    sfx = "f32" if d.dtype == np.float32 else "f64"
    o = spx_lasso_opts(L, mode, int(pos), -1, PREC[precision], FORM[form],
                       waves_per_signal)
    lists = spx_lists_dev(_tptr(V), _tptr(R), _tptr(count), Lout, _tptr(stats))
    f = getattr(lib(), "spx_lasso_mask_dev_" + sfx)
    f.argtypes = [_P, _P, _P, _I, _I, _P, _I, _P, _I, _ct(d.dtype),
                  ctypes.POINTER(spx_lasso_opts), ctypes.POINTER(spx_lists_dev)]
    _check(f(ctx.h, d.h, _tptr(X), X.stride(0), n, _tptr(B), B.stride(0), _tptr(lam),
             lam.numel(), lambda2, ctypes.byref(o), ctypes.byref(lists)))


def lasso_dtx_dev(ctx, d: Dictionary, DtX, norms, lam, V, R, count, lambda2=0.0,
                  L=-1, mode=PENALTY, pos=False, max_length_path=-1,
                  form="batched", waves_per_signal=0, stats=None, noalloc=False):
    """Device LARS from precomputed D'X (torch (n, ldd)) and ||x||^2 (n,).
    ``noalloc`` (FP32 only) selects the allocation-free entry point, whose
    scratch is a static 64 MiB device array (it raises when that is too
    small for the call)."""
    n = DtX.shape[0]
    Lout = V.shape[1]
    sfx = "f32" if d.dtype == np.float32 else "f64"
    if noalloc and sfx != "f32":
        raise ValueError("lasso_dtx_dev: the allocation-free entry point is FP32")
    o = spx_lasso_opts(L, mode, int(pos), max_length_path, 0, FORM[form],
                       waves_per_signal)
    lists = spx_lists_dev(_tptr(V), _tptr(R), _tptr(count), Lout, _tptr(stats))
    f = getattr(lib(), ("spx_lasso_dtx_dev_noalloc_" if noalloc else "spx_lasso_dtx_dev_") + sfx)
    f.argtypes = [_P, _P, _P, _I, _P, _I, _P, _I, _ct(d.dtype),
                  ctypes.POINTER(spx_lasso_opts), ctypes.POINTER(spx_lists_dev)]
    _check(f(ctx.h, d.h, _tptr(DtX), DtX.shape[1], _tptr(norms), n, _tptr(lam),
             lam.numel(), lambda2, ctypes.byref(o), ctypes.byref(lists)))


def sqnorms_dev(ctx, X, out):
    """out (n,) = squared column norms of X torch (n, m)."""
    n, m = X.shape
    sfx = "f32" if str(X.dtype) == "torch.float32" else "f64"
    f = getattr(lib(), "spx_sqnorms_dev_" + sfx)
    f.argtypes = [_P, _P, _I, _I, _I, _P]
    _check(f(ctx.h, _tptr(X), m, m, n, _tptr(out)))


def omp_dev(ctx, d: Dictionary, X, V, R, count, L=None, eps=None, lambda1=None,
            precision="fp32", form="batched", waves_per_signal=0):
    """Device ORMP; L (int32), eps, lambda1: torch device tensors or None."""
    n = X.shape[0]
    Lout = V.shape[1]
    sfx = "f32" if d.dtype == np.float32 else "f64"
    o = spx_omp_opts(PREC[precision], FORM[form], waves_per_signal)
    lists = spx_lists_dev(_tptr(V), _tptr(R), _tptr(count), Lout, None)
    f = getattr(lib(), "spx_omp_dev_" + sfx)
    f.argtypes = [_P, _P, _P, _I, _I, _P, _I, _P, _I, _P, _I,
                  ctypes.POINTER(spx_omp_opts), ctypes.POINTER(spx_lists_dev)]
    nn = lambda t: 0 if t is None else t.numel()
    _check(f(ctx.h, d.h, _tptr(X), X.shape[1], n, _tptr(L), nn(L), _tptr(eps),
             nn(eps), _tptr(lambda1), nn(lambda1), ctypes.byref(o),
             ctypes.byref(lists)))


def omp_mask_dev(ctx, d: Dictionary, X, B, V, R, count, L=None, eps=None, lambda1=None,
                 precision="fp32", form="batched", waves_per_signal=0):
    """Device masked ORMP: B torch uint8 (n, m) (column-major m x n mask),
    other arguments as omp_dev."""
    n = X.shape[0]
    Lout = V.shape[1]
    sfx = "f32" if d.dtype == np.float32 else "f64"
    o = spx_omp_opts(PREC[precision], FORM[form], waves_per_signal)
    lists = spx_lists_dev(_tptr(V), _tptr(R), _tptr(count), Lout, None)
    f = getattr(lib(), "spx_omp_mask_dev_" + sfx)
    f.argtypes = [_P, _P, _P, _I, _I, _P, _I, _P, _I, _P, _I, _P, _I,
                  ctypes.POINTER(spx_omp_opts), ctypes.POINTER(spx_lists_dev)]
    nn = lambda t: 0 if t is None else t.numel()
    _check(f(ctx.h, d.h, _tptr(X), X.stride(0), n, _tptr(B), B.stride(0), _tptr(L), nn(L),
             _tptr(eps), nn(eps), _tptr(lambda1), nn(lambda1), ctypes.byref(o),
             ctypes.byref(lists)))


def fista_flat_dev(ctx, d: Dictionary, X, W0, W, info, lambda1, regul="l1",
                   lambda2=0.0, max_it=1000, L0=1.0, fixed_step=False,
                   gamma=1.5, tol=1e-6, it0=100, max_iter_backtracking=1000,
                   pos=False, ista=False, precision="fp32", graph=True, compute_gram=False):
    """Device FISTA: X torch (n, m), W0/W torch (n, p), info torch (n, 4).
    With a fixed step and graph=True the iteration loop is captured once per
    (buffers, sizes, options) as a hipGraph cached on the context and
    replayed.  compute_gram: gradient G alpha - D'x (prox/fista.h:364-413)."""
    n = X.shape[0]
    sfx = "f32" if d.dtype == np.float32 else "f64"
    o = spx_fista_opts(max_it, L0, int(fixed_step), gamma, lambda1, lambda2,
                       tol, it0, max_iter_backtracking,
                       {"l1": 0, "elastic-net": 1}[regul], int(pos), int(ista),
                       PREC[precision], int(bool(graph)), int(bool(compute_gram)))
    f = getattr(lib(), "spx_fista_flat_dev_" + sfx)
    f.argtypes = [_P, _P, _P, _I, _P, _P, ctypes.POINTER(spx_fista_opts), _P]
    _check(f(ctx.h, d.h, _tptr(X), n, _tptr(W0), _tptr(W), ctypes.byref(o),
             _tptr(info)))


def lists_to_csc_dev(ctx, V, R, count, pB, r, v):
    """Device CSC assembly; returns nnz."""
    n, Lout = V.shape
    sfx = "f32" if str(V.dtype) == "torch.float32" else "f64"
    lists = spx_lists_dev(_tptr(V), _tptr(R), _tptr(count), Lout, None)
    nnz = ctypes.c_longlong()
    f = getattr(lib(), "spx_lists_to_csc_dev_" + sfx)
    f.argtypes = [_P, ctypes.POINTER(spx_lists_dev), _I, _P, _P, _P,
                  ctypes.POINTER(ctypes.c_longlong)]
    _check(f(ctx.h, ctypes.byref(lists), n, _tptr(pB), _tptr(r), _tptr(v),
             ctypes.byref(nnz)))
    return nnz.value


def dtx_dev(ctx, d: Dictionary, X, out, precision="fp32"):
    """out (n, ldc) torch = D'X (column-major p x n with leading dim ldc)."""
    n = X.shape[0]
    if d.dtype == np.float32:
        f = lib().spx_dtx_dev_f32
        f.argtypes = [_P, _P, _P, _I, _I, _P, _I, _I]
        _check(f(ctx.h, d.h, _tptr(X), X.shape[1], n, _tptr(out), out.shape[1],
                 PREC[precision]))
    else:
        f = lib().spx_dtx_dev_f64
        f.argtypes = [_P, _P, _P, _I, _I, _P, _I]
        _check(f(ctx.h, d.h, _tptr(X), X.shape[1], n, _tptr(out), out.shape[1]))


def dict_gram(ctx, d: Dictionary, G):
    """G torch (p, p) receives the resident Gram matrix (column-major)."""
    sfx = "f32" if d.dtype == np.float32 else "f64"
    f = getattr(lib(), "spx_dict_gram_" + sfx)
    f.argtypes = [_P, _P, _P]
    _check(f(ctx.h, d.h, _tptr(G)))


def exported_symbols():
    """Names declared in include/spx.h (for the load check)."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "spx.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(spx_[a-z0-9_]+)\s*\(", txt)))
