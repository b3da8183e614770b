"""The oracle's Fortran-ABI BLAS subset (oracle/blas_body.inc) against numpy.

The oracle restatement and the reference build compiled in place link the
same BLAS, so their bit equality cannot catch a defect in it.  Here every
routine the decomposition path calls (linalg/cblas_alt_template.h:291-370)
is checked on its own: FP64 against numpy's FP64 products within 1e-12
(relative to the result's scale), FP32 within 1e-5, with non-unit strides,
leading dimensions above the row count, both triangles, both transposes,
unit and non-unit diagonals, alpha/beta scaling and the index tie rules of
i?amax (first index of the largest |x|) and the custom iamin (first index of
the smallest |x|, strict <; linalg/cblas_alt_template.h:1094-1105).
"""
import ctypes

import numpy as np
import pytest

_P = ctypes.c_void_p
_I = ctypes.c_int


@pytest.fixture(scope="module")
def blas(orc):
    return orc.lib()


def ip(v):
    return ctypes.byref(ctypes.c_int(v))


def ch(c):
    return ctypes.c_char_p(c.encode())


def sc(v, dt):
    return ctypes.byref((ctypes.c_float if dt == np.float32 else ctypes.c_double)(v))


def ptr(a):
    return a.ctypes.data_as(_P)


def pre(dt):
    return "s" if dt == np.float32 else "d"


def tol(dt):
    return 1e-5 if dt == np.float32 else 1e-12


def close(got, ref, dt):
    scale = max(1.0, float(np.abs(ref).max(initial=0)))
    assert np.abs(np.asarray(got, np.float64) - ref).max(initial=0) <= tol(dt) * scale


def fmat(rng, rows, cols, ld, dt):
    """Column-major rows x cols inside an ld x cols buffer (padding rows hold
    a sentinel that must stay untouched)."""
    buf = np.full((cols, ld), 7.0, dt)
    buf[:, :rows] = rng.standard_normal((cols, rows))
    return buf


def view(buf, rows):
    return buf[:, :rows].T.astype(np.float64)


DTS = [np.float32, np.float64]


@pytest.mark.parametrize("dt", DTS)
def test_level1(blas, dt):
    rng = np.random.default_rng(1)
    n, inc = 37, 3
    x = rng.standard_normal(n * inc).astype(dt)
    y = rng.standard_normal(n * 2).astype(dt)
    xs, ys = x[::inc].astype(np.float64), y[::2].astype(np.float64)
    p = pre(dt)
    rt = ctypes.c_float if dt == np.float32 else ctypes.c_double
    for name, ref in (("dot", xs @ ys),):
        f = getattr(blas, p + name + "_")
        f.restype = rt
        close(f(ip(n), ptr(x), ip(inc), ptr(y), ip(2)), ref, dt)
    f = getattr(blas, p + "nrm2_")
    f.restype = rt
    close(f(ip(n), ptr(x), ip(inc)), np.linalg.norm(xs), dt)
    f = getattr(blas, p + "asum_")
    f.restype = rt
    close(f(ip(n), ptr(x), ip(inc)), np.abs(xs).sum(), dt)
    # axpy: y += a x; a = 0 leaves y untouched
    y2 = y.copy()
    getattr(blas, p + "axpy_")(ip(n), sc(0.75, dt), ptr(x), ip(inc), ptr(y2), ip(2))
    close(y2[::2], ys + 0.75 * xs, dt)
    assert np.array_equal(y2[1::2], y[1::2])
    y3 = y.copy()
    getattr(blas, p + "axpy_")(ip(n), sc(0.0, dt), ptr(x), ip(inc), ptr(y3), ip(2))
    assert np.array_equal(y3, y)
    # axpby: y = a x + b y
    y4 = y.copy()
    getattr(blas, p + "axpby_")(ip(n), sc(-1.5, dt), ptr(x), ip(inc), sc(0.25, dt), ptr(y4), ip(2))
    close(y4[::2], -1.5 * xs + 0.25 * ys, dt)
    # scal, copy
    x2 = x.copy()
    getattr(blas, p + "scal_")(ip(n), sc(-2.0, dt), ptr(x2), ip(inc))
    close(x2[::inc], -2.0 * xs, dt)
    assert np.array_equal(x2[1::inc], x[1::inc])
    z = np.zeros(n * 2, dt)
    getattr(blas, p + "copy_")(ip(n), ptr(x), ip(inc), ptr(z), ip(2))
    assert np.array_equal(z[::2], x[::inc]) and not z[1::2].any()


@pytest.mark.parametrize("dt", DTS)
def test_iamax_and_iamin_tie_rules(blas, orc, dt):
    p = pre(dt)
    f = getattr(blas, "i" + p + "amax_")
    f.restype = _I
    x = np.array([0.5, -3.0, 2.0, 3.0, -3.0, 1.0], dt)
    assert f(ip(6), ptr(x), ip(1)) == 2          # 1-based, first of the tied |x| = 3
    x2 = np.array([1.0, 9.0, -4.0, 0.0, -4.0, 9.0], dt)
    assert f(ip(3), ptr(x2), ip(2)) == 2         # strided: 1, -4, -4 -> first -4
    assert f(ip(0), ptr(x), ip(1)) == 0
    rng = np.random.default_rng(2)
    v = rng.standard_normal(101).astype(dt)
    assert f(ip(101), ptr(v), ip(1)) - 1 == int(np.argmax(np.abs(v)))
    g = getattr(orc.lib(), "orc_iamin_" + ("f32" if dt == np.float32 else "f64"))
    g.restype = _I
    y = np.array([0.7, -0.2, 0.5, 0.2, -0.2], dt)
    assert g(5, ptr(y)) == 1                      # 0-based, first of the tied |x| = 0.2
    assert g(101, ptr(v)) == int(np.argmin(np.abs(v)))
    w = np.array([-0.0, 0.0, 1.0], dt)
    assert g(3, ptr(w)) == 0                      # |-0| == |0|: strict < keeps the first


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("trans", ["N", "T"])
def test_gemv(blas, dt, trans):
    rng = np.random.default_rng(3)
    m, n, ld = 13, 9, 17
    A = fmat(rng, m, n, ld, dt)
    Ad = view(A, m)
    lx, ly = (n, m) if trans == "N" else (m, n)
    x = rng.standard_normal(lx * 2).astype(dt)
    y = rng.standard_normal(ly).astype(dt)
    ref = 1.25 * (Ad if trans == "N" else Ad.T) @ x[::2].astype(np.float64) - 0.5 * y
    getattr(blas, pre(dt) + "gemv_")(ch(trans), ip(m), ip(n), sc(1.25, dt), ptr(A), ip(ld),
                                     ptr(x), ip(2), sc(-0.5, dt), ptr(y), ip(1))
    close(y, ref, dt)
    assert np.all(A[:, m:] == 7.0)
    # beta = 0 overwrites y even where it held a NaN
    y0 = np.full(ly, np.nan, dt)
    getattr(blas, pre(dt) + "gemv_")(ch(trans), ip(m), ip(n), sc(1.0, dt), ptr(A), ip(ld),
                                     ptr(x), ip(2), sc(0.0, dt), ptr(y0), ip(1))
    close(y0, (Ad if trans == "N" else Ad.T) @ x[::2].astype(np.float64), dt)


@pytest.mark.parametrize("dt", DTS)
def test_ger(blas, dt):
    rng = np.random.default_rng(4)
    m, n, ld = 8, 11, 10
    A = fmat(rng, m, n, ld, dt)
    x = rng.standard_normal(m).astype(dt)
    y = rng.standard_normal(n * 3).astype(dt)
    ref = view(A, m) + 0.5 * np.outer(x, y[::3])
    getattr(blas, pre(dt) + "ger_")(ip(m), ip(n), sc(0.5, dt), ptr(x), ip(1), ptr(y), ip(3),
                                    ptr(A), ip(ld))
    close(view(A, m), ref, dt)


def sym_from(A, n, uplo):
    """The symmetric matrix a routine reads from the `uplo` triangle."""
    M = view(A, n)
    T = np.triu(M) if uplo == "U" else np.tril(M)
    return T + T.T - np.diag(np.diag(M))


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_symv_and_syr(blas, dt, uplo):
    rng = np.random.default_rng(5)
    n, ld = 12, 15
    A = fmat(rng, n, n, ld, dt)
    S = sym_from(A, n, uplo)
    x = rng.standard_normal(n * 2).astype(dt)
    y = rng.standard_normal(n).astype(dt)
    ref = 0.75 * S @ x[::2].astype(np.float64) + 2.0 * y
    getattr(blas, pre(dt) + "symv_")(ch(uplo), ip(n), sc(0.75, dt), ptr(A), ip(ld), ptr(x),
                                     ip(2), sc(2.0, dt), ptr(y), ip(1))
    close(y, ref, dt)
    # syr updates only the `uplo` triangle: A += a x x'
    before = view(A, n).copy()
    xs = x[::2].astype(np.float64)
    getattr(blas, pre(dt) + "syr_")(ch(uplo), ip(n), sc(-0.3, dt), ptr(x), ip(2), ptr(A), ip(ld))
    after = view(A, n)
    tri = np.triu(np.ones((n, n), bool)) if uplo == "U" else np.tril(np.ones((n, n), bool))
    close(after[tri], (before - 0.3 * np.outer(xs, xs))[tri], dt)
    assert np.array_equal(after[~tri], before[~tri])


def tri_from(A, n, uplo, diag):
    M = view(A, n)
    T = np.triu(M) if uplo == "U" else np.tril(M)
    if diag == "U":
        T = T - np.diag(np.diag(T)) + np.eye(n)
    return T


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trmv(blas, dt, uplo, trans, diag):
    rng = np.random.default_rng(6)
    n, ld = 10, 13
    A = fmat(rng, n, n, ld, dt)
    T = tri_from(A, n, uplo, diag)
    x = rng.standard_normal(n * 2).astype(dt)
    ref = (T if trans == "N" else T.T) @ x[::2].astype(np.float64)
    getattr(blas, pre(dt) + "trmv_")(ch(uplo), ch(trans), ch(diag), ip(n), ptr(A), ip(ld),
                                     ptr(x), ip(2))
    close(x[::2], ref, dt)


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_trmm(blas, dt, side, uplo, trans):
    rng = np.random.default_rng(7)
    m, n, lda, ldb = 6, 8, 11, 9
    k = m if side == "L" else n
    A = fmat(rng, k, k, lda, dt)
    T = tri_from(A, k, uplo, "N")
    op = T if trans == "N" else T.T
    B = fmat(rng, m, n, ldb, dt)
    Bd = view(B, m)
    ref = -0.5 * (op @ Bd if side == "L" else Bd @ op)
    getattr(blas, pre(dt) + "trmm_")(ch(side), ch(uplo), ch(trans), ch("N"), ip(m), ip(n),
                                     sc(-0.5, dt), ptr(A), ip(lda), ptr(B), ip(ldb))
    close(view(B, m), ref, dt)
    assert np.all(B[:, m:] == 7.0)


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trtri(blas, dt, uplo, diag):
    rng = np.random.default_rng(8)
    n, ld = 9, 12
    A = fmat(rng, n, n, ld, dt)
    A[np.arange(n), np.arange(n)] = 2.0 + rng.uniform(size=n)   # well-conditioned diagonal
    T = tri_from(A, n, uplo, diag)
    info = ctypes.c_int(-1)
    getattr(blas, pre(dt) + "trtri_")(ch(uplo), ch(diag), ip(n), ptr(A), ip(ld),
                                      ctypes.byref(info))
    assert info.value == 0
    inv = tri_from(A, n, uplo, diag)
    close(inv, np.linalg.inv(T), dt)
    # a zero on a non-unit diagonal is reported (1-based), the matrix left as is
    Z = fmat(rng, 4, 4, 4, dt)
    Z[2, 2] = 0.0
    getattr(blas, pre(dt) + "trtri_")(ch(uplo), ch("N"), ip(4), ptr(Z), ip(4), ctypes.byref(info))
    assert info.value == 3


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_gemm(blas, dt, ta, tb):
    rng = np.random.default_rng(9)
    m, n, k = 7, 5, 6
    lda, ldb, ldc = 10, 9, 8
    A = fmat(rng, m if ta == "N" else k, k if ta == "N" else m, lda, dt)
    B = fmat(rng, k if tb == "N" else n, n if tb == "N" else k, ldb, dt)
    C = fmat(rng, m, n, ldc, dt)
    Ad = view(A, m if ta == "N" else k)
    Bd = view(B, k if tb == "N" else n)
    ref = 1.5 * (Ad if ta == "N" else Ad.T) @ (Bd if tb == "N" else Bd.T) - 0.25 * view(C, m)
    getattr(blas, pre(dt) + "gemm_")(ch(ta), ch(tb), ip(m), ip(n), ip(k), sc(1.5, dt), ptr(A),
                                     ip(lda), ptr(B), ip(ldb), sc(-0.25, dt), ptr(C), ip(ldc))
    close(view(C, m), ref, dt)


@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_syrk(blas, dt, uplo, trans):
    rng = np.random.default_rng(10)
    n, k, lda, ldc = 6, 9, 12, 8
    A = fmat(rng, n if trans == "N" else k, k if trans == "N" else n, lda, dt)
    Ad = view(A, n if trans == "N" else k)
    C = fmat(rng, n, n, ldc, dt)
    before = view(C, n).copy()
    prod = Ad @ Ad.T if trans == "N" else Ad.T @ Ad
    getattr(blas, pre(dt) + "syrk_")(ch(uplo), ch(trans), ip(n), ip(k), sc(2.0, dt), ptr(A),
                                     ip(lda), sc(0.5, dt), ptr(C), ip(ldc))
    tri = np.triu(np.ones((n, n), bool)) if uplo == "U" else np.tril(np.ones((n, n), bool))
    close(view(C, n)[tri], (2.0 * prod + 0.5 * before)[tri], dt)
    assert np.array_equal(view(C, n)[~tri], before[~tri])
