"""Pins the C restatement oracle (CPU only).

Every oracle function is checked against
  * the golden fixtures written by the reference compiled in place
    (tests/golden, bit-exact),
  * the in-place reference build itself when present (bit-exact, random
    cases),
  * closed forms (D = I), optimality conditions (lasso KKT, least-squares
    orthogonality of OMP) and brute force on tiny problems,
chosen so that a dropped term, a wrong sign or a transposed operand fails.
"""
import itertools

import numpy as np
import pytest

from conftest import csc_equal, golden_files, load_golden

import workloads as W


def dense(a):
    return np.asarray(a.todense())


# ------------------------------------------------------------ golden vectors

@pytest.mark.parametrize("name", golden_files("lasso_"))
def test_lasso_matches_golden(orc, name):
    g = load_golden(name)
    A = orc.lasso(g["X"], g["D"], g["lam"] if g["lam"].ndim else g["lam"].item(),
                  lambda2=g["lambda2"].item(), mode=int(g["mode"]), pos=bool(g["pos"]))
    import scipy.sparse as sp
    ref = sp.csc_matrix((g["v"], g["r"], g["pB"]), shape=(g["D"].shape[1], g["X"].shape[1]))
    assert csc_equal(A, ref)


@pytest.mark.parametrize("name", golden_files("omp_"))
def test_omp_matches_golden(orc, name):
    g = load_golden(name)
    L = None if int(np.atleast_1d(g["L"])[0]) < 0 else g["L"]
    eps = None if np.isnan(g["eps"]) else g["eps"].item()
    lam = None if np.isnan(g["lam"]) else g["lam"].item()
    A = orc.omp(g["X"], g["D"], L=L, eps=eps, lambda1=lam)
    import scipy.sparse as sp
    ref = sp.csc_matrix((g["v"], g["r"], g["pB"]), shape=(g["D"].shape[1], g["X"].shape[1]))
    assert csc_equal(A, ref)


@pytest.mark.parametrize("name", golden_files("fista_") + golden_files("fistagap_")
                         + golden_files("fistaista_") + golden_files("fistapos_"))
def test_fista_matches_golden(orc, name):
    """All four optim_info rows (loss, best dual, relative gap, iterations)
    and the iterates, bit for bit; the fistagap fixtures stop on the
    relative duality gap before max_it, the fistaista ones run without
    momentum, the fistapos ones through the positivity prox."""
    g = load_golden(name)
    Wt, info = orc.fista_flat(g["X"], g["D"], g["W0"], g["lambda1"].item(),
                              regul=str(g["regul"]), lambda2=g["lambda2"].item(),
                              max_it=int(g["max_it"]), L0=g["L0"].item(),
                              fixed_step=bool(g["fixed"]), tol=g["tol"].item(),
                              it0=int(g["it0"]), ista=bool(g.get("ista", 0)),
                              pos=bool(g.get("pos", 0)))
    assert np.array_equal(Wt, g["W"])
    assert np.array_equal(info, g["info"], equal_nan=True)
    if name.startswith("fistagap_"):
        assert info[3].max() < int(g["max_it"]) and np.all(info[3] % int(g["it0"]) == 0)
        assert np.all(info[2] < g["tol"])
    if name.startswith("fistapos_"):
        assert Wt.min() >= 0 and Wt.max() > 0


# ------------------------------------------------ reference build, random

def _rand(seed, m, p, n, dt, k=4, noise=0.05, norm=True):
    rng = np.random.default_rng(seed)
    D = W.unit_dict(rng, m, p, dt)
    X = W.sparse_signals(rng, D, n, min(k, p), noise, norm, dt)
    return X, D


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lasso_matches_reference_build(orc, dt, seed):
    if not orc.has_ref():
        pytest.skip("reference build absent")
    m, p = [(20, 40), (32, 16), (48, 96)][seed]
    X, D = _rand(100 + seed, m, p, 30, dt)
    for mode, lam in ((2, 0.1), (0, 0.6), (1, 0.2)):
        for pos in (False, True):
            a = orc.lasso(X, D, lam, mode=mode, pos=pos)
            b = orc.ref_lasso(X, D, lam, mode=mode, pos=pos)
            assert csc_equal(a, b), (mode, pos)
    lamv = np.linspace(0.02, 0.4, 30).astype(dt)
    assert csc_equal(orc.lasso(X, D, lamv), orc.ref_lasso_pervec(X, D, lamv))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_omp_matches_reference_build(orc, dt):
    if not orc.has_ref():
        pytest.skip("reference build absent")
    X, D = _rand(7, 24, 50, 30, dt, k=6, noise=0.01, norm=False)
    for L, eps, lam in ((6, None, None), (None, 0.01, None), (10, 0.05, 0.001)):
        assert csc_equal(orc.omp(X, D, L=L, eps=eps, lambda1=lam),
                         orc.ref_omp(X, D, L=L, eps=eps, lambda1=lam))


# ------------------------------------------------------------- closed forms

def soft(x, t):
    return np.sign(x) * np.maximum(np.abs(x) - t, 0.0)


def test_lasso_identity_dictionary_is_soft_threshold(orc):
    rng = np.random.default_rng(3)
    p = 24
    D = np.eye(p)
    X = np.asfortranarray(rng.standard_normal((p, 15)))
    for lam in (0.1, 0.5, 1.2):
        A = dense(orc.lasso(X, D, lam))
        # G = I + 1e-10 I (decomp/decomp.h:1569)
        np.testing.assert_allclose(A, soft(X, lam) / (1 + 1e-10), rtol=0, atol=1e-12)
        Apos = dense(orc.lasso(X, D, lam, pos=True))
        np.testing.assert_allclose(Apos, np.maximum(X - lam, 0) / (1 + 1e-10), atol=1e-12)


def test_omp_identity_dictionary_is_top_L(orc):
    rng = np.random.default_rng(4)
    p = 30
    X = np.asfortranarray(rng.standard_normal((p, 20)))
    A = dense(orc.omp(X, np.eye(p), L=5))
    for j in range(20):
        top = np.argsort(-np.abs(X[:, j]), kind="stable")[:5]
        expect = np.zeros(p)
        expect[top] = X[top, j]
        np.testing.assert_allclose(A[:, j], expect, atol=1e-13)


def test_fista_elastic_net_identity_closed_form(orc):
    rng = np.random.default_rng(5)
    p = 16
    X = np.asfortranarray(rng.standard_normal((p, 6)))
    W0 = np.zeros((p, 6), order="F")
    lam, lam2 = 0.3, 0.2
    Wt, _ = orc.fista_flat(X, np.eye(p), W0, lam, regul="elastic-net", lambda2=lam2,
                           max_it=500, L0=1.0, fixed_step=True, tol=0.0, it0=1000)
    np.testing.assert_allclose(Wt, soft(X, lam) / (1 + lam2), atol=1e-10)


# ---------------------------------------------------- optimality conditions

def test_lasso_kkt_conditions(orc):
    X, D = _rand(21, 40, 120, 25, np.float64)
    lam = 0.08
    A = dense(orc.lasso(X, D, lam))
    # the Gram carries +1e-10 on its diagonal (decomp/decomp.h:1569)
    C = D.T @ (X - D @ A) - 1e-10 * A
    assert np.all(np.abs(C) <= lam * (1 + 1e-9))
    on = A != 0
    np.testing.assert_allclose(C[on], lam * np.sign(A[on]), rtol=1e-9)


def test_kkt_check_detects_a_planted_error(orc):
    """A 1e-3 perturbation of one coefficient must break the KKT check."""
    X, D = _rand(21, 40, 120, 25, np.float64)
    lam = 0.08
    A = dense(orc.lasso(X, D, lam))
    i, j = np.argwhere(A != 0)[0]
    A[i, j] += 1e-3
    C = D.T @ (X - D @ A) - 1e-10 * A
    on = A != 0
    ok = np.all(np.abs(C) <= lam * (1 + 1e-9)) and np.allclose(
        C[on], lam * np.sign(A[on]), rtol=1e-9)
    assert not ok


def test_lasso_l1coeffs_reaches_the_l1_budget_at_a_kkt_point(orc):
    """L1COEFFS (mode 0): the path stops where ||alpha||_1 reaches lambda1,
    and alpha is a lasso solution at the path's final correlation level c:
    |d_j'r| <= c for every atom, with equality and matching sign on the
    support (r the residual, the Gram's 1e-10 ridge included)."""
    X, D = _rand(31, 40, 120, 25, np.float64)
    t = 0.6
    A = dense(orc.lasso(X, D, t, mode=0))
    np.testing.assert_allclose(np.abs(A).sum(0), t, rtol=1e-9)
    C = D.T @ (X - D @ A) - 1e-10 * A
    for j in range(X.shape[1]):
        on = A[:, j] != 0
        c = np.abs(C[on, j])
        np.testing.assert_allclose(c, c.mean(), rtol=1e-8)
        assert np.all(np.abs(C[:, j]) <= c.mean() * (1 + 1e-8))
        assert np.all(np.sign(C[on, j]) == np.sign(A[on, j]))
    # a planted error (budget off by 1e-4) fails the first check
    assert not np.allclose(np.abs(A).sum(0) * (1 + 1e-4), t, rtol=1e-9, atol=0)


def test_lasso_lambda2_kkt_conditions(orc):
    """lambda2 > 0 (elastic-net lasso): the Gram carries +lambda2 on its
    diagonal (decomp/decomp.h:1566-1569), so the solution satisfies
    |d_j'r - lambda2 alpha_j| <= lambda1, with equality and sign on the
    support; the same check without the lambda2 term fails."""
    X, D = _rand(32, 40, 120, 25, np.float64)
    lam, lam2 = 0.08, 0.05
    A = dense(orc.lasso(X, D, lam, lambda2=lam2))
    C = D.T @ (X - D @ A) - lam2 * A
    assert np.all(np.abs(C) <= lam * (1 + 1e-9))
    on = A != 0
    np.testing.assert_allclose(C[on], lam * np.sign(A[on]), rtol=1e-9)
    C0 = D.T @ (X - D @ A)
    assert not np.allclose(C0[on], lam * np.sign(A[on]), rtol=1e-6)


def _ls_resid2(X, D, A):
    out = np.empty(X.shape[1])
    for j in range(X.shape[1]):
        S = np.flatnonzero(A[:, j])
        r = X[:, j] - (D[:, S] @ np.linalg.lstsq(D[:, S], X[:, j], rcond=None)[0]
                       if len(S) else 0)
        out[j] = r @ r
    return out


def test_omp_eps_stops_at_the_first_residual_below_eps(orc):
    """eps stop (decomp/decomp.h:599-600): after the last selected atom the
    squared residual is <= eps; with that atom left out (the greedy prefix,
    i.e. the same solver at L = count - 1) it was still > eps."""
    rng = np.random.default_rng(33)
    D = W.unit_dict(rng, 30, 80, np.float64)
    X = W.sparse_signals(rng, D, 40, np.arange(40) % 10 + 1, 0.003, False, np.float64)
    eps = 0.01
    A = dense(orc.omp(X, D, eps=eps))
    k = (A != 0).sum(0)
    r_end = _ls_resid2(X, D, A)
    stopped = (k < 30) & (k > 0)
    assert stopped.sum() >= 30
    assert np.all(r_end[stopped] <= eps)
    Lprev = np.maximum(k - 1, 0).astype(np.int32)
    Ap = dense(orc.omp(X, D, L=Lprev))
    r_prev = _ls_resid2(X, D, Ap)
    assert np.all(r_prev[stopped] > eps)


def test_omp_lambda_stop_rule(orc):
    """lambda stop (decomp/decomp.h:566-573): an atom is admitted only if it
    lowers the squared residual by at least 2 lambda, and the path stops at
    the first best atom that would lower it by less: the last admitted atom
    lowered it by >= 2 lambda, and no further atom lowers it by 2 lambda
    (the largest decrease over candidates a is (d_a'r)^2 / ||P d_a||^2, P the
    projector off the support)."""
    X, D = _rand(34, 30, 80, 30, np.float64, k=6, noise=0.02, norm=False)
    lam = 0.01
    A = dense(orc.omp(X, D, lambda1=lam))
    k = (A != 0).sum(0)
    assert k.min() >= 1 and k.max() < 30
    r_end = _ls_resid2(X, D, A)
    r_prev = _ls_resid2(X, D, dense(orc.omp(X, D, L=(k - 1).astype(np.int32))))
    assert np.all(r_prev - r_end >= 2 * lam * (1 - 1e-9))
    for j in range(X.shape[1]):
        S = np.flatnonzero(A[:, j])
        Q, _ = np.linalg.qr(D[:, S])
        r = X[:, j] - Q @ (Q.T @ X[:, j])
        Pd = D - Q @ (Q.T @ D)
        nrm = np.sum(Pd * Pd, axis=0)
        ok = nrm > 1e-8
        dec = np.zeros(D.shape[1])
        dec[ok] = (D[:, ok].T @ r) ** 2 / nrm[ok]
        dec[S] = 0
        assert dec.max() < 2 * lam


def test_fista_backtracking_is_beck_teboulle(orc):
    """FISTA with backtracking (prox/fista.h:2950-2960, test :400-413): L
    grows by gamma until ||D(y - z)||^2 <= L ||y - z||^2 (for the square
    loss this is the quadratic upper bound of Beck & Teboulle, SIAM J.
    Imaging Sci. 2009, section 4), L persists across iterations, and the
    iterates follow the published FISTA momentum.  Restated here from the
    paper in numpy (FP64) for 6 iterations and compared with the oracle."""
    X, D = _rand(35, 20, 50, 5, np.float64, k=3, noise=0.05, norm=False)
    lam, gamma, L0, its = 0.05, 1.5, 0.1, 6
    p = D.shape[1]
    Wt, info = orc.fista_flat(X, D, np.zeros((p, 5), order="F"), lam, regul="l1",
                              max_it=its, L0=L0, fixed_step=False, gamma=gamma, tol=0.0,
                              it0=1000)
    for j in range(5):
        x = X[:, j]
        xk = np.zeros(p)
        y = xk.copy()
        t, L = 1.0, L0
        for _ in range(its):
            g = D.T @ (D @ y - x)
            while True:
                z = soft(y - g / L, lam / L)
                dz = y - z
                if np.sum((D @ dz) ** 2) <= L * np.sum(dz * dz):
                    break
                L *= gamma
            tn = (1 + np.sqrt(1 + 4 * t * t)) / 2
            y = z + ((t - 1) / tn) * (z - xk)
            xk, t = z, tn
        np.testing.assert_allclose(Wt[:, j], xk, atol=1e-12)
    assert np.all(info[3] == its + 1)


def test_omp_least_squares_on_support(orc):
    X, D = _rand(22, 30, 80, 20, np.float64, k=6, noise=0.01, norm=False)
    A = dense(orc.omp(X, D, L=8))
    for j in range(X.shape[1]):
        S = np.flatnonzero(A[:, j])
        coef, *_ = np.linalg.lstsq(D[:, S], X[:, j], rcond=None)
        np.testing.assert_allclose(A[S, j], coef, rtol=1e-9, atol=1e-12)
        r = X[:, j] - D[:, S] @ coef
        assert np.abs(D[:, S].T @ r).max() < 1e-10


# ------------------------------------------------------------- brute force

def test_ormp_matches_exhaustive_greedy(orc):
    """ORMP = at each step add the atom whose least-squares refit on the
    enlarged support leaves the smallest residual."""
    X, D = _rand(23, 10, 12, 8, np.float64, k=3, noise=0.01, norm=False)
    L = 4
    A = dense(orc.omp(X, D, L=L))
    for j in range(X.shape[1]):
        x = X[:, j]
        S = []
        for _ in range(L):
            best, bestr = None, np.inf
            for a in range(D.shape[1]):
                if a in S:
                    continue
                c, *_ = np.linalg.lstsq(D[:, S + [a]], x, rcond=None)
                r = np.linalg.norm(x - D[:, S + [a]] @ c)
                if r < bestr - 1e-12:
                    best, bestr = a, r
            S.append(best)
        assert set(np.flatnonzero(A[:, j])) == set(S)


def test_lasso_matches_sign_pattern_enumeration(orc):
    """Tiny lasso solved exactly by enumerating supports and signs."""
    X, D = _rand(24, 6, 5, 6, np.float64, k=2, noise=0.1, norm=True)
    lam = 0.05
    A = dense(orc.lasso(X, D, lam))
    G = D.T @ D
    for j in range(X.shape[1]):
        x = X[:, j]
        best, bestobj = np.zeros(5), 0.5 * x @ x
        for k in range(1, 6):
            for S in itertools.combinations(range(5), k):
                S = list(S)
                for sg in itertools.product((-1.0, 1.0), repeat=k):
                    sg = np.array(sg)
                    try:
                        a = np.linalg.solve(G[np.ix_(S, S)], D[:, S].T @ x - lam * sg)
                    except np.linalg.LinAlgError:
                        continue
                    if np.any(np.sign(a) != sg):
                        continue
                    full = np.zeros(5)
                    full[S] = a
                    obj = 0.5 * np.sum((x - D @ full) ** 2) + lam * np.abs(full).sum()
                    if obj < bestobj - 1e-14:
                        best, bestobj = full, obj
        np.testing.assert_allclose(A[:, j], best, atol=1e-8)


def test_fista_l1_converges_to_lars(orc):
    X, D = _rand(25, 20, 40, 6, np.float64, k=3, noise=0.05, norm=True)
    lam = 0.05
    A = dense(orc.lasso(X, D, lam))
    L0 = float(np.linalg.eigvalsh(D.T @ D)[-1])
    Wt, info = orc.fista_flat(X, D, np.zeros((40, 6), order="F"), lam, regul="l1",
                              max_it=20000, L0=L0, fixed_step=True, tol=1e-14, it0=100)
    np.testing.assert_allclose(Wt, A, atol=1e-6)
    assert np.all(info[2] < 1e-6)


# ------------------------------------ paths, Gram-given form, least squares

@pytest.mark.parametrize("name", golden_files("lassopath_"))
def test_lasso_path_matches_golden(orc, name):
    """Signal 0's regularisation path: the restatement writes exactly the
    entries the reference writes (NaN in the fixture marks the others,
    which the reference leaves uninitialised) with the same values."""
    g = load_golden(name)
    A, path = orc.lasso(g["X"], g["D"], g["lam"].item(), return_reg_path=True)
    import scipy.sparse as sp
    ref = sp.csc_matrix((g["v"], g["r"], g["pB"]), shape=(g["D"].shape[1], g["X"].shape[1]))
    assert csc_equal(A, ref)
    assert np.array_equal(np.isnan(path), np.isnan(g["path"]))
    w = ~np.isnan(path)
    assert w.sum() > 10 and np.array_equal(path[w], g["path"][w])


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_lasso_path_matches_reference_build(orc, dt):
    if not orc.has_ref():
        pytest.skip("reference build not present")
    rng = np.random.default_rng(31)
    D = W.unit_dict(rng, 32, 96, dt)
    X = W.sparse_signals(rng, D, 6, 4, 0.05, True, dt)
    for lam, mlp in ((0.02, -1), (0.1, 7)):
        _, path = orc.lasso(X, D, lam, max_length_path=mlp, return_reg_path=True)
        _, rpath = orc.ref_lasso(X, D, lam, max_length_path=mlp, return_reg_path=True)
        w = ~np.isnan(path)
        assert w.any() and np.array_equal(path[w], rpath[w])


@pytest.mark.parametrize("name", golden_files("lassoqq_"))
def test_lasso_qq_matches_golden(orc, name):
    g = load_golden(name)
    A = orc.lasso_qq(g["X"], g["Q"], g["q"], g["lam"].item())
    import scipy.sparse as sp
    ref = sp.csc_matrix((g["v"], g["r"], g["pB"]), shape=(g["Q"].shape[1], g["X"].shape[1]))
    assert csc_equal(A, ref)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_lasso_qq_matches_reference_build_and_uses_q_as_given(orc, dt):
    """Bit-exact with the reference build; and Q is used as given (no
    ridge added): Q = D'D + 0.1 I reproduces the D form with lambda2 = 0.1,
    whose Gram carries that ridge."""
    if not orc.has_ref():
        pytest.skip("reference build not present")
    rng = np.random.default_rng(32)
    D = W.unit_dict(rng, 32, 96, dt)
    X = W.sparse_signals(rng, D, 8, 4, 0.05, True, dt)
    Q = (D.T.astype(np.float64) @ D).astype(dt)
    q = (D.T.astype(np.float64) @ X).astype(dt)
    a = orc.lasso_qq(X, Q, q, 0.05)
    b = orc.ref_lasso_qq(X, Q, q, 0.05)
    assert csc_equal(a, b)
    if dt == np.float64:
        Qs = Q + 0.1 * np.eye(96)
        c = orc.lasso_qq(X, Qs, q, 0.05).toarray()
        d = orc.lasso(X, D, 0.05, lambda2=0.1).toarray()
        np.testing.assert_allclose(c, d, atol=1e-10)


def test_reference_ols_is_least_squares_on_the_lasso_support(orc):
    """ols (decomp/decomp.h:1087-1091) returns, on each signal's final
    active set A, (D_A'D_A + 1e-10 I)^-1 D_A'x; the active set is the one
    of the plain path (FP64)."""
    if not orc.has_ref():
        pytest.skip("reference build not present")
    rng = np.random.default_rng(33)
    D = W.unit_dict(rng, 48, 128, np.float64)
    X = W.sparse_signals(rng, D, 12, 5, 0.05, True, np.float64)
    base = orc.lasso(X, D, 0.05).toarray()
    ols = orc.ref_lasso(X, D, 0.05, ols=True).toarray()
    assert np.array_equal(base != 0, ols != 0)
    for j in range(X.shape[1]):
        S = np.nonzero(ols[:, j])[0]
        Ds = D[:, S]
        x = np.linalg.solve(Ds.T @ Ds + 1e-10 * np.eye(len(S)), Ds.T @ X[:, j])
        np.testing.assert_allclose(ols[S, j], x, rtol=1e-9, atol=1e-12)
    # a planted error (the lasso coefficients) is far from the fit
    assert np.abs(base - ols).max() > 1e-3


@pytest.mark.parametrize("name", [f for f in golden_files("lassools_") if "f64_mode" in f])
def test_reference_ols_fixtures_are_least_squares_fits(orc, name):
    """The ols fixtures (reference _lassoD with ols in modes 0-2, FP64):
    every column is the least-squares fit (D_A'D_A + 1e-10 I)^-1 D_A'x on its
    support; in L2ERROR mode the path stops on the least-squares residual,
    so the fit's squared residual is within the target lambda (the homotopy
    residual of the plain path at that support would not be)."""
    g = load_golden(name)
    import scipy.sparse as sp
    A = sp.csc_matrix((g["v"], g["r"], g["pB"]),
                      shape=(g["D"].shape[1], g["X"].shape[1])).toarray()
    X, D, lam = g["X"], g["D"], g["lam"].item()
    for j in range(X.shape[1]):
        S = np.nonzero(A[:, j])[0]
        if not len(S):
            continue
        Ds = D[:, S]
        c = np.linalg.solve(Ds.T @ Ds + 1e-10 * np.eye(len(S)), Ds.T @ X[:, j])
        np.testing.assert_allclose(A[S, j], c, rtol=1e-8, atol=1e-11)
        if int(g["mode"]) == 1:
            r = X[:, j] - Ds @ c
            assert r @ r <= lam + 1e-12
    if int(g["mode"]) == 1:
        # the least-squares stop ends earlier than coreLARS2's homotopy stop
        B = orc.lasso(X, D, lam, mode=1).toarray()
        assert (A != 0).sum() < (B != 0).sum()


@pytest.mark.parametrize("name", golden_files("lassochol_"))
def test_lasso_cholesky_matches_golden(orc, name):
    """cholesky=True (coreLARS, decomp/decomp.h:844-1093): the restatement
    keeps an explicit inverse where the reference updates a Cholesky factor,
    so agreement is to rounding: FP64 same supports and 1e-12; FP32 (where
    an atom at rounding distance from a tie may enter or not) same supports
    on 95 % of the signals and 2e-5 on those."""
    g = load_golden(name)
    A = orc.lasso(g["X"], g["D"], g["lam"].item(), lambda2=g["lambda2"].item(),
                  mode=int(g["mode"]), pos=bool(g["pos"]), cholesky=True).toarray()
    import scipy.sparse as sp
    ref = sp.csc_matrix((g["v"], g["r"], g["pB"]),
                        shape=(g["D"].shape[1], g["X"].shape[1])).toarray()
    if g["X"].dtype == np.float64:
        assert np.array_equal(A != 0, ref != 0)
        np.testing.assert_allclose(A, ref, atol=1e-12)
    else:
        same = np.all((A != 0) == (ref != 0), axis=0)
        assert same.mean() >= 0.95
        assert np.abs(A[:, same] - ref[:, same]).max() <= 2e-5


def _breakpoint_gaps(X, D, A, pos):
    """(c_active - max inactive |correlation|) / c_active per signal."""
    gaps = []
    for j in range(X.shape[1]):
        c = D.T @ (X[:, j] - D @ A[:, j])
        if pos:
            c = np.maximum(c, 0)
        S = A[:, j] != 0
        ca = np.abs(c[S]).mean()
        gaps.append((ca - np.abs(c[~S]).max()) / ca)
    return np.asarray(gaps)


@pytest.mark.parametrize("name", [f for f in golden_files("lassochol_") if "f64_mode1" in f])
def test_cholesky_l2error_ends_at_a_breakpoint_below_the_target(orc, name):
    """In L2ERROR mode coreLARS takes whole homotopy steps (no cut at the
    residual target, decomp/decomp.h:997-1033) and tests the target after
    admissions (:1076): every signal ends at a breakpoint of the path (the
    next atom's correlation ties the active ones) with its residual below
    lambda.  coreLARS2's cut step (the planted mistake: the rule of the other
    variant) lands on the target between breakpoints and fails the check."""
    g = load_golden(name)
    X, D, lam, pos = g["X"], g["D"], g["lam"].item(), bool(g["pos"])
    import scipy.sparse as sp
    ref = sp.csc_matrix((g["v"], g["r"], g["pB"]), shape=(D.shape[1], X.shape[1])).toarray()
    for A in (orc.lasso(X, D, lam, mode=1, pos=pos, cholesky=True).toarray(), ref):
        res = ((X - D @ A) ** 2).sum(0)
        assert (res <= lam + 1e-12).all()
        assert np.abs(_breakpoint_gaps(X, D, A, pos)).max() < 1e-8
    B = orc.lasso(X, D, lam, mode=1, pos=pos).toarray()
    np.testing.assert_allclose(((X - D @ B) ** 2).sum(0), lam, atol=1e-9)
    assert _breakpoint_gaps(X, D, B, pos).min() > 1e-5


@pytest.mark.parametrize("name", golden_files("omppath_"))
def test_omp_path_fixture_is_consistent(orc, name):
    """The reference's OMP path (decomp/decomp.h:605-613, :634-639): the
    last written column holds the final coefficients, earlier column j holds
    a (j+1)-atom least-squares fit on the first j+1 selected atoms."""
    g = load_golden(name)
    A = orc.omp(g["X"], g["D"], L=int(g["L"]))
    import scipy.sparse as sp
    ref = sp.csc_matrix((g["v"], g["r"], g["pB"]), shape=(g["D"].shape[1], g["X"].shape[1]))
    assert csc_equal(A, ref)
    path = g["path"].astype(np.float64)
    a0 = ref.toarray()[:, 0]
    k = int((a0 != 0).sum())
    np.testing.assert_array_equal(path[:, k - 1], ref.toarray()[:, 0])
    D = g["D"].astype(np.float64)
    x = g["X"][:, 0].astype(np.float64)
    for j in range(k - 1):
        S = np.nonzero(path[:, j])[0]
        assert len(S) == j + 1
        c = np.linalg.lstsq(D[:, S], x, rcond=None)[0]
        tol = 1e-9 if g["X"].dtype == np.float64 else 1e-3
        np.testing.assert_allclose(path[S, j], c, rtol=tol, atol=tol)
    assert not path[:, k:].any()


# ------------------------------------------------------- weighted penalty

@pytest.mark.parametrize("name", golden_files("lassoweighted_"))
def test_lasso_weighted_matches_golden(orc, name):
    import scipy.sparse as sp
    g = load_golden(name)
    A = orc.lasso_weighted(g["X"], g["D"], g["W"], g["lam"].item(), mode=int(g["mode"]),
                           pos=bool(g["pos"]))
    ref = sp.csc_matrix((g["v"], g["r"], g["pB"]), shape=(g["D"].shape[1], g["X"].shape[1]))
    assert csc_equal(A, ref)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_lasso_weighted_matches_reference_build(orc, dt):
    if not orc.has_ref():
        pytest.skip("reference build absent")
    X, D = _rand(31, 24, 60, 40, dt)
    Wt = np.random.default_rng(32).uniform(0.3, 3.0, (60, 40)).astype(dt)
    for mode, lam in ((2, 0.1), (0, 0.9), (1, 0.25)):
        for pos in (False, True):
            a = orc.lasso_weighted(X, D, Wt, lam, mode=mode, pos=pos)
            b = orc.ref_lasso_weighted(X, D, Wt, lam, mode=mode, pos=pos)
            assert csc_equal(a, b), (mode, pos)


def test_lasso_weighted_identity_dictionary_is_weighted_soft_threshold(orc):
    # D = I: a_j = soft(x_j, lambda w_j) / (1 + 1e-10)
    rng = np.random.default_rng(33)
    p = 20
    X = np.asfortranarray(rng.standard_normal((p, 12)))
    Wt = np.asfortranarray(rng.uniform(0.5, 2.0, (p, 12)))
    for lam in (0.1, 0.4):
        A = dense(orc.lasso_weighted(X, np.eye(p), Wt, lam))
        np.testing.assert_allclose(A, soft(X, lam * Wt) / (1 + 1e-10), rtol=0, atol=1e-12)


def _weighted_kkt_violation(X, D, Wt, A, lam):
    # |d_j' r| <= lam w_j, with d_j' r = lam w_j sign(a_j) on the support
    R = X - D @ A
    C = D.T @ R
    worst = 0.0
    for i in range(X.shape[1]):
        sup = A[:, i] != 0
        worst = max(worst, np.max(np.abs(C[:, i]) - lam * Wt[:, i]))
        if sup.any():
            worst = max(worst, np.max(np.abs(C[sup, i] - lam * Wt[sup, i] * np.sign(A[sup, i]))))
    return worst


def test_lasso_weighted_kkt_conditions(orc):
    X, D = _rand(34, 32, 80, 30, np.float64)
    Wt = np.asfortranarray(np.random.default_rng(35).uniform(0.5, 2.0, (80, 30)))
    lam = 0.08
    A = dense(orc.lasso_weighted(X, D, Wt, lam))
    assert _weighted_kkt_violation(X, D, Wt, A, lam) < 1e-8
    # a transposed or dropped weight breaks the conditions
    B = dense(orc.lasso_weighted(X, D, np.ones_like(Wt), lam))
    assert _weighted_kkt_violation(X, D, Wt, B, lam) > 1e-3


# ------------------------------------------------------ partly observed signals

def _golden_csc_dense(g):
    import scipy.sparse as sp
    # dense: the reference stores L explicit zeros at row 0 for a column
    # with no observed row
    return np.asarray(sp.csc_matrix((g["v"], g["r"], g["pB"]),
                                    shape=(g["D"].shape[1], g["X"].shape[1])).todense())


@pytest.mark.parametrize("name", golden_files("lassomask_"))
def test_lasso_mask_matches_golden(orc, name):
    g = load_golden(name)
    A = dense(orc.lasso_mask(g["X"], g["D"], g["B"], g["lam"].item(), lambda2=g["lambda2"].item(),
                             mode=int(g["mode"]), pos=bool(g["pos"])))
    assert np.array_equal(A, _golden_csc_dense(g))


@pytest.mark.parametrize("name", golden_files("ompmask_"))
def test_omp_mask_matches_golden(orc, name):
    g = load_golden(name)
    L = None if int(np.atleast_1d(g["L"])[0]) < 0 else g["L"]
    eps = None if np.isnan(g["eps"]) else g["eps"].item()
    lam = None if np.isnan(g["lam"]) else g["lam"].item()
    A = dense(orc.omp_mask(g["X"], g["D"], g["B"], L=L, eps=eps, lambda1=lam))
    assert np.array_equal(A, _golden_csc_dense(g))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_masks_match_reference_build(orc, dt):
    if not orc.has_ref():
        pytest.skip("reference build absent")
    X, D = _rand(41, 24, 50, 30, dt)
    B = (np.random.default_rng(42).uniform(size=X.shape) < 0.7).astype(np.uint8)
    B[:, 3] = 0
    B[:, 4] = 1
    for mode, lam in ((2, 0.1), (0, 0.6), (1, 0.2)):
        a = dense(orc.lasso_mask(X, D, B, lam, mode=mode))
        b = dense(orc.ref_lasso_mask(X, D, B, lam, mode=mode))
        assert np.array_equal(a, b), mode
    for L, eps in ((5, None), (None, 0.02)):
        a = dense(orc.omp_mask(X, D, B, L=L, eps=eps))
        b = dense(orc.ref_omp_mask(X, D, B, L=L, eps=eps))
        assert np.array_equal(a, b), (L, eps)


def test_masks_reduce_to_the_plain_solvers(orc):
    """All rows observed = the plain path; no row observed = empty; a masked
    signal = the plain path over its observed rows with lambda (eps) scaled
    by |M| / m and 1e-10 added to the Gram diagonal."""
    X, D = _rand(43, 20, 40, 12, np.float64)
    ones = np.ones(X.shape, np.uint8)
    np.testing.assert_array_equal(dense(orc.lasso_mask(X, D, ones, 0.1)),
                                  dense(orc.lasso(X, D, 0.1)))
    np.testing.assert_array_equal(dense(orc.omp_mask(X, D, ones, L=4)),
                                  dense(orc.omp(X, D, L=4)))
    assert orc.lasso_mask(X, D, np.zeros_like(ones), 0.1).nnz == 0
    assert orc.omp_mask(X, D, np.zeros_like(ones), L=4).nnz == 0
    B = ones.copy()
    B[::3, :] = 0
    keep = B[:, 0] != 0
    cnt = int(keep.sum())
    A = dense(orc.lasso_mask(X, D, B, 0.1))
    # the plain solver on the observed rows: lambda * cnt / m; its ridge
    # 1e-10 is the masked path's default
    Aref = dense(orc.lasso(np.asfortranarray(X[keep]), np.asfortranarray(D[keep]),
                           0.1 * cnt / X.shape[0]))
    np.testing.assert_allclose(A, Aref, atol=1e-12)
    # OMP: 1e-10 on the masked Gram's diagonal changes nothing at 1e-9
    Ao = dense(orc.omp_mask(X, D, B, L=4))
    Aoref = dense(orc.omp(np.asfortranarray(X[keep]), np.asfortranarray(D[keep]), L=4))
    np.testing.assert_allclose(Ao, Aoref, atol=1e-8)
