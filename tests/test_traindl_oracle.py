"""Dictionary learning (trainDL): the oracle restatement of Trainer::train
(dictLearn/dicts.h:346-759) against the reference's golden fixtures, and
against what the mathematics of the method fixes independently of the
reference: the minibatch statistics A = sum a a', B = sum x a' of the lasso
codes, and the block-coordinate dictionary update as a minimiser of the
surrogate 1/2 Tr(D'DA) - Tr(D'B) over unit-ball columns."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import golden_files, load_golden

import workloads as W

PRM = ("batchsize", "lambda1", "lambda2", "iter", "t0", "mode", "posAlpha", "posD", "clean",
       "rho", "iter_updateD", "modeParam", "batch")


def golden_params(g):
    out = {}
    for k in PRM:
        v = g[k].item()
        out[k] = v
    return out


@pytest.mark.parametrize("name", golden_files("traindl_"))
def test_train_dl_matches_golden(orc, name):
    g = load_golden(name)
    prm = golden_params(g)
    X = g["X"]
    if "D0" in g:
        D1, A1, B1, it1 = orc.train_dl(X, D=g["D0"], **dict(prm, iter=7))
        assert it1 == int(g["it1"])
        for a, b in ((D1, g["D1"]), (A1, g["A1"]), (B1, g["B1"])):
            assert np.array_equal(a, b)
        D, A, B, it = orc.train_dl(X, D=D1, A=A1, B=B1, itercount=it1, **prm)
    else:
        D, A, B, it = orc.train_dl(X, D=g.get("Dinit"), K=int(g["K"]), **prm)
    assert it == int(g["itercount"])
    # the restatement performs the reference's operations in its order
    assert np.array_equal(D, g["D"])
    assert np.array_equal(A, g["A"])
    assert np.array_equal(B, g["B"])


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_train_dl_matches_reference_build(orc, dt):
    if not orc.has_ref():
        pytest.skip("reference build absent")
    rng = np.random.default_rng(41)
    D0 = W.unit_dict(rng, 12, 20, dt)
    X = W.sparse_signals(rng, D0, 150, 3, 0.05, True, dt)
    for kw in (dict(mode=2, lambda1=0.12, batchsize=32, iter=15),
               dict(mode=5, lambda1=0.02, batchsize=32, iter=15),
               dict(mode=2, lambda1=0.12, batch=True, iter=3, iter_updateD=3)):
        a = orc.train_dl(X, K=20, **kw)
        b = orc.ref_train_dl(X, K=20, **kw)
        assert a[3] == b[3]
        for u, v in zip(a[:3], b[:3]):
            assert np.array_equal(u, v)


def _problem(dt=np.float64, seed=43, m=12, K=20, n=120):
    rng = np.random.default_rng(seed)
    D0 = W.unit_dict(rng, m, K, dt)
    X = W.sparse_signals(rng, D0, n, 3, 0.05, True, dt)
    return X, D0


def test_batch_statistics_are_the_lasso_code_moments(orc):
    """Batch mode, one iteration from a given unit-norm D: A and B are the
    second moments of the lasso codes of all signals, A = sum_i a_i a_i',
    B = sum_i x_i a_i' (dicts.h:600-609), with a_i from the (separately
    pinned) lasso restatement."""
    X, D0 = _problem()
    lam = 0.1
    D, A, B, it = orc.train_dl(X, D=D0, lambda1=lam, lambda2=1e-9, iter=1, batch=True,
                               iter_updateD=0, clean=False)
    assert it == 1
    assert np.allclose(D, D0, atol=1e-15)          # no update sweep asked for
    alpha = orc.lasso(X, D0, lam, lambda2=1e-9).toarray()
    assert np.allclose(A, alpha @ alpha.T, rtol=1e-12, atol=1e-13)
    assert np.allclose(B, X @ alpha.T, rtol=1e-12, atol=1e-13)


def _surrogate(D, A, B):
    return 0.5 * np.trace(D.T @ D @ A) - np.trace(D.T @ B)


def test_dictionary_update_decreases_the_surrogate_and_reaches_its_kkt_point(orc):
    """Each sweep of the update minimises the surrogate exactly over one
    column at a time (then projects on the unit ball), so the surrogate does
    not increase from sweep to sweep, and many sweeps reach its KKT point:
    for every used atom j, g_j = D a_j - b_j is 0 inside the ball, or
    -mu d_j with mu >= 0 on its boundary."""
    X, D0 = _problem(seed=44)
    lam = 0.1
    _, A, B, _ = orc.train_dl(X, D=D0, lambda1=lam, iter=1, batch=True, iter_updateD=0,
                              clean=False)
    vals = []
    for sweeps in range(0, 6):
        D, A2, B2, _ = orc.train_dl(X, D=D0, lambda1=lam, iter=1, batch=True,
                                    iter_updateD=sweeps, clean=False)
        assert np.array_equal(A2, A) and np.array_equal(B2, B)
        vals.append(_surrogate(D, A, B))
    assert all(b <= a + 1e-12 * abs(a) for a, b in zip(vals, vals[1:]))
    assert vals[-1] < vals[0]
    D, _, _, _ = orc.train_dl(X, D=D0, lambda1=lam, iter=1, batch=True, iter_updateD=400,
                              clean=False)
    used = np.diag(A) > 1e-6
    G = D @ A - B
    for j in np.flatnonzero(used):
        d, g = D[:, j], G[:, j]
        nrm = np.linalg.norm(d)
        assert nrm <= 1 + 1e-12
        if nrm < 1 - 1e-9:
            assert np.linalg.norm(g) < 1e-7 * (1 + np.linalg.norm(B[:, j]))
        else:
            mu = -g @ d
            assert mu >= -1e-9
            assert np.linalg.norm(g + mu * d) < 1e-7 * (1 + np.linalg.norm(B[:, j]))


def test_kkt_check_detects_a_planted_update_error(orc):
    """The KKT test above fails on a dictionary update with a planted
    mistake (the update drops its + d_k term)."""
    X, D0 = _problem(seed=45)
    _, A, B, _ = orc.train_dl(X, D=D0, lambda1=0.1, iter=1, batch=True, iter_updateD=0,
                              clean=False)
    D = D0.copy()
    for _ in range(400):
        for k in range(D.shape[1]):
            if A[k, k] <= 1e-6:
                continue
            nd = (B[:, k] - D @ A[:, k]) / A[k, k]   # planted: + D[:, k] dropped
            nd /= max(1.0, np.linalg.norm(nd))
            D[:, k] = nd
    G = D @ A - B
    res = []
    for j in np.flatnonzero(np.diag(A) > 1e-6):
        d, g = D[:, j], G[:, j]
        mu = -g @ d
        res.append(np.linalg.norm(g + mu * d) if np.linalg.norm(d) > 1 - 1e-9
                   else np.linalg.norm(g))
    assert max(res) > 1e-4
