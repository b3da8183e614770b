"""The spams.h-compatible C++ templates (include/spams_amd.h) run on the GPU
as a reference C++ caller would use them: tests/cpp/dropin_run.cpp is built
against libspx.so and run in a child process (_lassoD, lasso with ols,
lasso2 with D and with G/DtX, _omp, _lassoWeighted, _lassoMask, _ompMask,
_fistaFlat, cppLasso, cppOMP, _alltrainDL), and its results are checked
against the oracle (FP64; within 1e-9), the least-squares refit against the
least-squares fit on its support."""
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as ssp

import workloads as W
from conftest import assert_csc_layout

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_csc_raw(path, p):
    """The SpMatrix<double> a C++ caller received, as written by
    dropin_run.cpp: n, nnz, pB[n+1], r[nnz], v[nnz]."""
    b = open(path, "rb").read()
    n = int(np.frombuffer(b, np.int32, 1, 0)[0])
    nnz = int(np.frombuffer(b, np.int64, 1, 4)[0])
    o = 12
    pB = np.frombuffer(b, np.int32, n + 1, o); o += 4 * (n + 1)
    r = np.frombuffer(b, np.int32, nnz, o); o += 4 * nnz
    v = np.frombuffer(b, np.float64, nnz, o)
    return ssp.csc_matrix((v, r, pB), shape=(p, n))


def read_csc(path, p):
    return read_csc_raw(path, p).toarray()


def test_cpp_dropin_runs_and_matches_oracle(tmp_path):
    import oracle
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ absent")
    lib_dir = os.path.join(ROOT, "spams-devel_amd")
    exe = str(tmp_path / "dropin_run")
    r = subprocess.run([gxx, "-std=c++14", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "dropin_run.cpp"), "-o", exe,
                        "-L", lib_dir, "-lspx", "-Wl,-rpath," + lib_dir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    rng = np.random.default_rng(71)
    m, p, n = 32, 64, 120
    D = W.unit_dict(rng, m, p, np.float64)
    X = W.sparse_signals(rng, D, n, 4, 0.05, True, np.float64)
    Wt = rng.uniform(0.5, 2.0, (p, n))
    B = (rng.uniform(size=(m, n)) < 0.8).astype(np.float64)
    B[:, 3] = 0.0                                    # a signal with no observed row
    L0 = float(np.linalg.eigvalsh(D.T @ D).max())
    np.asfortranarray(Wt).T.tofile(str(tmp_path / "W.bin"))
    np.asfortranarray(B).T.tofile(str(tmp_path / "B.bin"))
    np.array([L0]).tofile(str(tmp_path / "L0.bin"))
    (tmp_path / "dims.txt").write_text("%d %d %d" % (m, n, p))
    np.asfortranarray(X).T.tofile(str(tmp_path / "X.bin"))   # column-major bytes
    np.asfortranarray(D).T.tofile(str(tmp_path / "D.bin"))
    print("dropin_run: m=%d n=%d p=%d" % (m, n, p), flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # _lassoD and the Gram-given lasso2 against the lasso restatement
    refs = oracle.lasso(X, D, 0.15)
    ref = refs.toarray()
    got = read_csc_raw(str(tmp_path / "lassoD.bin"), p)
    assert np.abs(got.toarray() - ref).max() < 1e-9
    # the SpMatrix arrays themselves: pB and the sorted rows, bit for bit
    assert_csc_layout(got, refs.indptr, refs.indices)
    G = D.T @ D + 1e-10 * np.eye(p)
    refq = oracle.lasso_qq(X, G, D.T @ X, 0.15).toarray()
    assert np.abs(read_csc(str(tmp_path / "lasso2_gram.bin"), p) - refq).max() < 1e-9
    # lasso with ols: the same supports, least squares on each
    A = read_csc(str(tmp_path / "lasso_ols.bin"), p)
    for j in range(n):
        s = np.flatnonzero(ref[:, j])
        assert set(np.flatnonzero(A[:, j])) <= set(s)
        if len(s):
            ls = np.linalg.lstsq(D[:, s], X[:, j], rcond=None)[0]
            assert np.abs(A[s, j] - ls).max() < 1e-8
    # _omp (L = 8)
    refos = oracle.omp(X, D, L=8)
    refo = refos.toarray()
    got = read_csc_raw(str(tmp_path / "omp.bin"), p)
    assert np.abs(got.toarray() - refo).max() < 1e-9
    assert_csc_layout(got, refos.indptr, refos.indices)
    # _alltrainDL: 5 iterations of 64 signals from the given D, one thread's order
    Do, Ao, _, ito = oracle.train_dl(X, D=D, lambda1=0.15, lambda2=1e-9, batchsize=64, iter=5)
    Dl = np.fromfile(str(tmp_path / "traindl_D.bin"), np.float64).reshape(p, m).T
    Al = np.fromfile(str(tmp_path / "traindl_A.bin"), np.float64).reshape(p, p).T
    it = int(np.fromfile(str(tmp_path / "traindl_iter.bin"), np.int32)[0])
    assert it == ito == 5
    assert np.abs(Dl - Do).max() < 1e-10
    assert np.abs(Al - Ao).max() < 1e-10

    # weighted and masked decompositions
    refw = oracle.lasso_weighted(X, D, Wt, 0.1).toarray()
    assert np.abs(read_csc(str(tmp_path / "lasso_weighted.bin"), p) - refw).max() < 1e-9
    Bu = B.astype(np.uint8)
    refm = oracle.lasso_mask(X, D, Bu, 0.1).toarray()
    assert np.abs(read_csc(str(tmp_path / "lasso_mask.bin"), p) - refm).max() < 1e-9
    refom = oracle.omp_mask(X, D, Bu, L=8).toarray()
    assert np.abs(read_csc(str(tmp_path / "omp_mask.bin"), p) - refom).max() < 1e-9
    # FISTA (l1, fixed step L0 = ||D'D||, 60 iterations)
    Wf = np.fromfile(str(tmp_path / "fista_W.bin"), np.float64).reshape(n, p).T
    Wo, _ = oracle.fista_flat(X, D, np.zeros((p, n)), 0.05, regul="l1", max_it=60, L0=L0,
                              fixed_step=True, tol=0.0, it0=100)
    assert np.abs(Wf - Wo).max() < 1e-9
    # the dictionary given as an SpMatrix: the same result
    Ws = np.fromfile(str(tmp_path / "fista_sparseD_W.bin"), np.float64).reshape(n, p).T
    assert np.abs(Ws - Wo).max() < 1e-9
    # compute_gram with backtracking (L0 = 0.1): the Gram-form gradient and
    # backtracking test of the reference (prox/fista.h:364-413)
    Wg = np.fromfile(str(tmp_path / "fista_gram_W.bin"), np.float64).reshape(n, p).T
    Wgo, _ = oracle.fista_flat(X, D, np.zeros((p, n)), 0.05, regul="l1", max_it=60, L0=0.1,
                               fixed_step=False, tol=0.0, it0=100, compute_gram=True)
    assert np.abs(Wg - Wgo).max() < 1e-9
    # cppLasso (L = 10), cppOMP (L = 8), lasso2 with D (L = 10)
    refl = oracle.lasso(X, D, 0.15, L=10).toarray()
    assert np.abs(read_csc(str(tmp_path / "cpp_lasso.bin"), p) - refl).max() < 1e-9
    assert np.abs(read_csc(str(tmp_path / "lasso2_d.bin"), p) - refl).max() < 1e-9
    assert np.abs(read_csc(str(tmp_path / "cpp_omp.bin"), p) - refo).max() < 1e-9
