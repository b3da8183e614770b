"""GPU parity of dictionary learning (trainDL, Trainer::train of
dictLearn/dicts.h:346-759) against the reference's fixtures and the oracle,
called through the C ABI.

Tolerances (stated here, used below):
  * FP64: D, A and B within 1e-10 (absolute; entries are O(1)) of the
    reference's run and the same iteration count.  The engine sums in a
    different order (MFMA GEMMs for G, D'X, A_t and B_t; a blocked
    dictionary update), so bit equality is not expected.
  * FP32: D within 2e-4, A and B within 5e-5 of the reference's FP32 run
    (one homotopy per signal and iteration: a near-tie can still branch
    differently, which these small gap-separated cases do not hit).
  * The m=256, K=1024 dictionary (the headline decomposition size) in FP64:
    within 1e-9 of the oracle after three iterations.
"""
import numpy as np
import pytest

from conftest import golden_files, load_golden

import workloads as W

pytestmark = pytest.mark.gpu

PRM = ("batchsize", "lambda1", "lambda2", "iter", "t0", "mode", "posAlpha", "posD", "clean",
       "rho", "iter_updateD", "modeParam", "batch")


@pytest.fixture(scope="module")
def spx():
    import spx as mod
    print("spx:", mod.lib().spx_version(), flush=True)
    return mod


def _tol(dt):
    return (1e-10, 1e-10) if dt == np.float64 else (2e-4, 5e-5)


@pytest.mark.parametrize("name", golden_files("traindl_"))
def test_train_dl_matches_golden(spx, name):
    g = load_golden(name)
    prm = {k: g[k].item() for k in PRM}
    X = g["X"]
    dt = X.dtype.type
    tD, tAB = _tol(dt)
    print("trainDL fixture", name, flush=True)
    if "D0" in g:
        D1, mdl = spx.trainDL(X, D=g["D0"], return_model=True, **dict(prm, iter=7))
        assert mdl["iter"] == int(g["it1"])
        assert np.abs(D1 - g["D1"]).max() < tD
        D, mdl = spx.trainDL(X, D=D1, model=mdl, return_model=True, **prm)
    else:
        D0 = g["Dinit"] if "Dinit" in g else None
        D, mdl = spx.trainDL(X, D=D0, K=int(g["K"]), return_model=True, **prm)
    assert mdl["iter"] == int(g["itercount"])
    assert D.dtype == dt and D.shape == g["D"].shape
    assert np.abs(D - g["D"]).max() < tD
    assert np.abs(mdl["A"] - g["A"]).max() < tAB
    assert np.abs(mdl["B"] - g["B"]).max() < tAB


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_train_dl_device_entry_equals_host_entry(spx, dt):
    import torch
    rng = np.random.default_rng(51)
    D0 = W.unit_dict(rng, 32, 64, dt)
    X = W.sparse_signals(rng, D0, 700, 4, 0.05, True, dt)
    kw = dict(lambda1=0.12, batchsize=128, iter=9, t0=1e-5)
    Dh, mh = spx.trainDL(X, D=D0, return_model=True, **kw)
    ctx = spx.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    tdt = torch.float32 if dt == np.float32 else torch.float64

    def colmajor(a):
        return torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda", tdt).T

    Xt, Dt = colmajor(X), colmajor(D0)
    At = colmajor(np.zeros((64, 64), dt))
    Bt = colmajor(np.zeros((32, 64), dt))
    it = spx.train_dl_dev(ctx, Xt, Dt, At, Bt, itercount=0, D_given=True, **kw)
    torch.cuda.synchronize()
    assert it == mh["iter"] == 9
    assert np.array_equal(Dt.cpu().numpy(), Dh)
    assert np.array_equal(At.cpu().numpy(), mh["A"])
    assert np.array_equal(Bt.cpu().numpy(), mh["B"])


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_train_dl_headline_dictionary_matches_oracle(spx, dt):
    """m=256, K=1024 from a given dictionary (an initial draw of 1024 atoms
    from a few thousand signals would duplicate atoms, whose exact ties no
    two summation orders break alike): the dictionary update runs in many
    blocks here.  FP32 within 1e-4 of the FP32 oracle."""
    import oracle
    rng = np.random.default_rng(1003)
    Dtrue = W.unit_dict(rng, 256, 1024, dt)
    X = W.sparse_signals(rng, Dtrue, 3000, 5, 0.05, True, dt)
    D0 = W.unit_dict(np.random.default_rng(1004), 256, 1024, dt)
    kw = dict(lambda1=0.15, batchsize=256, iter=3)
    print("trainDL m=256 K=1024", dt.__name__, flush=True)
    D, mdl = spx.trainDL(X, D=D0, return_model=True, **kw)
    Do, Ao, Bo, ito = oracle.train_dl(X, D=D0, **kw)
    assert mdl["iter"] == ito == 3
    err = [np.abs(D - Do).max(), np.abs(mdl["A"] - Ao).max(), np.abs(mdl["B"] - Bo).max()]
    print("max |diff| D, A, B:", err, flush=True)
    tol = 1e-9 if dt == np.float64 else 1e-4
    assert max(err) < tol


def test_train_dl_bf16x3_stays_close_to_fp32(spx):
    """BF16x3 products (G, D'X, A_t, B_t) give a dictionary of the same
    quality: the lasso objective of the signals under it (FP64 oracle)
    within 1e-3 relative of the FP32 run's."""
    import oracle
    rng = np.random.default_rng(52)
    Dtrue = W.unit_dict(rng, 64, 256, np.float32)
    X = W.sparse_signals(rng, Dtrue, 4000, 5, 0.05, True, np.float32)
    kw = dict(lambda1=0.15, K=256, batchsize=512, iter=10)
    objs = []
    for prec in ("fp32", "bf16x3"):
        D = spx.trainDL(X, precision=prec, **kw)
        X64, D64 = X.astype(np.float64), D.astype(np.float64)
        A = oracle.lasso(X64, D64, 0.15).toarray()
        R = X64 - D64 @ A
        objs.append(0.5 * np.sum(R * R) + 0.15 * np.abs(A).sum())
    assert abs(objs[1] - objs[0]) <= 1e-3 * objs[0]


def test_train_dl_time_limit_and_errors(spx):
    rng = np.random.default_rng(53)
    Dtrue = W.unit_dict(rng, 16, 32, np.float32)
    X = W.sparse_signals(rng, Dtrue, 300, 3, 0.05, True, np.float32)
    # iter < 0: train for -iter seconds (dicts.h:517-518)
    D, mdl = spx.trainDL(X, K=32, lambda1=0.15, batchsize=64, iter=-1, return_model=True)
    assert mdl["iter"] >= 1
    assert np.all(np.linalg.norm(D, axis=0) <= 1 + 1e-5)
    # the ridge-only problem (conjugate-gradient branch of the reference) is refused
    with pytest.raises(RuntimeError):
        spx.trainDL(X, K=32, lambda1=0.0, lambda2=0.1, iter=2)
    # a model without its dictionary is refused by the C entry point
    import ctypes
    o = spx.spx_train_dl_opts(64, 1, 0.15, 1e-9, 2, 1e-5, 2, 0, 0, 1, 1.0, 1, 0, 0, 0, 0, 0)
    Dn = np.zeros((16, 32), np.float32, order="F")
    An = np.zeros((32, 32), np.float32, order="F")
    Bn = np.zeros((16, 32), np.float32, order="F")
    it = ctypes.c_int(3)
    f = spx.lib().spx_train_dl_f32
    f.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int] * 2 + [ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                  ctypes.POINTER(spx.spx_train_dl_opts)]
    Xf = np.asfortranarray(X)
    rc = f(spx.default_context().h, Xf.ctypes.data, 16, 300, Dn.ctypes.data, 32, 0,
           An.ctypes.data, Bn.ctypes.data, ctypes.byref(it), ctypes.byref(o))
    assert rc != 0 and b"dictionary" in spx.lib().spx_last_error()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_train_dl_graph_replays_equal_eager_launches(spx, dt):
    """The dictionary-update sweeps replayed as a captured hipGraph (many
    replays in one call, with the solver's eager work between them) give
    the eager launches' dictionary bit for bit, in batch mode (5 sweeps per
    replay) and online."""
    rng = np.random.default_rng(54)
    D0 = W.unit_dict(rng, 64, 200, dt)
    X = W.sparse_signals(rng, D0, 1500, 5, 0.05, True, dt)
    for kw in (dict(batchsize=256, iter=12), dict(batch=True, iter=3, iter_updateD=5)):
        Dg, mg = spx.trainDL(X, D=D0, lambda1=0.12, return_model=True, **kw)
        De, me = spx.trainDL(X, D=D0, lambda1=0.12, return_model=True, eager_update=True, **kw)
        assert np.array_equal(Dg, De)
        assert np.array_equal(mg["A"], me["A"]) and np.array_equal(mg["B"], me["B"])


def test_train_dl_long_signals_fp64_matches_oracle(spx):
    """m = 520: each thread of the dictionary update's block walk owns three
    rows (every other test has one or two); same bar as the FP64 fixtures."""
    import oracle
    rng = np.random.default_rng(55)
    Dtrue = W.unit_dict(rng, 520, 48, np.float64)
    X = W.sparse_signals(rng, Dtrue, 400, 4, 0.05, True, np.float64)
    D0 = W.unit_dict(np.random.default_rng(56), 520, 48, np.float64)
    kw = dict(lambda1=0.12, batchsize=64, iter=6)
    D, mdl = spx.trainDL(X, D=D0, return_model=True, **kw)
    Do, Ao, Bo, ito = oracle.train_dl(X, D=D0, **kw)
    assert mdl["iter"] == ito == 6
    assert max(np.abs(D - Do).max(), np.abs(mdl["A"] - Ao).max(),
               np.abs(mdl["B"] - Bo).max()) < 1e-10


def test_train_dl_batch_mode_over_several_chunks_fp64(spx):
    """Batch mode with 20000 signals: each iteration codes them in two
    solver chunks (16384 + 3616) whose statistics accumulate through the
    ping-pong buffers; FP64 against the oracle."""
    import oracle
    rng = np.random.default_rng(57)
    Dtrue = W.unit_dict(rng, 32, 64, np.float64)
    X = W.sparse_signals(rng, Dtrue, 20000, 3, 0.05, True, np.float64)
    D0 = W.unit_dict(np.random.default_rng(58), 32, 64, np.float64)
    kw = dict(lambda1=0.1, batch=True, iter=2, iter_updateD=2)
    D, mdl = spx.trainDL(X, D=D0, return_model=True, **kw)
    Do, Ao, Bo, ito = oracle.train_dl(X, D=D0, **kw)
    assert mdl["iter"] == ito == 2
    # A and B sum 20000 codes: compare relative to their scale
    assert np.abs(mdl["A"] - Ao).max() < 1e-10 * max(1.0, np.abs(Ao).max())
    assert np.abs(mdl["B"] - Bo).max() < 1e-10 * max(1.0, np.abs(Bo).max())
    assert np.abs(D - Do).max() < 1e-10


@pytest.mark.parametrize("case", ["iter0", "small_n"])
def test_train_dl_edge_cases_fp64(spx, case):
    """iter = 0 returns the normalised start and the fresh online model
    (A = t0 I, B = t0 D); fewer signals than the batch size codes all of
    them each iteration (the permutation wraps)."""
    import oracle
    rng = np.random.default_rng(59)
    Dtrue = W.unit_dict(rng, 12, 20, np.float64)
    n = 300 if case == "iter0" else 50
    X = W.sparse_signals(rng, Dtrue, n, 3, 0.05, True, np.float64)
    D0 = 3.0 * W.unit_dict(np.random.default_rng(60), 12, 20, np.float64)  # not unit norm
    kw = dict(lambda1=0.12, batchsize=64, iter=0 if case == "iter0" else 5)
    D, mdl = spx.trainDL(X, D=D0, return_model=True, **kw)
    Do, Ao, Bo, ito = oracle.train_dl(X, D=D0, **kw)
    assert mdl["iter"] == ito
    assert max(np.abs(D - Do).max(), np.abs(mdl["A"] - Ao).max(),
               np.abs(mdl["B"] - Bo).max()) < 1e-10
    if case == "iter0":
        assert np.allclose(np.linalg.norm(D, axis=0), 1.0)
        assert np.allclose(mdl["A"], 1e-5 * np.eye(20))


@pytest.mark.parametrize("which", ["nmf", "nnsc"])
def test_nmf_nnsc_match_oracle_fp64(spx, which):
    """nmf and nnsc (swig/python/spams.py:530-562): trainDL with nonnegative
    codes and dictionary (lambda1 = 0 for nmf), then the codes by lasso
    (nonnegative for nmf, unconstrained for nnsc, as the reference calls
    it); FP64 against the oracle's trainer and lasso."""
    import oracle
    rng = np.random.default_rng(61)
    U0 = np.abs(rng.standard_normal((16, 24)))
    V0 = np.abs(rng.standard_normal((24, 300))) * (rng.random((24, 300)) < 0.2)
    X = np.asfortranarray(U0 @ V0 + 0.01 * np.abs(rng.standard_normal((16, 300))))
    lam = 0.0 if which == "nmf" else 0.05
    kw = dict(K=24, batchsize=64, iter=8)
    if which == "nmf":
        U, V = spx.nmf(X, return_lasso=True, **kw)
    else:
        U, V = spx.nnsc(X, return_lasso=True, lambda1=lam, **kw)
    Uo, _, _, _ = oracle.train_dl(X, lambda1=lam, lambda2=10e-10, posAlpha=True, posD=True, **kw)
    assert np.abs(U - Uo).max() < 1e-10
    assert U.min() >= 0
    Vo = oracle.lasso(X, Uo, lam, pos=(which == "nmf")).toarray()
    assert np.abs(V.toarray() - Vo).max() < 1e-9
    if which == "nmf":
        assert V.toarray().min() >= 0


@pytest.mark.parametrize("mode,lam", [(3, 5), (4, 0.05)])
def test_train_dl_ormp_modes_mid_size_fp64(spx, mode, lam):
    """SPARSITY and L2ERROR2 coding (coreORMPB over the trainer's ridged Gram)
    at m=64, K=256 from a given dictionary; FP64 against the oracle."""
    import oracle
    rng = np.random.default_rng(62)
    Dtrue = W.unit_dict(rng, 64, 256, np.float64)
    X = W.sparse_signals(rng, Dtrue, 2000, 5, 0.05, True, np.float64)
    D0 = W.unit_dict(np.random.default_rng(63), 64, 256, np.float64)
    kw = dict(lambda1=lam, mode=mode, batchsize=512, iter=3)
    D, mdl = spx.trainDL(X, D=D0, return_model=True, **kw)
    Do, Ao, Bo, ito = oracle.train_dl(X, D=D0, **kw)
    assert mdl["iter"] == ito == 3
    assert max(np.abs(D - Do).max(), np.abs(mdl["A"] - Ao).max(),
               np.abs(mdl["B"] - Bo).max()) < 1e-9


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_train_dl_unused_atoms_between_used_ones(spx, dt):
    """Batch mode with a dictionary whose odd atoms span the orthogonal
    complement of the signals (never selected: A_kk = 0, so the update skips
    them and their step has no norm reduction), interleaved with used even
    atoms: the update's block walk alternates steps with and without a
    barrier.  FP64 within 1e-10 of the oracle, FP32 at the module's bar."""
    import oracle
    rng = np.random.default_rng(71)
    m, K, n = 64, 64, 600
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    E = Q[:, :24] @ rng.standard_normal((24, K // 2))
    O = Q[:, 24:48] @ rng.standard_normal((24, K // 2))
    D0 = np.empty((m, K))
    D0[:, 0::2] = E / np.linalg.norm(E, axis=0)
    D0[:, 1::2] = O / np.linalg.norm(O, axis=0)
    codes = np.zeros((K // 2, n))
    for j in range(n):
        codes[rng.choice(K // 2, 3, replace=False), j] = rng.standard_normal(3)
    X = D0[:, 0::2] @ codes
    X /= np.linalg.norm(X, axis=0)
    X = np.asfortranarray(X.astype(dt))
    D0 = np.asfortranarray(D0.astype(dt))
    # batch mode: the binding's default of 5 update sweeps, given to both
    kw = dict(lambda1=0.05, batchsize=n, iter=3, batch=True, clean=False, iter_updateD=5)
    D, mdl = spx.trainDL(X, D=D0, return_model=True, **kw)
    Do, Ao, Bo, ito = oracle.train_dl(X, D=D0, **kw)
    err = [np.abs(D - Do).max(), np.abs(mdl["A"] - Ao).max(), np.abs(mdl["B"] - Bo).max()]
    print("max |diff| D, A, B:", err, flush=True)
    assert mdl["iter"] == ito
    assert np.all(np.diag(Ao)[1::2] == 0)
    td, tab = _tol(dt)
    assert err[0] < td and max(err[1:]) < tab, err
