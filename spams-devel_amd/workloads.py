"""Seeded synthetic workloads of the decomposition path.

Each generator returns Fortran-ordered numpy arrays (column-major, like the
reference's Matrix<T>).  Shapes and recipes follow the benchmark
configurations of this engine:

  lars_small    D 64x256, X 64x10000, 5-sparse codes + 0.05 noise,
                columns normalised; lambda 0.15                  (CPU case)
  omp_batch     D 256x1024, X 256x100000, k ~ U{1..32} codes + 0.01 noise;
                L = 32, eps = 1e-3
  lars_headline D 256x1024, X 256x200000 as lars_small, lambda per signal
                log-uniform in [0.05, 0.5]
  fista_flat    D 512x2048, X 512x8192 from 20-sparse codes + 0.02 noise
  served        4 resident dictionaries, Poisson arrivals of LARS / OMP jobs
"""
from __future__ import annotations

import numpy as np


def unit_dict(rng, m, p, dtype=np.float32):
    D = rng.standard_normal((m, p))
    D /= np.linalg.norm(D, axis=0, keepdims=True)
    return np.asfortranarray(D.astype(dtype))


def sparse_signals(rng, D, n, k, noise, normalize, dtype=np.float32,
                   chunk=20000):
    """X = D A + noise * N(0,1); A has k nonzeros per column (k int or an
    array of per-column counts) at uniform random rows with N(0,1) values."""
    m, p = D.shape
    Dd = D.astype(np.float64)
    X = np.empty((m, n), dtype=dtype, order="F")
    ks = np.broadcast_to(np.asarray(k), (n,))
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        nc = c1 - c0
        A = np.zeros((p, nc))
        kk = ks[c0:c1]
        kmax = int(kk.max()) if nc else 0
        # random supports without replacement: argsort of uniforms
        if kmax:
            sup = np.argsort(rng.random((nc, p)), axis=1)[:, :kmax]
            vals = rng.standard_normal((nc, kmax))
            mask = np.arange(kmax)[None, :] < kk[:, None]
            cols = np.repeat(np.arange(nc)[:, None], kmax, axis=1)
            A[sup[mask], cols[mask]] = vals[mask]
        Xc = Dd @ A + noise * rng.standard_normal((m, nc))
        if normalize:
            Xc /= np.linalg.norm(Xc, axis=0, keepdims=True)
        X[:, c0:c1] = Xc
    return X


def lars_small(seed=1001, n=10000, dtype=np.float32):
    rng = np.random.default_rng(seed)
    D = unit_dict(rng, 64, 256, dtype)
    X = sparse_signals(rng, D, n, 5, 0.05, True, dtype)
    return X, D, 0.15


def omp_batch(seed=1002, n=100000, dtype=np.float32):
    rng = np.random.default_rng(seed)
    D = unit_dict(rng, 256, 1024, dtype)
    k = rng.integers(1, 33, size=n)
    X = sparse_signals(rng, D, n, k, 0.01, False, dtype)
    return X, D, 32, 1e-3


def lars_headline(seed=1003, n=200000, dtype=np.float32, m=256, p=1024):
    rng = np.random.default_rng(seed)
    D = unit_dict(rng, m, p, dtype)
    X = sparse_signals(rng, D, n, 5, 0.05, True, dtype)
    lam = np.exp(rng.uniform(np.log(0.05), np.log(0.5), size=n)).astype(dtype)
    return X, D, lam


def fista_flat(seed=1004, n=8192, dtype=np.float32, m=512, p=2048):
    rng = np.random.default_rng(seed)
    D = unit_dict(rng, m, p, dtype)
    X = sparse_signals(rng, D, n, 20, 0.02, False, dtype)
    return X, D


SERVED_DICTS = ((64, 256), (128, 512), (256, 1024), (192, 512))


def served_jobs(seed=1005, n_jobs=200, rate=200.0, max_n=20000,
                dtype=np.float32):
    """Dictionaries and a Poisson job trace.  Each job: (arrival_s, dict_id,
    kind 'lasso'|'omp', X, param) where param is lambda (lasso) or L (omp)."""
    rng = np.random.default_rng(seed)
    dicts = [unit_dict(rng, m, p, dtype) for (m, p) in SERVED_DICTS]
    t = 0.0
    jobs = []
    for _ in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        di = int(rng.integers(0, len(dicts)))
        n = max(1, int(np.exp(rng.uniform(0.0, np.log(max_n)))))
        D = dicts[di]
        X = sparse_signals(rng, D, n, 5, 0.05, True, dtype)
        if rng.random() < 0.5:
            lam = float(np.exp(rng.uniform(np.log(0.05), np.log(0.5))))
            jobs.append((t, di, "lasso", X, lam))
        else:
            jobs.append((t, di, "omp", X, int(rng.integers(4, 33))))
    return dicts, jobs
