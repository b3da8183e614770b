"""Served form: continuous batching over resident dictionaries.

Each job's result must be the same when it runs alone and when it is fused
with other jobs in the mix (the solvers take lambda / L per signal, and each
signal's arithmetic does not depend on its neighbours), and must match the
oracle within the FP32 tolerances of test_gpu_parity.py.
"""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def _jobs(seed, n_jobs, max_n):
    dicts, trace = W.served_jobs(seed=seed, n_jobs=n_jobs, rate=2000.0, max_n=max_n)
    from server import Job
    return dicts, [Job(t, di, kind, X, prm) for (t, di, kind, X, prm) in trace]


def test_alone_equals_in_the_mix():
    import server
    print("case server alone-vs-mix", flush=True)
    dicts, jobs = _jobs(11, 24, 400)
    srv = server.SparseCodingServer(dicts, lanes=3, hbm_budget=2 << 30)
    try:
        mixed = srv.run_trace(jobs, realtime=False)
        assert all(j.error is None for j in mixed)
        assert srv.n_batches < len(jobs)  # fusion happened
        mix_res = {id(j): j.result for j in mixed}
        for j in mixed:
            from server import Job
            alone = Job(0.0, j.dict_id, j.kind, j.X, j.param)
            srv2 = srv
            srv2.fuse = False
            out = srv2.run_trace([alone], realtime=False)[0]
            srv2.fuse = True
            assert out.error is None
            a = out.result.toarray()
            b = mix_res[id(j)].toarray()
            assert np.array_equal(a != 0, b != 0)
            assert np.array_equal(a, b), "results differ alone vs in the mix"
    finally:
        srv.close()


def test_served_results_match_oracle(orc):
    import server
    print("case server vs oracle", flush=True)
    dicts, jobs = _jobs(12, 10, 200)
    srv = server.SparseCodingServer(dicts, lanes=2, hbm_budget=2 << 30)
    try:
        done = srv.run_trace(jobs, realtime=True)
    finally:
        srv.close()
    for j in done:
        assert j.error is None
        D = dicts[j.dict_id]
        if j.kind == "lasso":
            ref = orc.lasso(j.X.astype(np.float64), D.astype(np.float64), j.param).toarray()
        else:
            ref = orc.omp(j.X.astype(np.float64), D.astype(np.float64), L=int(j.param)).toarray()
        got = j.result.toarray()
        same = np.all((got != 0) == (ref != 0), axis=0)
        assert same.mean() >= 0.95
        assert np.abs(got[:, same] - ref[:, same]).max(initial=0) <= 5e-4


def test_admission_rejects_oversized_job():
    import server
    from server import Job
    print("case server admission", flush=True)
    dicts, jobs = _jobs(13, 3, 50)
    srv = server.SparseCodingServer(dicts, lanes=1, hbm_budget=4 << 20)
    try:
        big = Job(0.0, 2, "lasso", np.asfortranarray(np.ones((256, 5000), np.float32)), 0.1)
        out = srv.run_trace([big], realtime=False)[0]
        assert out.error is not None and "budget" in out.error
    finally:
        srv.close()


def test_hbm_budget_holds_with_large_jobs_on_several_lanes():
    """Large LARS and OMP jobs on three lanes under a 1.25 GiB budget: every
    job completes, and the device memory in use beyond the resident
    dictionaries (read after every batch, and at the end) stays within the
    budget, lane scratch and torch's cache included."""
    import torch

    import server
    from server import Job
    print("case server hbm budget", flush=True)
    dicts, _ = _jobs(14, 2, 10)
    rng = np.random.default_rng(15)
    D = dicts[2]                                   # (256, 1024)
    jobs = []
    for k in range(6):
        X = W.sparse_signals(rng, D, 12000, 5, 0.05, True, np.float32)
        kind = "lasso" if k % 2 == 0 else "omp"
        jobs.append(Job(0.0, 2, kind, X, 0.05 if kind == "lasso" else 24))
    budget = 5 << 28
    srv = server.SparseCodingServer(dicts, lanes=3, hbm_budget=budget)
    try:
        assert all(srv.footprint(2, j.kind, j.n, j.param) <= budget for j in jobs)
        done = srv.run_trace(jobs, realtime=False)
        assert all(j.error is None for j in done), [j.error for j in done]
        torch.cuda.synchronize()
        end_used = srv.base_free - torch.cuda.mem_get_info()[0]
        print("  peak %.0f MB, end %.0f MB, budget %.0f MB"
              % (srv.peak_used / 1e6, end_used / 1e6, budget / 1e6), flush=True)
        assert srv.peak_used <= budget
        assert end_used <= budget
    finally:
        srv.close()
