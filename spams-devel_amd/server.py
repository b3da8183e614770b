"""Served sparse coding: a continuous-batching request loop over the engine.

Jobs (LARS-Lasso with its own lambda, or ORMP with its own L) arrive over
time against one of several dictionaries kept resident in HBM (D, D' and the
Gram matrix).  The scheduler

  * admits a job when a lane (an engine context with its own HIP stream) is
    free and the job's device footprint fits the HBM budget still unclaimed;
  * fuses every queued job with the same dictionary and solver into the
    launch (up to `max_batch_signals`): the solvers take lambda / L per
    signal, so a fused launch computes exactly what separate launches would
    -- each signal's arithmetic does not depend on its neighbours;
  * runs lanes concurrently, so jobs that arrive while others are on the GPU
    overlap with them on the device.

A job whose footprint exceeds the whole budget is rejected with an error
(admission failure) instead of being queued forever.  Latency = completion
time - arrival time.

HBM accounting (the budget covers everything the served jobs hold beyond the
resident dictionaries):
  * each lane's engine scratch is counted while the lane holds it (it is
    kept between batches: the context's slots only grow), read back from the
    context after every batch, and released when it exceeds the lane's share
    of the budget;
  * a batch claims its device tensors plus the engine scratch it may add to
    its lane beyond what the lane holds (an upper bound of the solver's
    scratch slots, with their allocation slack);
  * device memory torch keeps cached but unused is counted too, and returned
    to the device when it exceeds a quarter of the budget.
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as ssp

import spx


@dataclass
class Job:
    arrival: float              # seconds from the start of the trace
    dict_id: int
    kind: str                   # "lasso" | "omp"
    X: np.ndarray               # m x n, Fortran order, float32
    param: float                # lambda (lasso) or L (omp)
    result: ssp.csc_matrix | None = None
    error: str | None = None
    t_admit: float = 0.0
    t_done: float = 0.0
    batch_id: int = -1

    @property
    def n(self) -> int:
        return self.X.shape[1]

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


@dataclass
class _Lane:
    ctx: object
    stream: object
    busy: bool = False
    thread: threading.Thread | None = None
    retained: int = 0           # engine scratch the lane's context holds (bytes)


class SparseCodingServer:
    def __init__(self, dicts, lanes=4, hbm_budget=8 << 30, max_batch_signals=65536,
                 device=0, fuse=True):
        import torch

        self.torch = torch
        self.dev = torch.device("cuda:%d" % device)
        self.dicts_np = dicts
        self.ctx0 = spx.Context(device)
        self.dicts = [spx.Dictionary(self.ctx0, D) for D in dicts]
        self.ctx0.synchronize()
        self.lanes = []
        for _ in range(lanes):
            s = torch.cuda.Stream(device=self.dev)  # default priority
            c = spx.Context(device)
            c.set_stream(s.cuda_stream)
            self.lanes.append(_Lane(c, s))
        self.budget = int(hbm_budget)
        self.in_use = 0             # claims of the batches on the GPU
        self.num_cus = torch.cuda.get_device_properties(self.dev).multi_processor_count
        # device memory in use once the dictionaries are resident: the served
        # jobs' usage is measured against it (peak_used, after every batch)
        torch.cuda.synchronize(self.dev)
        self.base_free = torch.cuda.mem_get_info(self.dev)[0]
        self.peak_used = 0
        self.max_batch = int(max_batch_signals)
        self.fuse = fuse
        self.lock = threading.Condition()
        self.n_batches = 0

    # ----------------------------------------------------------- autotune
    def autotune(self, sample_jobs, max_signals=16384):
        """Tunes the wavefronts-per-signal of each (dictionary, solver) pair
        present in `sample_jobs` on their signals, before serving.  The width
        is then fixed for every later batch on that pair, so a job's result
        does not depend on which batch it lands in.  Returns
        {(dict_id, kind): (waves, {width: ms})}."""
        torch = self.torch
        groups = {}
        for j in sample_jobs:
            groups.setdefault((j.dict_id, j.kind), []).append(j)
        out = {}
        for (di, kind), js in sorted(groups.items()):
            d = self.dicts[di]
            X = np.concatenate([j.X for j in js], axis=1)[:, :max_signals]
            n = X.shape[1]
            Xt = torch.from_numpy(np.ascontiguousarray(X.T)).to(self.dev)
            par = np.concatenate([np.full(j.n, j.param) for j in js])[:n]
            if kind == "lasso":
                lam = torch.from_numpy(par.astype(np.float32)).to(self.dev)
                out[(di, kind)] = d.autotune_lasso(Xt, lam)
            else:
                L = torch.from_numpy(par.astype(np.int32)).to(self.dev)
                out[(di, kind)] = d.autotune_omp(Xt, L=L)
        return out

    # ---------------------------------------------------------- footprint
    def _lout(self, dict_id: int, kind: str, L) -> int:
        """List length of a batch: min(m, p) for LARS (lasso_L), the largest
        job L (clamped) for ORMP."""
        d = self.dicts[dict_id]
        Lout = min(d.m, d.p)
        return Lout if kind == "lasso" or L is None else max(1, min(Lout, int(L)))

    def _tensor_bytes(self, dict_id: int, kind: str, n: int, L=None) -> int:
        """Device tensors of a batch of n signals: X, lambda / L, the (value,
        index) lists and counts, and the CSC (pB, r, v); 2 MiB of caching
        allocator rounding per tensor."""
        d = self.dicts[dict_id]
        Lout = self._lout(dict_id, kind, L)
        return n * (4 * d.m + 4 + 16 * Lout + 12) + 16 + 9 * (2 << 20)

    def _scratch_bytes(self, dict_id: int, kind: str, n: int, L=None) -> int:
        """Upper bound of the engine scratch a batch of n signals takes on its
        lane's context: D'X and norms; for LARS the work order, overflow list
        and keys, the sort buffers, the hand-over states (one per signal) and
        the inverse slots of the full-capacity pass (one per resident
        workgroup, at most n); for ORMP its inverse-factor slots when they do
        not fit LDS; the CSC scan.  Each slot is allocated 1/8 larger, and
        8 MiB cover allocation granularity."""
        d = self.dicts[dict_id]
        m, p = d.m, d.p
        ldg = (p + 15) // 16 * 16
        L = self._lout(dict_id, kind, L)
        b = n * ldg * 4 + n * 4                      # D'X, ||x||^2
        b += 8 * n + 8 * 1024 + 64                   # CSC count scan
        if kind == "lasso":
            state = (4 * (ldg + 24 + 24 * 24 + 2) + 4 * 28 + 15) // 16 * 16
            b += 16 * n + 16 * n + 4 * 256 * (n // 2048 + 1) + 4096   # order, lists, sort
            b += n * state                                           # hand-over states
            b += min(n, 4 * self.num_cus) * 2 * L * L * 4            # full-pass inverse slots
        else:
            if L * L * 4 > 32 * 1024:
                b += min(n, 16 * self.num_cus) * L * L * 4
            b += 64
        return b + b // 8 + (8 << 20)

    def footprint(self, dict_id: int, kind: str, n: int, L=None) -> int:
        """Device bytes a job of n signals (ORMP: at most L atoms) may need on
        a lane that holds no scratch (the admission test for a job alone)."""
        return (self._tensor_bytes(dict_id, kind, n, L) +
                self._scratch_bytes(dict_id, kind, n, L))

    def _claim(self, lane: _Lane, dict_id: int, kind: str, n: int, L=None) -> int:
        """A batch's claim on lane `lane`: its tensors plus the scratch it may
        add beyond what the lane already holds."""
        grow = max(0, self._scratch_bytes(dict_id, kind, n, L) - lane.retained)
        return self._tensor_bytes(dict_id, kind, n, L) + grow

    def _cached_unused(self) -> int:
        torch = self.torch
        return max(0, torch.cuda.memory_reserved(self.dev) - torch.cuda.memory_allocated(self.dev))

    def _held(self) -> int:
        """Bytes counted against the budget now (lock held)."""
        return self.in_use + sum(l.retained for l in self.lanes) + self._cached_unused()

    # --------------------------------------------------------------- run
    def _run_batch(self, lane: _Lane, jobs, claim: int):
        torch = self.torch
        d = self.dicts[jobs[0].dict_id]
        kind = jobs[0].kind
        m = d.m
        sizes = [j.n for j in jobs]
        N = sum(sizes)
        try:
            with torch.cuda.stream(lane.stream):
                Xh = np.concatenate([j.X for j in jobs], axis=1) if len(jobs) > 1 else jobs[0].X
                Xt = torch.from_numpy(np.ascontiguousarray(Xh.T)).pin_memory().to(
                    self.dev, non_blocking=True)
                if kind == "lasso":
                    lam = np.concatenate([np.full(j.n, j.param, np.float32) for j in jobs])
                    lam_t = torch.from_numpy(lam).to(self.dev, non_blocking=True)
                    Lout = spx.lasso_L(m, d.p, -1, spx.PENALTY, float(lam.max()))
                else:
                    Ls = np.concatenate([np.full(j.n, int(j.param), np.int32) for j in jobs])
                    L_t = torch.from_numpy(Ls).to(self.dev, non_blocking=True)
                    Lout = spx.omp_L(m, d.p, Ls, n=N)
                V = torch.empty((N, Lout), dtype=torch.float32, device=self.dev)
                R = torch.empty((N, Lout), dtype=torch.int32, device=self.dev)
                cnt = torch.empty((N,), dtype=torch.int32, device=self.dev)
                if kind == "lasso":
                    spx.lasso_dev(lane.ctx, d, Xt, lam_t, V, R, cnt)
                else:
                    spx.omp_dev(lane.ctx, d, Xt, V, R, cnt, L=L_t)
                pB = torch.empty((N + 1,), dtype=torch.int32, device=self.dev)
                r = torch.empty((N * Lout + 1,), dtype=torch.int32, device=self.dev)
                v = torch.empty((N * Lout + 1,), dtype=torch.float32, device=self.dev)
                nnz = spx.lists_to_csc_dev(lane.ctx, V, R, cnt, pB, r, v)
                pBh = pB.cpu().numpy()
                rh = r[:nnz].cpu().numpy()
                vh = v[:nnz].cpu().numpy()
            lane.ctx.synchronize()
            t = time.perf_counter() - self.t0
            c0 = 0
            for j in jobs:
                a, b = pBh[c0], pBh[c0 + j.n]
                j.result = ssp.csc_matrix((vh[a:b], rh[a:b], pBh[c0:c0 + j.n + 1] - a),
                                          shape=(d.p, j.n))
                j.t_done = t
                c0 += j.n
        except Exception as e:  # per-job error propagation
            t = time.perf_counter() - self.t0
            for j in jobs:
                j.error = str(e)
                j.t_done = t
        finally:
            torch.cuda.synchronize(self.dev)
            used = self.base_free - torch.cuda.mem_get_info(self.dev)[0]
            held = lane.ctx.scratch_bytes()
            if held > self.budget // max(len(self.lanes), 1):
                lane.ctx.release_scratch()
                held = 0
            with self.lock:
                self.peak_used = max(self.peak_used, used)
                self.in_use -= claim
                lane.retained = held
                if self._cached_unused() > self.budget // 4:
                    torch.cuda.empty_cache()
                lane.busy = False
                self.lock.notify_all()

    def run_trace(self, jobs, realtime=True):
        """Replays `jobs` (sorted by arrival) and returns them completed."""
        jobs = sorted(jobs, key=lambda j: j.arrival)
        pending = deque()
        nxt = 0
        self.t0 = time.perf_counter()
        threads = []
        done = 0
        while done < len(jobs):
            now = time.perf_counter() - self.t0
            while nxt < len(jobs) and (not realtime or jobs[nxt].arrival <= now):
                j = jobs[nxt]
                if self.footprint(j.dict_id, j.kind, j.n, j.param) > self.budget:
                    j.error = "admission: job footprint exceeds the HBM budget"
                    j.t_done = now
                    done += 1
                else:
                    pending.append(j)
                nxt += 1
            launched = False
            with self.lock:
                # the free lane holding the most scratch (its claim grows least)
                free = [l for l in self.lanes if not l.busy]
                lane = max(free, key=lambda l: l.retained) if free else None
                if lane is not None and pending:
                    head = pending[0]
                    held = self._held()
                    claim = self._claim(lane, head.dict_id, head.kind, head.n, head.param)
                    if held + claim > self.budget and not any(l.busy for l in self.lanes):
                        # an idle server whose retained scratch or cached
                        # tensors block the head job gives them back first
                        for l in self.lanes:
                            l.ctx.release_scratch()
                            l.retained = 0
                        self.torch.cuda.empty_cache()
                        held = self._held()
                        claim = self._claim(lane, head.dict_id, head.kind, head.n, head.param)
                    if held + claim <= self.budget:
                        batch = [pending.popleft()]
                        tot = head.n
                        Lb = head.param
                        if self.fuse:
                            rest = deque()
                            while pending:
                                j = pending.popleft()
                                L2 = max(Lb, j.param)
                                c2 = self._claim(lane, head.dict_id, head.kind, tot + j.n, L2)
                                if (j.dict_id == head.dict_id and j.kind == head.kind
                                        and tot + j.n <= self.max_batch
                                        and held + c2 <= self.budget):
                                    batch.append(j)
                                    tot += j.n
                                    Lb = L2
                                    claim = c2
                                else:
                                    rest.append(j)
                            pending = rest
                        self.in_use += claim
                        for j in batch:
                            j.t_admit = now
                            j.batch_id = self.n_batches
                        self.n_batches += 1
                        lane.busy = True
                        th = threading.Thread(target=self._run_batch, args=(lane, batch, claim))
                        th.start()
                        threads.append((th, batch))
                        launched = True
            # reap finished batches
            alive = []
            for th, b in threads:
                if th.is_alive():
                    alive.append((th, b))
                else:
                    th.join()
                    done += len(b)
            threads = alive
            if not launched:
                with self.lock:
                    self.lock.wait(timeout=0.0005)
        return jobs

    def close(self):
        for l in self.lanes:
            l.ctx.synchronize()
            l.ctx.close()
        for d in self.dicts:
            d.close()
        self.ctx0.close()


def latency_report(jobs):
    ok = [j for j in jobs if j.error is None]
    lat = np.array([j.latency for j in ok]) * 1e3
    sig = sum(j.n for j in ok)
    span = max(j.t_done for j in jobs) - min(j.arrival for j in jobs)
    return {"jobs": len(jobs), "failed": len(jobs) - len(ok), "signals": int(sig),
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "throughput_signals_per_s": sig / span if span > 0 else None,
            "makespan_s": span}
