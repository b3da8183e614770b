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
        self.in_use = 0
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
    def footprint(self, dict_id: int, kind: str, n: int) -> int:
        d = self.dicts[dict_id]
        m, p = d.m, d.p
        ldg = (p + 15) // 16 * 16
        L = min(m, p)
        lists = n * L * 8 + n * 4
        per = n * m * 4 + n * ldg * 4 + n * 4 + lists + n * 4
        # solver scratch: one slot per resident workgroup, bounded by n
        slot = (2 if kind == "lasso" else 1) * L * L * 4
        wg = min(n, 256 * 16)
        # LARS hand-over state between its two passes: one slot per signal
        handover = n * (4 * (ldg + 24 + 24 * 24 + 2) + 4 * 28) if kind == "lasso" else 0
        return per + wg * slot + handover + (1 << 20)

    # --------------------------------------------------------------- run
    def _run_batch(self, lane: _Lane, jobs):
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
            with self.lock:
                self.in_use -= sum(self.footprint(j.dict_id, j.kind, j.n) for j in jobs)
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
                if self.footprint(j.dict_id, j.kind, j.n) > self.budget:
                    j.error = "admission: job footprint exceeds the HBM budget"
                    j.t_done = now
                    done += 1
                else:
                    pending.append(j)
                nxt += 1
            launched = False
            with self.lock:
                lane = next((l for l in self.lanes if not l.busy), None)
                if lane is not None and pending:
                    head = pending[0]
                    fp = self.footprint(head.dict_id, head.kind, head.n)
                    if self.in_use + fp <= self.budget:
                        batch = [pending.popleft()]
                        self.in_use += fp
                        if self.fuse:
                            rest = deque()
                            tot = head.n
                            while pending:
                                j = pending.popleft()
                                f = self.footprint(j.dict_id, j.kind, j.n)
                                if (j.dict_id == head.dict_id and j.kind == head.kind
                                        and tot + j.n <= self.max_batch
                                        and self.in_use + f <= self.budget):
                                    batch.append(j)
                                    tot += j.n
                                    self.in_use += f
                                else:
                                    rest.append(j)
                            pending = rest
                        for j in batch:
                            j.t_admit = now
                            j.batch_id = self.n_batches
                        self.n_batches += 1
                        lane.busy = True
                        th = threading.Thread(target=self._run_batch, args=(lane, batch))
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
