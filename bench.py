"""Benchmark of the headline path: LARS-Lasso, m=256, p=1024, n=200000
signals, lambda log-uniform in [0.05, 0.5] per signal, FP32, on one MI355X.

One step = one pass of the path over the whole batch with inputs resident in
HBM: D'X on MFMA, ||x||^2, then the LARS solver kernel producing the
per-signal (value, index) lists.  `value` is the batched (persistent
work-queue) form's signals/s; the plain form (one workgroup per signal) is
measured beside it.  Kernel durations come from HIP events recorded on the
engine's stream; roofline figures use the algorithmic work of each kernel.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "spams-devel_amd"))

METRIC = "signals/sec LARS-Lasso m256 p1024 served vs plain GPU; p50/p99 request latency"
HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
FP32_MFMA_PEAK = 157.3e12  # FLOP/s, dense FP32 matrix


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(X, D, lam, seconds):
    """The reference's OpenMP LARS (compiled in place, oracle/_ref) on a
    bounded sample of the same workload; falls back to the C restatement."""
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    kind = "reference" if oracle.has_ref() else "port"
    run = (lambda Xs, l: oracle.ref_lasso_pervec(Xs, D, l, num_threads=threads)) \
        if kind == "reference" else (lambda Xs, l: oracle.lasso(Xs, D, l, num_threads=threads))
    # calibrate on growing prefixes, then time a sample of about `seconds`
    ns, dt = 256, 0.0
    while True:
        t = time.perf_counter()
        run(np.asfortranarray(X[:, :ns]), lam[:ns])
        dt = time.perf_counter() - t
        if dt >= 0.25 * seconds or ns >= X.shape[1]:
            break
        ns = int(min(X.shape[1], ns * max(2.0, min(16.0, 0.5 * seconds / max(dt, 1e-3)))))
    if dt < 0.5 * seconds and ns < X.shape[1]:
        ns = int(min(X.shape[1], ns * seconds / dt))
        t = time.perf_counter()
        run(np.asfortranarray(X[:, :ns]), lam[:ns])
        dt = time.perf_counter() - t
    return {"value": ns / dt, "unit": "signals/s", "cores": threads, "kind": kind,
            "sample": f"first {ns} signals of the same workload, {dt:.1f} s, "
                      f"per-signal-lambda lasso2/coreLARS2 loop, {threads} OpenMP threads"}


def measured_traffic(prefix):
    """HBM bytes per dispatch of the kernel whose name starts with `prefix`,
    from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this
    workload (null when absent)."""
    path = os.path.join(ROOT, "profiles", "headline_counters.json")
    try:
        with open(path) as f:
            d = json.load(f)["bytes_per_dispatch"]
    except (OSError, ValueError, KeyError):
        return None
    for k, v in d.items():
        if k.startswith(prefix):
            return v
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=200000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 256 MiB read + 256 MiB write copy kernel first "
                         "(calibrates HBM byte counters under a profiler)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported "
              "(signals shard across GPUs with no exchange, see the design notes)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    import spx
    import workloads as W

    log("bench: generating workload n=%d" % args.n)
    X, D, lam = W.lars_headline(n=args.n)
    m, n = X.shape
    p = D.shape[1]
    dev = torch.device("cuda:0")
    Xt = torch.from_numpy(X.T).to(dev)            # column-major m x n
    lam_t = torch.from_numpy(lam).to(dev)
    stream = torch.cuda.current_stream(dev)
    ctx = spx.Context(0)
    ctx.set_stream(stream.cuda_stream)
    d = spx.Dictionary(ctx, D)
    Lout = spx.lasso_L(m, p, -1, spx.PENALTY, float(lam.max()))
    ldg = (p + 15) // 16 * 16
    DtX = torch.empty((n, ldg), dtype=torch.float32, device=dev)
    norms = torch.empty((n,), dtype=torch.float32, device=dev)
    V = torch.empty((n, Lout), dtype=torch.float32, device=dev)
    R = torch.empty((n, Lout), dtype=torch.int32, device=dev)
    cnt = torch.empty((n,), dtype=torch.int32, device=dev)
    stats = torch.empty((n, 2), dtype=torch.int32, device=dev)

    def step(form, ev=None):
        if ev:
            ev[0].record(stream)
        spx.dtx_dev(ctx, d, Xt, DtX)
        spx.sqnorms_dev(ctx, Xt, norms)
        if ev:
            ev[1].record(stream)
        spx.lasso_dtx_dev(ctx, d, DtX, norms, lam_t, V, R, cnt, form=form,
                          waves_per_signal=args.waves, stats=stats)
        if ev:
            ev[2].record(stream)

    if args.calib_copy:
        a = torch.ones((64 << 20,), dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.add(a, 0.0, out=b)
        torch.cuda.synchronize()
        del a, b
    log("bench: warmup %d" % args.warmup)
    for _ in range(args.warmup):
        step("batched")
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    log("bench: timing %d steps" % args.steps)
    t0 = time.perf_counter()
    for k in range(args.steps):
        step("batched", evs[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = (t1 - t0) / args.steps
    gemm_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    solve_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    value = n / wall

    st = stats.cpu().numpy()
    counts = cnt.cpu().numpy()
    admissions = st[:, 1].astype(np.int64)
    # algorithmic bytes of the solver: the D'X column in, one Gram column per
    # atom admission, (value, index) per final coefficient out
    solver_bytes = float(np.sum(4 * p * (1 + admissions) + 8 * counts))
    gemm_flops = 2.0 * m * p * n

    plain = None
    if not args.no_plain:
        log("bench: plain form")
        step("plain")
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        tp = time.perf_counter()
        step("plain", ev)
        torch.cuda.synchronize()
        tp = time.perf_counter() - tp
        plain = {"value": n / tp, "unit": "signals/s", "solver_ms": ev[1].elapsed_time(ev[2]),
                 "batched_over_plain": value / (n / tp)}

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline")
        cpu = cpu_baseline(X, D, lam, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "signals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "LARS-Lasso penalty, D 256x1024 unit-norm N(0,1), X 256x%d "
                               "5-sparse + 0.05 noise normalised, lambda log-U[0.05,0.5] "
                               "per signal, FP32 MFMA D'X" % n,
                   "m": m, "p": p, "n": n, "form": "batched"},
        "roofline": {"bound": "hbm", "kernel": "lars_kernel",
                     "achieved": solver_bytes / (solve_ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": solver_bytes / (solve_ms * 1e-3) / HBM_PEAK,
                     "traffic": measured_traffic("lars_kernel"),
                     "algorithmic_bytes": solver_bytes},
        "roofline_gemm": {"bound": "mfma", "kernel": "gemm_tn_kernel<float>",
                          "achieved": gemm_flops / (gemm_ms * 1e-3) / 1e12,
                          "peak": FP32_MFMA_PEAK / 1e12, "unit": "TFLOP/s",
                          "frac": gemm_flops / (gemm_ms * 1e-3) / FP32_MFMA_PEAK},
        "kernel_ms": {"dtx_and_norms": gemm_ms, "lars_solver": solve_ms},
        "solver_stats": {"mean_admissions": float(admissions.mean()),
                         "mean_steps": float(st[:, 0].mean()),
                         "mean_nnz": float(counts.mean())},
        "plain_gpu": plain,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
