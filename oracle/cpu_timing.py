"""CPU-baseline timing of the reference's GEMM path, run as a child process.

TEST INFRASTRUCTURE ONLY: started by bench.py's CPU-baseline leg (never by the
product).  It times the oracle's C restatement of jmul! (src/gemm.jl:244-348:
register tile per pick_kernel_size, no packing, no cache blocking) and, beside
it, the host BLAS (numpy dgemm), in a fresh process that holds no GPU runtime
threads.  The parent sets the OpenMP placement in this process's environment
(OMP_PROC_BIND, OMP_PLACES, and OMP_WAIT_POLICY=passive so that idle OpenMP
threads do not spin beside the BLAS threads) before it starts, so the thread
pool is pinned from its creation; every leg is warmed once before it is timed,
and the restatement's legs run before the BLAS legs.

Output: one JSON object on stdout.
  panel      the restatement on D[:, 0:panel] = A*B[:, 0:panel] of the n^3
             wide-range workload (A, B from the seeded host generator, the
             same streams as the GPU run), all threads, and on `panel1t`
             columns with one thread;
  blas       numpy dgemm A*B[:, 0:512], median of 5, all threads and one;
  sizes      the reference's CPU protocol: square uniform operands at 256^3,
             1024^3, 4096^3, median of 5 (the one-thread restatement at
             4096^3, ~20 s a repetition, is left out).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for _p in (ROOT, os.path.join(ROOT, "jblas.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

from oracle import oracle as O  # noqa: E402


def median_seconds(fn, reps: int) -> float:
    fn()  # warm-up (thread pool, caches, page faults)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--panel", type=int, default=6144)
    ap.add_argument("--panel1t", type=int, default=320)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    from jblas_amd import generators as G

    th = a.threads
    out = {"threads_requested": th, "threads_used": O.threads_used(th), "nproc": os.cpu_count(),
           "omp": {k: os.environ.get(k) for k in ("OMP_PROC_BIND", "OMP_PLACES", "OMP_WAIT_POLICY")}}
    try:
        with open("/proc/cpuinfo") as f:
            out["cpu_model"] = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        out["cpu_model"] = None
    n = a.n
    A = G.matrix(42, n, n, kind="wide")
    # the first `panel` columns of B are the first n*panel values of B's stream
    Bp = np.asfortranarray(G.wide(43, n * a.panel).reshape(a.panel, n).T)
    flops = 2.0 * n * n * a.panel
    O.jmul(A[:256, :], Bp[:, :16], nthreads=th)  # warm the pool
    t0 = time.perf_counter()
    O.jmul(A, Bp, nthreads=th)
    t_mt = time.perf_counter() - t0
    B1 = np.asfortranarray(Bp[:, : a.panel1t])
    t0 = time.perf_counter()
    O.jmul(A, B1, nthreads=1)
    t_1 = time.perf_counter() - t0
    out["panel"] = {"mt_tflops": flops / t_mt / 1e12, "mt_seconds": t_mt,
                    "st_tflops": 2.0 * n * n * a.panel1t / t_1 / 1e12, "st_seconds": t_1,
                    "sample": f"D[:, 0:{a.panel}] = A*B[:, 0:{a.panel}] of the {n}x{n}x{n} workload "
                              f"(single-thread leg: {a.panel1t} columns)"}
    # the restatement's legs first, then the BLAS legs: two thread pools never
    # wait for cores side by side
    sizes = {}
    rng = np.random.default_rng(7)
    mats = {}
    for m in (256, 1024, 4096):
        X = np.asfortranarray(rng.uniform(-1.0, 1.0, (m, m)))
        Y = np.asfortranarray(rng.uniform(-1.0, 1.0, (m, m)))
        mats[m] = (X, Y)
        row = {}
        gf = 2.0 * m ** 3 / 1e9
        row[f"jmul_restatement_{th}t_gflops"] = round(gf / median_seconds(lambda: O.jmul(X, Y, nthreads=th), a.reps), 3)
        if m <= 1024:
            row["jmul_restatement_1t_gflops"] = round(gf / median_seconds(lambda: O.jmul(X, Y, nthreads=1), a.reps), 3)
        sizes[f"{m}^3"] = row
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        threadpool_limits = None
    blas = {}
    if threadpool_limits is not None:
        Bq = np.asfortranarray(Bp[:, :512])
        for t in (th, 1):
            with threadpool_limits(limits=t, user_api="blas"):
                blas[str(t)] = 2.0 * n * n * 512 / median_seconds(lambda: A @ Bq, a.reps) / 1e12
                for m, (X, Y) in mats.items():
                    sizes[f"{m}^3"][f"blas_{t}t_gflops"] = round(2.0 * m ** 3 / 1e9 /
                                                                 median_seconds(lambda: X @ Y, a.reps), 3)
    out["blas_tflops"] = blas
    out["sizes"] = sizes
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
