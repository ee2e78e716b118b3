"""Benchmark of the GEMM hot path on one MI355X.

Metric: GEMM TFLOP/s at 8192^3, split-BF16 emulated DGEMM vs native FP64
DGEMM, with the max relative error of each against the FP64 (Dot2) oracle.

A step is one C = A*B at M = N = K = 8192 (column-major, NN) on inputs already
resident in HBM, a_ij = u*2^e with u uniform in [-1,1) and e uniform in
[-12, 12] (the emulated-DGEMM workload), computed by the split-BF16 emulated
DGEMM with `--nslices` BF16 slices of 9 bits (default 6: the smallest count
whose componentwise error is at or below native FP64's on this workload).  `value` is that step's
throughput, 2*M*N*K / step time.  The native FP64-MFMA DGEMM is timed on the
same inputs for the comparison, and a slice-count table (3..8) is measured
after the timed region.

Only one GPU: `--gpus N` with N > 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "jblas.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "GEMM TFLOP/s at 8192³: split-BF16 emulated vs native FP64 DGEMM, max rel err"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nslices", type=int, default=6, help="BF16 slices (9 bits each)")
    ap.add_argument("--nslices-int8", type=int, default=7, help="INT8 slices (8 bits each)")
    ap.add_argument("--skip-cpu", action="store_true", help="omit the CPU baseline and the oracle check")
    ap.add_argument("--skip-table", action="store_true", help="omit the per-slice-count table")
    ap.add_argument("--only", choices=["ozaki", "native", "ozaki_int8", "grouped", "grouped_static"], default=None,
                    help="run only one method, or only the batched and rrule workloads, or only the batched "
                         "workload's static plans (for profiler runs)")
    ap.add_argument("--skip-grouped", action="store_true",
                    help="omit the batched (1024 uneven problems) and rrule (4096x2048x4096) measurements")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy a MIB-sized device buffer (a kernel of known traffic for counter scaling)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed(torch, fn, steps, warm=False):
    if warm:
        fn()  # first-launch costs (code-object load, workspace growth) are not steady state
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def grouped_and_rrule(torch, jb, tune=True, rrule=True):
    """Secondary workloads, measured after the headline step:
    * batched: 1024 independent problems, m, n, k log-uniform in [16, 2048]
      (multiples of 8), one persistent launch captured in a hipGraph, FP64 and
      FP32 (3xBF16); static tile choice, then after the autotuner (tune);
    * rrule: F32 A 4096x2048, B 2048x4096: forward + pullback (dA, dB in one
      grouped launch) (rrule)."""
    import numpy as np

    from jblas_amd import generators as G
    from jblas_amd.grouped import GemmDesc
    from jblas_amd.rrule import pullback_

    out = {}
    samples = {}
    shapes = G.loguniform_shapes(52, 1024)
    pick = sorted(int(i) for i in np.random.default_rng(53).choice(len(shapes), 8, replace=False))
    flops = float(sum(2.0 * m * n * k for m, n, k in shapes))
    for method, dtype in (("f64", torch.float64), ("f32_bf16x3", torch.float32)):
        descs, keep = [], []
        for i, (m, n, k) in enumerate(shapes):
            A = torch.empty((int(k), int(m)), dtype=dtype, device="cuda")
            B = torch.empty((int(n), int(k)), dtype=dtype, device="cuda")
            C = torch.zeros((int(n), int(m)), dtype=dtype, device="cuda")
            jb.fill_uniform_(A, 1000 + 3 * i)
            jb.fill_uniform_(B, 1001 + 3 * i)
            descs.append(GemmDesc("N", "N", 1.0, A.T, B.T, 0.0, C.T))
            keep += [A, B, C]
        res = {}
        for phase in ("static", "tuned") if tune else ("static",):
            if phase == "tuned":
                t0 = time.perf_counter()
                jb.autotune(method, descs, reps=2)
                res["autotune_seconds"] = round(time.perf_counter() - t0, 2)
            plan = jb.GroupedPlan(method, descs)
            ms = timed(torch, plan.launch, 1)  # warm
            ms = timed(torch, plan.launch, 10)
            plan.destroy()
            res[f"{phase}_ms"] = round(ms, 4)
            res[f"{phase}_tflops"] = round(flops / ms / 1e9, 3)
        # the grouped kernels against the matrix-core peak of the arithmetic they
        # run: FP64 MFMA (78.6 TFLOP/s), or BF16 MFMA for 3xBF16 (2516.6 TFLOP/s;
        # six BF16 products per FP32 product)
        best = max(res["static_tflops"], res.get("tuned_tflops", 0.0))
        ach, peak = (best, 78.6) if method == "f64" else (6 * best, 2516.6)
        res["roofline"] = {"bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                           "frac": round(ach / peak, 4),
                           "dtype": "fp64" if method == "f64" else "bf16 (6 products per fp32 product)"}
        torch.cuda.synchronize()
        # host copies of a few problems (operands, result) for the per-problem error check
        samples[method] = [(keep[3 * i].cpu().numpy().T, keep[3 * i + 1].cpu().numpy().T,
                            keep[3 * i + 2].cpu().numpy().T) for i in pick]
        out[method] = res
        del descs, keep
        torch.cuda.empty_cache()
    out["batched_config"] = "1024 problems, m,n,k log-uniform in [16,2048] (multiples of 8), seed 52; sum 2mnk = %.4g" % flops
    if not rrule:
        return out, samples
    # rrule sizes (F32): forward C = A*B and pullback dA = W*B^T, dB = A^T*W
    M, K, N = 4096, 2048, 4096
    # N(0,1)/sqrt(2048) and N(0,1), column-major, drawn in HBM (device Box-Muller)
    A = jb.fill_normal_(torch.empty((K, M), dtype=torch.float32, device="cuda"), 42, 2048 ** -0.5).T
    B = jb.fill_normal_(torch.empty((N, K), dtype=torch.float32, device="cuda"), 43, 2048 ** -0.5).T
    W = jb.fill_normal_(torch.empty((N, M), dtype=torch.float32, device="cuda"), 45).T
    Cm = jb.device_matrix(M, N, torch.float32)
    dA = jb.device_matrix(M, K, torch.float32)
    dB = jb.device_matrix(K, N, torch.float32)
    rr = {}
    for method in ("f32", "f32_bf16x3"):
        fwd = "native" if method == "f32" else "bf16x3"
        ms_f = timed(torch, lambda: jb.gemm_("N", "N", 1.0, A, B, 0.0, Cm, method=fwd), 5, warm=True)
        ms_b = timed(torch, lambda: pullback_("N", "N", 1.0, A, B, W, dA, dB, method=method), 5, warm=True)
        fl = 2.0 * M * N * K
        rr[method] = {"forward_ms": round(ms_f, 4), "pullback_ms": round(ms_b, 4),
                      "total_tflops": round(3 * fl / (ms_f + ms_b) / 1e9, 3)}
    out["rrule_4096x2048x4096_f32"] = rr
    return out, samples


def native_and_boundary(torch, jb):
    """Secondary lines, after the timed region:
    * native FP64 MFMA DGEMM, C = 1.5*op(A)*op(B) + 0.5*C, M=N=K in {1024,
      4096, 8192}, op in {N,T}^2 (device-resident operands);
    * jmul! at 256^3 (the reference's CPU-runnable case): device-resident
      call and the host-pointer call (operands staged over PCIe);
    * the host-pointer DGEMM at 8192^3 (PCIe-inclusive: A and B copied in,
      C copied out), native and split-BF16 s = 6;
    * torch's BF16 and FP64 GEMM at 8192^3 (vendor yardsticks)."""
    import numpy as np

    out = {"c2_native_tflops": {}}
    for n in (1024, 4096, 8192):
        A = torch.empty((n, n), dtype=torch.float64, device="cuda")
        B = torch.empty_like(A)
        C = torch.empty_like(A)
        jb.fill_uniform_(A, 42)
        jb.fill_uniform_(B, 43)
        jb.fill_uniform_(C, 44)
        for ta in "NT":
            for tb in "NT":
                def f(ta=ta, tb=tb):
                    jb.gemm_(ta, tb, 1.5, A.T, B.T, 0.5, C.T, method="native")
                f()
                reps = 20 if n < 8192 else 5
                ms = timed(torch, f, reps)
                out["c2_native_tflops"][f"{ta}{tb}_{n}"] = round(2.0 * n ** 3 / ms / 1e9, 2)
        del A, B, C
        torch.cuda.empty_cache()
    # jmul! 256^3: device-resident and host-pointer
    n = 256
    Ad = torch.empty((n, n), dtype=torch.float64, device="cuda")
    Xd = torch.empty_like(Ad)
    Dd = torch.empty_like(Ad)
    jb.fill_uniform_(Ad, 42)
    jb.fill_uniform_(Xd, 43)
    jb.jmul_(Dd.T, Ad.T, Xd.T)
    ms_dev = timed(torch, lambda: jb.jmul_(Dd.T, Ad.T, Xd.T), 200)
    Ah, Xh = np.asfortranarray(Ad.cpu().numpy().T), np.asfortranarray(Xd.cpu().numpy().T)
    Dh = np.zeros((n, n), order="F")
    jb.jmul_(Dh, Ah, Xh)
    t0 = time.perf_counter()
    for _ in range(50):
        jb.jmul_(Dh, Ah, Xh)
    ms_host = (time.perf_counter() - t0) / 50 * 1e3
    out["c1_jmul_256"] = {"device_us": round(ms_dev * 1e3, 2), "device_gflops": round(2 * n ** 3 / ms_dev / 1e6, 2),
                          "host_pointer_us": round(ms_host * 1e3, 2),
                          "host_pointer_gflops": round(2 * n ** 3 / ms_host / 1e6, 2)}
    # host-pointer DGEMM at 8192^3 (PCIe-inclusive)
    n = 8192
    Ah = np.asfortranarray(np.random.default_rng(1).uniform(-1, 1, (n, n)))
    Bh = np.asfortranarray(np.random.default_rng(2).uniform(-1, 1, (n, n)))
    Ch = np.zeros((n, n), order="F")
    host = {}
    for meth, kw in (("native", {}), ("ozaki_s6", {"method": "ozaki", "nslices": 6})):
        jb.gemm_("N", "N", 1.0, Ah, Bh, 0.0, Ch, **kw)
        t0 = time.perf_counter()
        for _ in range(2):
            jb.gemm_("N", "N", 1.0, Ah, Bh, 0.0, Ch, **kw)
        ms = (time.perf_counter() - t0) / 2 * 1e3
        host[meth] = {"ms": round(ms, 2), "tflops": round(2.0 * n ** 3 / ms / 1e9, 2)}
    out["host_pointer_8192_pcie_inclusive"] = host
    # vendor yardsticks on the same device: torch (hipBLASLt / rocBLAS) BF16 and
    # FP64 GEMM at 8192^3, for the BF16 rate the slice-pair kernel reaches and
    # for the native FP64 kernel
    vendor = {}
    for name, dt in (("bf16", torch.bfloat16), ("fp64", torch.float64)):
        X = torch.randn(n, n, device="cuda").to(dt)
        Y = torch.randn(n, n, device="cuda").to(dt)
        Z = torch.matmul(X, Y)
        ms = timed(torch, lambda: torch.matmul(X, Y, out=Z), 5)
        vendor[f"torch_{name}_8192_tflops"] = round(2.0 * n ** 3 / ms / 1e9, 1)
        del X, Y, Z
    out["vendor_gemm_8192"] = vendor
    return out


def profiled_traffic(key):
    """HBM bytes per launch of a kernel at this workload, as measured by the
    counter passes recorded in profiles/traffic.json (None if not profiled)."""
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(tpath) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def cpu_baseline_leg(A_sub_rows, B_sub_cols, n, panel_cols, nthreads, grouped_samples=None, lines=None):
    """CPU side, rank 0 only: (1) the oracle's restatement of the reference
    jmul! recurrence timed on a bounded panel of the same workload, beside the
    host BLAS, in a child process (oracle/cpu_timing.py: no GPU runtime threads,
    OpenMP threads pinned from the start, warmed legs, medians of 5), (2) the
    Dot2 FP64 oracle on the sampled rows x columns, on whole rows and columns
    of C (`lines` = (A row block, B, A stored by rows, B column block)) and on
    sampled problems of the batched launch, used only as the checker."""
    import subprocess

    import numpy as np

    from oracle import oracle as O

    out = {}
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores", OMP_WAIT_POLICY="passive",
               OMP_NUM_THREADS=str(nthreads))
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_timing", "--n", str(n), "--panel", str(panel_cols),
                        "--panel1t", os.environ.get("JB_CPU_PANEL_1T", "320"), "--threads", str(nthreads)],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"CPU timing child failed ({r.returncode}): {r.stderr[-2000:]}")
    t = json.loads(r.stdout.strip().splitlines()[-1])
    out["cpu"] = {"mt_tflops": t["panel"]["mt_tflops"], "mt_seconds": t["panel"]["mt_seconds"],
                  "mt_threads": t["threads_used"], "st_tflops": t["panel"]["st_tflops"],
                  "st_seconds": t["panel"]["st_seconds"], "blas_tflops": t["blas_tflops"],
                  "cpu_model": t["cpu_model"], "nproc": t["nproc"], "sizes": t["sizes"], "omp": t["omp"],
                  "panel": t["panel"]["sample"]}
    # (2) checker: Dot2 on the sampled block
    out["ref"] = O.gemm_dot2("N", "N", 1.0, A_sub_rows, B_sub_cols, 0.0,
                             np.zeros((A_sub_rows.shape[0], B_sub_cols.shape[1])), nthreads=nthreads)
    if lines is not None:
        # whole rows C[i, :] = A[i, :] B, and whole columns as C[:, j]^T = B[:, j]^T A^T with
        # A^T read column by column from A's row-major copy (unit-stride k in both operands)
        a_rows, b_full, a_by_rows, b_cols = lines
        t0 = time.perf_counter()
        out["ref_rows"] = O.gemm_dot2("T", "N", 1.0, np.asfortranarray(a_rows.T), b_full, 0.0,
                                      np.zeros((a_rows.shape[0], b_full.shape[1])), nthreads=nthreads)
        out["ref_cols_t"] = O.gemm_dot2("T", "N", 1.0, b_cols, a_by_rows.T, 0.0,
                                        np.zeros((b_cols.shape[1], a_by_rows.shape[0])), nthreads=nthreads)
        out["lines_seconds"] = time.perf_counter() - t0
    # (3) checker for sampled problems of the batched launch: per-problem error vs Dot2
    if grouped_samples:
        errs = {}
        for method, probs in grouped_samples.items():
            worst_rel, worst_scaled = 0.0, 0.0
            for A, B, C in probs:
                if method == "f64":
                    ref = O.gemm_dot2("N", "N", 1.0, A, B, 0.0, np.zeros(C.shape), nthreads=nthreads)
                else:
                    ref = O.sgemm_dot2("N", "N", 1.0, A, B, 0.0, np.zeros(C.shape, np.float32), nthreads=nthreads)
                absab = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
                d = np.abs(C.astype(np.float64) - ref)
                big = np.abs(ref) > 1e-3 * np.max(np.abs(ref))
                worst_rel = max(worst_rel, float(np.max(d[big] / np.abs(ref[big]))))
                worst_scaled = max(worst_scaled, float(np.max(d / np.where(absab > 0, absab, 1.0))))
            errs[method] = {"problems": [list(p[2].shape) + [p[0].shape[1]] for p in probs],
                            "max_rel_err": worst_rel, "max_err_over_absAB": worst_scaled}
        out["grouped_errors"] = errs
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only: sharding the GEMM across GPUs "
                                   "(by tiles of C) is not launched", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    import jblas_amd as jb
    from jblas_amd import _lib

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    jb.init(0)
    info = jb.device_info()
    n, s, s8 = args.n, args.nslices, args.nslices_int8
    flops = 2.0 * n * n * n
    pairs, pairs8 = s * (s + 1) // 2, s8 * (s8 + 1) // 2
    log(f"device {info['arch']} CUs={info['compute_units']} clock={info['clock_khz'] / 1e6:.2f}GHz "
        f"peak fp64={info['peak_fp64_tflops']:.1f} bf16={info['peak_bf16_tflops']:.1f} TF")

    if args.only in ("grouped", "grouped_static"):
        full = args.only == "grouped"
        extra, _ = grouped_and_rrule(torch, jb, tune=full, rrule=full)
        print(json.dumps({"grouped_and_rrule": extra}), flush=True)
        return

    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration * 2 ** 20 // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    # inputs resident in HBM (seeded device generator, bit-identical to the host one)
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    B = torch.empty((n, n), dtype=torch.float64, device="cuda")
    C = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    jb.fill_wide_(A, 42, -12, 12)
    jb.fill_wide_(B, 43, -12, 12)
    Ac, Bc, Cc = A.T, B.T, C.T  # column-major views

    def step_ozaki():
        jb.gemm_("N", "N", 1.0, Ac, Bc, 0.0, Cc, method="ozaki", nslices=s)

    def step_native():
        jb.gemm_("N", "N", 1.0, Ac, Bc, 0.0, Cc, method="native")

    def step_int8():
        jb.gemm_("N", "N", 1.0, Ac, Bc, 0.0, Cc, method="ozaki_int8", nslices=s8)

    result = {}
    # ---- emulated DGEMM (the measured step) ----
    if args.only in (None, "ozaki"):
        for _ in range(args.warmup):
            step_ozaki()
        torch.cuda.synchronize()
        _lib.kernel_timing_reset()
        _lib.kernel_timing(True)
        ms = timed(torch, step_ozaki, args.steps)
        _lib.kernel_timing(False)
        k_ms, k_n = _lib.kernel_time("slice_pair_gemm")
        sp_ms, sp_n = _lib.kernel_time("ozaki_split")
        result["ozaki"] = {"ms": ms, "tflops": flops / ms / 1e9, "gemm_kernel_ms": k_ms / max(k_n, 1),
                           "split_ms": sp_ms / max(sp_n, 1), "launches": k_n}
        log(f"ozaki s={s}: {ms:.3f} ms/step  {flops / ms / 1e9:.2f} TF  (slice-pair kernel "
            f"{k_ms / max(k_n, 1):.3f} ms, split {sp_ms / max(sp_n, 1):.3f} ms)")
        C_oz = C.clone() if not args.skip_cpu else None
    # ---- native FP64 DGEMM (comparison) ----
    if args.only in (None, "native"):
        for _ in range(args.warmup):
            step_native()
        _lib.kernel_timing_reset()
        _lib.kernel_timing(True)
        ms_n = timed(torch, step_native, args.steps)
        _lib.kernel_timing(False)
        kn_ms, kn_n = _lib.kernel_time("dgemm_mfma")
        result["native"] = {"ms": ms_n, "tflops": flops / ms_n / 1e9, "kernel_ms": kn_ms / max(kn_n, 1)}
        log(f"native: {ms_n:.3f} ms/step  {flops / ms_n / 1e9:.2f} TF")
        C_nat = C.clone() if not args.skip_cpu else None
    # ---- emulated DGEMM on INT8 slices (same digit weights, int8 MFMA) ----
    if args.only in (None, "ozaki_int8"):
        for _ in range(args.warmup):
            step_int8()
        _lib.kernel_timing_reset()
        _lib.kernel_timing(True)
        ms_i = timed(torch, step_int8, args.steps)
        _lib.kernel_timing(False)
        ki_ms, ki_n = _lib.kernel_time("slice_pair_gemm_i8")
        si_ms, si_n = _lib.kernel_time("ozaki_i8_split")
        result["int8"] = {"ms": ms_i, "tflops": flops / ms_i / 1e9, "gemm_kernel_ms": ki_ms / max(ki_n, 1),
                          "split_ms": si_ms / max(si_n, 1)}
        log(f"ozaki_int8 s={s8}: {ms_i:.3f} ms/step  {flops / ms_i / 1e9:.2f} TF  (slice-pair kernel "
            f"{ki_ms / max(ki_n, 1):.3f} ms)")
        C_i8 = C.clone() if not args.skip_cpu else None

    # ---- slice-count table (after the timed region) ----
    table = []
    if args.only is None and not args.skip_table:
        for meth, srange in (("ozaki", range(3, 9)), ("ozaki_int8", range(5, 9))):
            for ss in srange:
                def f(ss=ss, meth=meth):
                    jb.gemm_("N", "N", 1.0, Ac, Bc, 0.0, Cc, method=meth, nslices=ss)
                f()
                ms_s = timed(torch, f, 2)
                table.append({"method": meth, "nslices": ss, "ms": round(ms_s, 3),
                              "tflops": round(flops / ms_s / 1e9, 2),
                              "out": C.clone() if not args.skip_cpu else None})
                log(f"  table {meth} s={ss}: {ms_s:.3f} ms  {flops / ms_s / 1e9:.2f} TF")

    # ---- Float32 through the 3xBF16 split at the same size (after the timed region) ----
    f32 = None
    if args.only is None and not args.skip_table:
        f32 = {"rows": []}
        A32, B32 = A.float(), B.float()  # the wide-range FP64 streams rounded to float (jb_fill_wide_f32_dev's values)
        C32 = torch.zeros((n, n), dtype=torch.float32, device="cuda")
        for meth, nprod in (("bf16x3", 3), ("bf16x3", 6), ("bf16x3", 9), ("native", None)):
            def f(meth=meth, nprod=nprod):
                jb.gemm_("N", "N", 1.0, A32.T, B32.T, 0.0, C32.T, method=meth, nprod=nprod or 6)
            for _ in range(args.warmup):
                f()
            _lib.kernel_timing_reset()
            _lib.kernel_timing(True)
            ms_f = timed(torch, f, max(2, args.steps // 2))
            _lib.kernel_timing(False)
            kname = "bf16x3_gemm" if meth == "bf16x3" else "sgemm_mfma"
            k_ms, k_n = _lib.kernel_time(kname)
            row = {"method": meth, "nprod": nprod, "ms": round(ms_f, 3), "tflops": round(flops / ms_f / 1e9, 2),
                   "kernel_ms": round(k_ms / max(k_n, 1), 4), "out": C32.clone() if not args.skip_cpu else None}
            if meth == "bf16x3":
                hw = nprod * flops / (k_ms / max(k_n, 1) * 1e-3) / 1e12
                # block width the library picks: 256 where that grid holds two tiles per CU
                bn = 256 if -(-n // 128) * -(-n // 256) >= 2 * info["compute_units"] else 128
                row["roofline"] = {"kernel": f"slice_pair_gemm_kernel<1,3,{ {3: 3, 6: 4, 9: 6}[nprod] },2,{bn}>",
                                   "bound": "mfma", "achieved": round(hw, 2),
                                   "peak": round(info["peak_bf16_tflops"], 1), "unit": "TFLOP/s",
                                   "frac": round(hw / info["peak_bf16_tflops"], 4)}
            f32["rows"].append(row)
            log(f"  f32 {meth} nprod={nprod}: {ms_f:.3f} ms  {flops / ms_f / 1e9:.2f} TF")
        if args.skip_cpu:
            for row in f32["rows"]:
                row.pop("out", None)
        del C32

    # ---- batched and rrule workloads (after the timed region) ----
    extra, gsamples = None, None
    if args.only is None and not args.skip_grouped:
        extra, gsamples = grouped_and_rrule(torch, jb)
        log(f"grouped/rrule: {extra}")
    secondary = None
    if args.only is None and not args.skip_table:
        secondary = native_and_boundary(torch, jb)
        log(f"native table / boundary: {secondary}")

    # ---- CPU baseline + checker (rank 0, N = 1) ----
    cpu = None
    errs = {}
    if not args.skip_cpu:
        rng = np.random.default_rng(88)
        ii = np.sort(rng.choice(n, 64, replace=False))
        jj = np.sort(rng.choice(n, 64, replace=False))
        # the column-major matrix of a row-major (n, n) storage tensor X is X.T:
        # X.cpu().numpy().T is that matrix as a Fortran-ordered host array
        Ah = A.cpu().numpy().T
        Bh = B.cpu().numpy().T
        A_rows = np.asfortranarray(Ah[ii, :])
        B_cols = np.asfortranarray(Bh[:, jj])
        panel_cols = int(os.environ.get("JB_CPU_PANEL", "6144"))
        nthreads = max(1, min(16, os.cpu_count() or 1))
        # 16 whole rows and 16 whole columns of C, besides the 64 x 64 sampled block
        li, lj = ii[::4], jj[::4]
        A_by_rows = np.ascontiguousarray(Ah)
        leg = cpu_baseline_leg(A_rows, B_cols, n, panel_cols, nthreads, gsamples,
                               lines=(A_rows[::4], Bh, A_by_rows, B_cols[:, ::4]))
        absab_rows = np.abs(A_rows[::4]) @ np.abs(Bh)
        absab_cols = np.abs(Ah) @ np.abs(B_cols[:, ::4])
        del Ah, Bh, A_by_rows
        ref = leg["ref"]
        absab = np.abs(A_rows) @ np.abs(B_cols)
        ref_all = np.concatenate([ref.ravel(), leg["ref_rows"].ravel(), leg["ref_cols_t"].T.ravel()])
        absab_all = np.concatenate([absab.ravel(), absab_rows.ravel(), absab_cols.ravel()])
        log(f"Dot2 on whole rows/columns: {leg['lines_seconds']:.2f} s")

        def err(out_t):
            """(max rel, max |err|/(|A||B|), normwise, max rel over entries with
            |C*| > 1e-3 max|C*|, count of entries excluded from that) on the Dot2-checked sample."""
            Cm = out_t.cpu().numpy().T
            got = np.concatenate([Cm[np.ix_(ii, jj)].ravel(), Cm[li, :].ravel(), Cm[:, lj].ravel()])
            return err_of(got, ref_all, absab_all)

        def err_of(got, ref_v, absab_v):
            d = np.abs(got - ref_v)
            big = np.abs(ref_v) > 1e-3 * np.max(np.abs(ref_v))
            return (float(np.max(d / np.abs(ref_v))), float(np.max(d / absab_v)),
                    float(np.linalg.norm(got - ref_v) / np.linalg.norm(ref_v)),
                    float(np.max(d[big] / np.abs(ref_v[big]))), int(np.count_nonzero(~big)),
                    float(np.mean(d[big] / np.abs(ref_v[big]))))

        if "ozaki" in result:
            errs["ozaki"] = err(C_oz)
        if "native" in result:
            errs["native"] = err(C_nat)
        if "int8" in result:
            errs["int8"] = err(C_i8)
        for row in table:
            e = err(row.pop("out"))
            row["max_rel_err"], row["max_err_over_absAB"], row["normwise_err"] = e[:3]
            row["max_rel_err_excluding_small"], row["mean_rel_err_excluding_small"] = e[3], e[5]
            if "native" in errs:
                row["err_ratio_vs_native"] = round(row["max_rel_err"] / errs["native"][0], 3)
        if f32 is not None:
            # FP64 Dot2 oracle on the float inputs (exact in FP64), same sample
            from oracle import oracle as O
            A32h = A32.double().cpu().numpy().T
            B32h = B32.double().cpu().numpy().T
            r_blk = O.gemm_dot2("N", "N", 1.0, np.asfortranarray(A32h[ii, :]), np.asfortranarray(B32h[:, jj]), 0.0,
                                np.zeros((ii.size, jj.size)), nthreads=nthreads)
            r_rows = O.gemm_dot2("T", "N", 1.0, np.asfortranarray(A32h[li, :].T), B32h, 0.0,
                                 np.zeros((li.size, n)), nthreads=nthreads)
            r_cols = O.gemm_dot2("T", "N", 1.0, np.asfortranarray(B32h[:, lj]), np.ascontiguousarray(A32h).T, 0.0,
                                 np.zeros((lj.size, n)), nthreads=nthreads)
            ref32 = np.concatenate([r_blk.ravel(), r_rows.ravel(), r_cols.T.ravel()])
            aA, aB = np.abs(A32h), np.abs(B32h)
            absab32 = np.concatenate([(aA[ii, :] @ aB[:, jj]).ravel(), (aA[li, :] @ aB).ravel(),
                                      (aA @ aB[:, lj]).ravel()])
            del A32h, B32h, aA, aB
            for row in f32["rows"]:
                Cm = row.pop("out").cpu().numpy().T.astype(np.float64)
                got = np.concatenate([Cm[np.ix_(ii, jj)].ravel(), Cm[li, :].ravel(), Cm[:, lj].ravel()])
                e = err_of(got, ref32, absab32)
                row.update({"max_err_over_absAB": e[1], "max_rel_err": e[0], "normwise_err": e[2],
                            "max_rel_err_excluding_small": e[3], "mean_rel_err_excluding_small": e[5],
                            "bar_2K_u32": 2 * n * 2.0 ** -24})
        cpu = leg["cpu"]
        if extra is not None and "grouped_errors" in leg:
            for method, e in leg["grouped_errors"].items():
                extra[method]["sampled_problem_errors"] = e
        log(f"errors: {errs}; cpu: {cpu}")
    else:
        for row in table:
            row.pop("out", None)

    oz = result.get("ozaki")
    nat = result.get("native")
    i8 = result.get("int8")
    main_ms = oz["ms"] if oz else (nat["ms"] if nat else i8["ms"])
    line = {
        "metric": METRIC,
        "value": round(flops / main_ms / 1e9, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if oz else ("f64" if nat else "int8"),
        "data": "synthetic (seeded SplitMix64 in HBM: a_ij = u*2^e, u in [-1,1), e in [-12,12])",
        "config": {"workload": f"dgemm {n}x{n}x{n} column-major NN, split-BF16 emulated with {s} slices of 9 bits "
                               f"({pairs} BF16 slice-pair products, FP32 exact level sums, FP64 combine) "
                               f"vs native FP64 MFMA",
                   "m": n, "n": n, "k": n, "nslices": s, "parallelism": "single GPU"},
    }
    if oz:
        line["emulated_tflops"] = round(oz["tflops"], 3)
    if nat:
        line["native_fp64_tflops"] = round(nat["tflops"], 3)
        line["native_fp64_frac_of_peak"] = round(nat["tflops"] / info["peak_fp64_tflops"], 4)
    if oz and nat:
        line["speedup_emulated_vs_native"] = round(nat["ms"] / oz["ms"], 4)
    if i8:
        line["emulated_int8"] = {"method": f"ozaki_int8, {s8} INT8 slices of 8 bits on v_mfma_i32_32x32x32_i8 "
                                           f"({pairs8} slice-pair products, exact int32 level sums, FP64 combine)",
                                 "tflops": round(i8["tflops"], 3), "ms_per_step": round(i8["ms"], 4),
                                 "split_kernels_ms": round(i8["split_ms"], 4)}
        if nat:
            line["emulated_int8"]["speedup_vs_native"] = round(nat["ms"] / i8["ms"], 4)
        hw = pairs8 * flops
        ach_i = hw / (i8["gemm_kernel_ms"] * 1e-3) / 1e12
        peak_i8 = info["peak_bf16_tflops"] * 2.0  # int8 MFMA: 8192 ops/clk/CU
        line["emulated_int8"]["roofline"] = {"kernel": f"slice_pair_gemm_kernel<2,{s8},{s8 + 1},1,64>", "bound": "mfma",
                                             "achieved": round(ach_i, 2), "peak": round(peak_i8, 1),
                                             "unit": "TOP/s", "frac": round(ach_i / peak_i8, 4),
                                             "traffic": profiled_traffic(f"slice_pair_gemm_i8_s{s8}_n{n}"),
                                             "avg_launch_ms": round(i8["gemm_kernel_ms"], 4)}
    if errs:
        if "ozaki" in errs:
            line["max_rel_err_emulated"] = errs["ozaki"][0]
        if "native" in errs:
            line["max_rel_err_native"] = errs["native"][0]
        if "ozaki" in errs and "native" in errs:
            line["err_ratio_emulated_vs_native"] = round(errs["ozaki"][0] / errs["native"][0], 4)
        if "int8" in errs and "emulated_int8" in line:
            line["emulated_int8"]["max_rel_err"] = errs["int8"][0]
            if "native" in errs:
                line["emulated_int8"]["err_ratio_vs_native"] = round(errs["int8"][0] / errs["native"][0], 4)
        line["error_sample"] = (f"Dot2 oracle on a 64 x 64 block of sampled rows x sampled columns plus 16 whole "
                                f"rows and 16 whole columns ({4096 + 32 * n} entries)")
        line["error_metrics"] = {
            name: {"componentwise_scaled": e[1], "max_rel": e[0], "normwise": e[2],
                   "max_rel_excluding_small": e[3], "excluded_entries": e[4], "mean_rel_excluding_small": e[5]}
            for name, e in (("emulated", errs.get("ozaki")), ("native", errs.get("native")),
                            ("emulated_int8", errs.get("int8"))) if e is not None}
    # roofline of the dominant kernel
    if oz:
        hw_flops = pairs * flops
        ach = hw_flops / (oz["gemm_kernel_ms"] * 1e-3) / 1e12
        traffic = profiled_traffic(f"slice_pair_gemm_s{s}_n{n}")
        line["roofline"] = {"kernel": f"slice_pair_gemm_kernel<0,{s},{s + 1},1,64>", "bound": "mfma",
                            "achieved": round(ach, 2), "peak": round(info["peak_bf16_tflops"], 1),
                            "unit": "TFLOP/s", "frac": round(ach / info["peak_bf16_tflops"], 4),
                            "traffic": traffic,
                            "flops_per_launch": hw_flops, "avg_launch_ms": round(oz["gemm_kernel_ms"], 4)}
        line["split_kernels_ms"] = round(oz["split_ms"], 4)
    elif nat:
        ach = flops / (nat["kernel_ms"] * 1e-3) / 1e12
        line["roofline"] = {"kernel": "gemm_mfma_kernel<double>", "bound": "mfma", "achieved": round(ach, 2),
                            "peak": round(info["peak_fp64_tflops"], 2), "unit": "TFLOP/s",
                            "frac": round(ach / info["peak_fp64_tflops"], 4), "traffic": None}
    if nat:
        line["roofline_native"] = {"kernel": "gemm_mfma_kernel<double>", "bound": "mfma",
                                   "achieved": round(flops / (nat["kernel_ms"] * 1e-3) / 1e12, 2),
                                   "peak": round(info["peak_fp64_tflops"], 2), "unit": "TFLOP/s",
                                   "frac": round(flops / (nat["kernel_ms"] * 1e-3) / 1e12 / info["peak_fp64_tflops"], 4)}
    if table:
        line["nslices_table"] = table
    if f32:
        line["sgemm_f32_8192"] = {"workload": f"sgemm {n}x{n}x{n} NN, Float32 inputs = the wide-range operands "
                                              f"rounded to float; 3xBF16 split with 3, 6, 9 slice products on BF16 "
                                              f"MFMA, and native FP32 MFMA", "rows": f32["rows"]}
    if extra:
        line["batched_and_rrule"] = extra
    if secondary:
        line.update(secondary)
    if cpu:
        line["cpu_baseline"] = {"value": round(cpu["mt_tflops"], 5), "unit": "TFLOP/s", "cores": cpu["mt_threads"],
                                "kind": "port",
                                "sample": cpu["panel"] + "; the oracle's C restatement of jmul!'s fma recurrence "
                                          "(register tile per pick_kernel_size, no packing, as the reference)",
                                "single_thread_tflops": round(cpu["st_tflops"], 5),
                                "seconds": round(cpu["mt_seconds"] + cpu["st_seconds"], 2),
                                "host": {"cpu_model": cpu["cpu_model"], "nproc": cpu["nproc"]},
                                "numpy_openblas_dgemm_tflops": {
                                    f"{th}_threads": round(v, 4) for th, v in cpu["blas_tflops"].items()},
                                "numpy_openblas_sample": "A (full) x B[:, 0:512], median of 5 after a warm-up",
                                "sizes_gflops": cpu["sizes"],
                                "sizes_protocol": "uniform [-1,1) square operands; median of 5 after a warm-up; "
                                                  "restatement = the oracle's jmul! port; run in a child process "
                                                  "without GPU runtime threads, OpenMP threads pinned "
                                                  f"({cpu['omp']})"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
