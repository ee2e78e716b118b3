"""GPU parity of the grouped (batched) GEMM, its hipGraph plans, the
autotuner and the gemm! pullback (rrule), through the C ABI.

Bars:
  * FP64 grouped: BIT-EXACT against the oracle's split-K restatement of the
    reference recurrence (oracle.gemm_recurrence_chunked with the k-chunk the
    plan chose for that problem, as grouped_choices reports it) -- each split
    is an fma chain in k order (f64 MFMA) and the partials are summed in split
    order.
  * FP32 (native MFMA and 3xBF16): |C - C*| <= 2*K*2^-24 (|alpha||A||B| + |beta||C0|)
    componentwise against the FP64 oracle.
  * graph replays are bit-identical to the eager launch, replay after replay.
  * rrule: closed form dA = W B^T, dB = A^T W; FP64 central differences
    (h = 1e-6) on a 64^3 problem agree to 1e-7 relative (the loss is linear, so
    the difference quotient is exact up to rounding ~ u*|L|/h).
"""
import math
import os
import tempfile

import numpy as np
import pytest

import jblas_amd as jb
from jblas_amd import generators as G
from jblas_amd.grouped import GemmDesc
from oracle import oracle as O

from conftest import to_dev, to_host

pytestmark = pytest.mark.gpu
U32 = 2.0 ** -24


def dev_uniform(torch, rows, cols, seed, dtype):
    t = torch.empty((cols, rows), dtype=dtype, device="cuda")
    jb.fill_uniform_(t, seed)
    return t.T  # column-major rows x cols, column-major fill of the stream


def dev_copy(torch, a):
    """Column-major device copy of a float64 matrix, empty shapes included."""
    t = torch.empty((a.shape[1], a.shape[0]), dtype=torch.float64, device="cuda")
    if a.size:
        t.copy_(torch.from_numpy(np.ascontiguousarray(a.T)))
    return t.T


def kchunks(method, descs):
    """k per split of every problem in the grouped launch of `descs` (the plan's choice now)."""
    return [c[2] if c is not None else 32 for c in jb.grouped_choices(method, descs)]


def write_table(path, method, shapes, cfg, splitk):
    """A tuning table (the library's JSON format) with one entry (cfg, splitk) for
    the size bucket of each shape: (floor(log2(floor(sqrt(m n)))), floor(log2 k))."""
    info = jb.device_info()
    pre = f"{info['name']}|{info['arch']}|m{jb.grouped.method_id(method)}|"
    keys = sorted({f"mn{max(math.isqrt(m * n), 1).bit_length() - 1}|k{max(k, 1).bit_length() - 1}"
                   for m, n, k in shapes})
    entries = ",\n".join(f'    {{"key": "{pre}{k}", "cfg": {cfg}, "splitk": {splitk}, "ms": 0.0}}' for k in keys)
    with open(path, "w") as f:
        f.write('{\n  "format": "jblas-tune-1",\n  "entries": [\n' + entries + "\n  ]\n}\n")
    jb.tune_load(path)


def make_group(torch, count, seed, dtype, ops=("N", "N"), alpha=1.0, beta=0.0):
    shapes = G.loguniform_shapes(seed, count)
    probs = []
    for i, (m, n, k) in enumerate(shapes):
        ta, tb = ops if isinstance(ops, tuple) else ops[i % len(ops)]
        A = dev_uniform(torch, m if ta == "N" else k, k if ta == "N" else m, 1000 + 3 * i, dtype)
        B = dev_uniform(torch, k if tb == "N" else n, n if tb == "N" else k, 1001 + 3 * i, dtype)
        C = dev_uniform(torch, m, n, 1002 + 3 * i, dtype)
        probs.append((ta, tb, int(m), int(n), int(k), A, B, C))
    descs = [GemmDesc(ta, tb, alpha, A, B, beta, C) for (ta, tb, m, n, k, A, B, C) in probs]
    return probs, descs


def test_grouped_f64_c4_bitexact(gpu):
    """Config: 1024 problems, m,n,k log-uniform in [16, 2048] (multiples of 8), FP64."""
    torch = gpu
    probs, descs = make_group(torch, 1024, 52, torch.float64)
    C0 = [to_host(p[7]) for p in probs[::8]]
    choices = jb.grouped_choices("f64", descs)
    kinds = {c[0] for c in choices}  # both persistent kernels and the small path take part
    assert "small" in kinds and "128x128" in kinds and kinds & {"32x32", "64x64"}, kinds
    jb.gemm_grouped_("f64", descs)
    torch.cuda.synchronize()
    for idx, (ta, tb, m, n, k, A, B, C) in enumerate(probs[::8]):
        tile, sk, kc = choices[8 * idx]
        ref = O.gemm_recurrence_chunked(ta, tb, 1.0, to_host(A), to_host(B), 0.0, C0[idx], kc)
        assert np.array_equal(to_host(C), ref), (m, n, k, tile, sk)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_grouped_f64_ops_alpha_beta(gpu, ta, tb):
    torch = gpu
    probs, descs = make_group(torch, 48, 60 + ord(ta) + 2 * ord(tb), torch.float64, ops=(ta, tb), alpha=1.5,
                              beta=0.5)
    C0 = [to_host(p[7]) for p in probs]
    kc = kchunks("f64", descs)
    jb.gemm_grouped_("f64", descs)
    torch.cuda.synchronize()
    for idx, (ta_, tb_, m, n, k, A, B, C) in enumerate(probs):
        ref = O.gemm_recurrence_chunked(ta, tb, 1.5, to_host(A), to_host(B), 0.5, C0[idx], kc[idx])
        assert np.array_equal(to_host(C), ref), (m, n, k, kc[idx])


def test_grouped_forced_splitk_bitexact(gpu):
    """Small outputs with long k take split-K; the combine is deterministic."""
    torch = gpu
    shapes = [(16, 16, 2048), (24, 40, 1536), (64, 64, 2000), (8, 8, 4096)]
    probs = []
    for i, (m, n, k) in enumerate(shapes):
        A = dev_uniform(torch, m, k, 10 + i, torch.float64)
        B = dev_uniform(torch, k, n, 20 + i, torch.float64)
        C = dev_uniform(torch, m, n, 30 + i, torch.float64)
        probs.append((m, n, k, A, B, C))
    C0 = [to_host(p[5]) for p in probs]
    descs = [GemmDesc("N", "N", 1.0, A, B, 1.0, C) for (m, n, k, A, B, C) in probs]
    choices = jb.grouped_choices("f64", descs)
    # statically one split each; the plan splits k while the group has fewer
    # work items than the device holds resident workgroups, down to 256 k a split
    assert all(jb.tune_lookup("f64", m, n, k)[1] == 1 for (m, n, k, *_r) in probs)
    assert all(sk > 1 and kc >= 256 for _, sk, kc in choices), choices
    for rep in range(2):  # the combine must give the same bits every time
        for i, p in enumerate(probs):
            p[5].copy_(to_dev(torch, C0[i]))
        jb.gemm_grouped_("f64", descs)
        torch.cuda.synchronize()
        for i, (m, n, k, A, B, C) in enumerate(probs):
            ref = O.gemm_recurrence_chunked("N", "N", 1.0, to_host(A), to_host(B), 1.0, C0[i], choices[i][2])
            assert np.array_equal(to_host(C), ref), (m, n, k, choices[i])


LARGE_SHAPES = [(256, 256, 160), (200, 130, 77), (128, 384, 48), (131, 129, 1000), (130, 140, 0)]


def strided(torch, rows, cols, seed, ld, offset):
    """rows x cols column-major view with leading dimension `ld` starting
    `offset` elements into a fresh buffer (odd ld / offset: no 16-byte vectors)."""
    buf = torch.empty(offset + ld * max(cols, 1), dtype=torch.float64, device="cuda")
    jb.fill_uniform_(buf, seed)
    return buf[offset:offset + ld * cols].view(cols, ld).T[:rows] if cols else buf[:0].view(0, rows).T


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("tile,splitk", [("128x128", 1), ("128x128", 3), ("128x64", 1), ("128x64", 2),
                                         ("64x128", 1), ("64x128", 3)])
def test_grouped_large_tile_bitexact(gpu, ta, tb, tile, splitk, tmp_path):
    """The FP64 tiles with 128 rows or columns (a persistent kernel per tile
    shape and transpose combination), forced through a tuning table on fringe
    shapes (partial tiles in m, n and k; k == 0), with 16-byte vector loads
    (aligned bases, even leading dimensions) and without (odd leading
    dimension or an odd element offset), one split or more: bit-exact against
    the chunked recurrence."""
    torch = gpu
    jb.tune_reset()
    cfg = {v: k for k, v in jb.grouped.TILES.items()}[tile]
    write_table(str(tmp_path / "large.json"), "f64", LARGE_SHAPES, cfg, splitk)
    try:
        probs = []
        for i, (m, n, k) in enumerate(LARGE_SHAPES):
            ra, ca = (m, k) if ta == "N" else (k, m)
            rb, cb = (k, n) if tb == "N" else (n, k)
            odd = i % 2 == 1  # every other problem on the element-wise loads
            A = strided(torch, ra, ca, 300 + 3 * i, ra + (1 if odd else 2), 1 if odd else 0)
            B = strided(torch, rb, cb, 301 + 3 * i, rb + (3 if odd else 0), 0)
            C = strided(torch, m, n, 302 + 3 * i, m + 5, 3)
            probs.append((m, n, k, A, B, C))
        C0 = [to_host(p[5]) for p in probs]
        for alpha, beta in [(1.5, 0.5), (-1.0, 0.0)]:
            for i, p in enumerate(probs):
                p[5].copy_(to_dev(torch, C0[i]))
            descs = [GemmDesc(ta, tb, alpha, A, B, beta, C) for (m, n, k, A, B, C) in probs]
            choices = jb.grouped_choices("f64", descs)
            assert all(c[0] == tile for c in choices), choices
            jb.gemm_grouped_("f64", descs)
            torch.cuda.synchronize()
            for i, (m, n, k, A, B, C) in enumerate(probs):
                ref = O.gemm_recurrence_chunked(ta, tb, alpha, to_host(A), to_host(B), beta, C0[i], choices[i][2])
                got = to_host(C)
                assert np.array_equal(got.view(np.int64), ref.view(np.int64)), (ta, tb, m, n, k, choices[i], alpha,
                                                                               beta)
    finally:
        jb.tune_reset()


def strided32(torch, rows, cols, seed, ld, offset):
    """Float32 `strided`: ld a multiple of 4 at offset 0 gives 16-byte vectors,
    an odd ld or offset element loads."""
    buf = torch.empty(offset + ld * max(cols, 1), dtype=torch.float32, device="cuda")
    jb.fill_uniform_(buf, seed)
    return buf[offset:offset + ld * cols].view(cols, ld).T[:rows] if cols else buf[:0].view(0, rows).T


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("tile,splitk", [("128x128", 1), ("128x128", 3), ("128x64", 1), ("128x64", 2),
                                         ("64x128", 1), ("64x128", 3)])
def test_grouped_bf16x3_large_tiles(gpu, ta, tb, tile, splitk, tmp_path):
    """The 3xBF16 tiles with 128 rows or columns (a persistent kernel per tile
    shape and transpose combination; 128x128 with eight waves), forced through
    a tuning table on fringe shapes (partial tiles in m, n and k; k == 0), with
    the full-tile 16-byte loads (aligned bases, leading dimensions a multiple of
    4) and the element-wise ones (odd leading dimension or offset), one split or
    more: within 2*K*2^-24 componentwise of the FP64 (Dot2) oracle, and
    beta*C for k == 0."""
    torch = gpu
    jb.tune_reset()
    cfg = {v: k for k, v in jb.grouped.TILES.items()}[tile]
    write_table(str(tmp_path / "large3.json"), "f32_bf16x3", LARGE_SHAPES, cfg, splitk)
    try:
        probs = []
        for i, (m, n, k) in enumerate(LARGE_SHAPES):
            ra, ca = (m, k) if ta == "N" else (k, m)
            rb, cb = (k, n) if tb == "N" else (n, k)
            odd = i % 2 == 1
            A = strided32(torch, ra, ca, 400 + 3 * i, ra + 1 if odd else -(-ra // 4) * 4, 1 if odd else 0)
            B = strided32(torch, rb, cb, 401 + 3 * i, rb + 3 if odd else -(-rb // 4) * 4 + 4, 0)
            C = strided32(torch, m, n, 402 + 3 * i, m + 5, 3)
            probs.append((m, n, k, A, B, C))
        C0 = [to_host(p[5]) for p in probs]
        for alpha, beta in [(1.5, 0.5), (-1.0, 0.0)]:
            for i, p in enumerate(probs):
                p[5].copy_(to_dev(torch, C0[i]))
            descs = [GemmDesc(ta, tb, alpha, A, B, beta, C) for (m, n, k, A, B, C) in probs]
            choices = jb.grouped_choices("f32_bf16x3", descs)
            assert all(c[0] == tile for c in choices), choices
            jb.gemm_grouped_("f32_bf16x3", descs)
            torch.cuda.synchronize()
            for i, (m, n, k, A, B, C) in enumerate(probs):
                Ah, Bh, got = to_host(A), to_host(B), to_host(C).astype(float)
                ref = O.sgemm_dot2(ta, tb, alpha, Ah, Bh, beta, C0[i])
                Ao = Ah if ta == "N" else Ah.T
                Bo = Bh if tb == "N" else Bh.T
                bound = (abs(alpha) * (np.abs(Ao).astype(float) @ np.abs(Bo).astype(float))
                         + abs(beta) * np.abs(C0[i]).astype(float))
                if k == 0:  # C = beta*C exactly (+0 for beta == 0)
                    assert np.array_equal(got, (np.float32(beta) * C0[i]).astype(float)), (ta, tb, m, n, alpha, beta)
                else:
                    err = np.max(np.abs(got - ref) / bound)
                    assert err <= 2 * k * U32, (ta, tb, m, n, k, choices[i], alpha, beta, err)
    finally:
        jb.tune_reset()


def test_grouped_concurrent_launches_mixed_ops(gpu):
    """One group whose problems need every persistent kernel at once: FP64
    problems of all four transpose combinations on the 128x128, 128x64 and
    64x128 tiles (twelve kernels) plus small-path and 32x32 problems, few
    enough that the plan also splits k to fill the device; the launches run
    concurrently on forked streams and join back.  Bit-exact per problem
    (eager launch and graph replays)."""
    torch = gpu
    jb.tune_reset()
    base = [(1024, 1024, 700), (1024, 64, 900), (64, 1024, 600), (40, 50, 60), (80, 90, 300)]
    probs = []
    for i in range(4 * len(base)):
        (m, n, k), (ta, tb) = base[i % len(base)], [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")][i // len(base)]
        A = dev_uniform(torch, m if ta == "N" else k, k if ta == "N" else m, 500 + 3 * i, torch.float64)
        B = dev_uniform(torch, k if tb == "N" else n, n if tb == "N" else k, 501 + 3 * i, torch.float64)
        C = dev_uniform(torch, m, n, 502 + 3 * i, torch.float64)
        probs.append((ta, tb, m, n, k, A, B, C))
    descs = [GemmDesc(ta, tb, 1.5, A, B, 0.5, C) for (ta, tb, m, n, k, A, B, C) in probs]
    choices = jb.grouped_choices("f64", descs)
    kinds = {c[0] for c in choices}
    assert {"128x128", "128x64", "64x128", "small"} <= kinds, kinds
    assert any(c[1] > 1 for c in choices), choices  # device-fill split-K took part
    C0 = [to_host(p[7]) for p in probs]
    refs = [O.gemm_recurrence_chunked(ta, tb, 1.5, to_host(A), to_host(B), 0.5, C0[i], choices[i][2])
            for i, (ta, tb, m, n, k, A, B, C) in enumerate(probs)]
    jb.gemm_grouped_("f64", descs)
    torch.cuda.synchronize()
    for i, p in enumerate(probs):
        assert np.array_equal(to_host(p[7]), refs[i]), (i, p[:5], choices[i])
    plan = jb.GroupedPlan("f64", descs)
    for rep in range(2):
        for i, p in enumerate(probs):
            p[7].copy_(to_dev(torch, C0[i]))
        plan.launch()
        torch.cuda.synchronize()
        for i, p in enumerate(probs):
            assert np.array_equal(to_host(p[7]), refs[i]), (rep, i, p[:5], choices[i])
    plan.destroy()


@pytest.mark.parametrize("tile", ["64x64", "32x32", "128x128", "128x64", "64x128"])
def test_grouped_bf16x3_flt_max_and_nonfinite(gpu, tile, tmp_path):
    """3xBF16 tiles split entries up to FLT_MAX exactly (the first plane
    truncates where rounding to nearest would give Inf), so products with small
    partners stay finite and within 2*K*2^-24 of the FP64 oracle; an Inf or NaN
    operand entry makes the row or column of C it enters NaN (as the
    single-problem 3xBF16 split does) and leaves the rest finite."""
    torch = gpu
    jb.tune_reset()
    cfg = {v: k for k, v in jb.grouped.TILES.items()}[tile]
    m, n, k = 100, 80, 96
    write_table(str(tmp_path / "big3.json"), "f32_bf16x3", [(m, n, k)], cfg, 1)
    try:
        A = G.matrix(31, m, k, dtype=np.float32)
        B = G.matrix(32, k, n, dtype=np.float32) * np.float32(2.0 ** -100)
        A[:, 0] = np.float32(3.4028235e38)   # FLT_MAX
        A[:, 1] = np.float32(-3.4e38)
        A[:, 2] = np.float32(3.3962e38)      # nearest BF16 is Inf
        A[5, 3] = np.float32(3.39e38)        # just below: nearest BF16 is finite
        Ad, Bd = to_dev(torch, A), to_dev(torch, B)
        Cd = to_dev(torch, np.zeros((m, n), dtype=np.float32))
        desc = [GemmDesc("N", "N", 1.0, Ad, Bd, 0.0, Cd)]
        assert jb.grouped_choices("f32_bf16x3", desc)[0][0] == tile
        jb.gemm_grouped_("f32_bf16x3", desc)
        torch.cuda.synchronize()
        got = to_host(Cd).astype(np.float64)
        ref = A.astype(np.float64) @ B.astype(np.float64)
        absab = np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)
        assert np.all(np.isfinite(got)), tile
        assert np.max(np.abs(got - ref) / absab) <= 2 * k * U32, tile
        A[7, 5], B[9, 30] = np.inf, np.nan
        Ad, Bd = to_dev(torch, A), to_dev(torch, B)
        desc = [GemmDesc("N", "N", 1.0, Ad, Bd, 0.0, Cd)]
        jb.gemm_grouped_("f32_bf16x3", desc)
        torch.cuda.synchronize()
        got = to_host(Cd)
        bad = np.zeros((m, n), bool)
        bad[7, :] = True
        bad[:, 30] = True
        assert np.all(np.isnan(got[bad])) and np.all(np.isfinite(got[~bad])), tile
    finally:
        jb.tune_reset()


def test_grouped_bf16x3_concurrent_launches(gpu, tmp_path):
    """3xBF16 group whose problems need every persistent kernel at once: a
    tuning table puts one size bucket each on 128x128 (eight-wave workgroups,
    two splits), 128x64, 64x128, 64x64, 32x32 and the small path, in all four
    transpose combinations; the launches run concurrently on forked streams.
    Every problem within 2*K*2^-24 componentwise of the FP64 (Dot2) oracle;
    graph replays bit-identical to the eager launch."""
    torch = gpu
    jb.tune_reset()
    base = [((1024, 1024, 700), 3, 2), ((1024, 64, 900), 4, 1), ((64, 1024, 300), 5, 1), ((200, 200, 100), 0, 1),
            ((80, 90, 300), 1, 1), ((40, 50, 60), 2, 1)]
    info = jb.device_info()
    pre = f"{info['name']}|{info['arch']}|m{jb.grouped.method_id('f32_bf16x3')}|"
    entries = ",\n".join(
        f'{{"key": "{pre}mn{max(math.isqrt(m * n), 1).bit_length() - 1}|k{max(k, 1).bit_length() - 1}", '
        f'"cfg": {cfg}, "splitk": {sk}, "ms": 0.0}}' for (m, n, k), cfg, sk in base)
    path = tmp_path / "mixed3.json"
    path.write_text('{"format": "jblas-tune-1", "entries": [' + entries + "]}")
    jb.tune_load(str(path))
    try:
        probs = []
        for i in range(4 * len(base)):
            (m, n, k), (ta, tb) = base[i % len(base)][0], [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")][i // len(base)]
            A = dev_uniform(torch, m if ta == "N" else k, k if ta == "N" else m, 700 + 3 * i, torch.float32)
            B = dev_uniform(torch, k if tb == "N" else n, n if tb == "N" else k, 701 + 3 * i, torch.float32)
            C = dev_uniform(torch, m, n, 702 + 3 * i, torch.float32)
            probs.append((ta, tb, m, n, k, A, B, C))
        descs = [GemmDesc(ta, tb, 1.5, A, B, 0.5, C) for (ta, tb, m, n, k, A, B, C) in probs]
        choices = jb.grouped_choices("f32_bf16x3", descs)
        want = {"128x128": 2, "128x64": 1, "64x128": 1, "64x64": 1, "32x32": 1, "small": 1}
        assert [(c[0], c[1]) for c in choices[:len(base)]] == list(want.items()), choices
        C0 = [to_host(p[7]) for p in probs]
        jb.gemm_grouped_("f32_bf16x3", descs)
        torch.cuda.synchronize()
        eager = [to_host(p[7]) for p in probs]
        for i, (ta, tb, m, n, k, A, B, C) in enumerate(probs):
            Ah, Bh = to_host(A), to_host(B)
            ref = O.sgemm_dot2(ta, tb, 1.5, Ah, Bh, 0.5, C0[i])
            Ao = Ah if ta == "N" else Ah.T
            Bo = Bh if tb == "N" else Bh.T
            bound = 1.5 * (np.abs(Ao).astype(float) @ np.abs(Bo).astype(float)) + 0.5 * np.abs(C0[i]).astype(float)
            err = np.max(np.abs(eager[i].astype(float) - ref) / bound)
            assert err <= 2 * k * U32, (i, ta, tb, m, n, k, choices[i], err)
        plan = jb.GroupedPlan("f32_bf16x3", descs)
        for rep in range(2):
            for i, p in enumerate(probs):
                p[7].copy_(to_dev(torch, C0[i]))
            plan.launch()
            torch.cuda.synchronize()
            for i, p in enumerate(probs):
                assert np.array_equal(to_host(p[7]), eager[i]), (rep, i, p[:5], choices[i])
        plan.destroy()
    finally:
        jb.tune_reset()


SMALL_SHAPES = [(1, 1, 1), (64, 64, 64), (33, 17, 5), (7, 64, 63), (64, 3, 1), (50, 40, 64)]
TILE_SHAPES = [(200, 72, 96), (16, 16, 2048)]


def small_group(torch, ta, tb, dtype, seed):
    probs = []
    for i, (m, n, k) in enumerate(SMALL_SHAPES + TILE_SHAPES):
        A = dev_uniform(torch, m if ta == "N" else k, k if ta == "N" else m, seed + 3 * i, dtype)
        B = dev_uniform(torch, k if tb == "N" else n, n if tb == "N" else k, seed + 1 + 3 * i, dtype)
        C = dev_uniform(torch, m, n, seed + 2 + 3 * i, dtype)
        probs.append((m, n, k, A, B, C))
    return probs


def load_small_table(tmpdir, method, shapes):
    """A tuning table naming the small path for the size buckets of `shapes`.
    (The static rule picks it for these shapes too; the table also covers the
    bucket of a problem that does not fit it.)"""
    write_table(os.path.join(tmpdir, "small.json"), method, shapes, 2, 1)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_grouped_small_path_f64_bitexact(gpu, ta, tb, tmp_path):
    """m, n, k <= 64 with a table entry for the small path: the whole problem
    in one work item on the vector ALUs (fastmul!'s form), launched together
    with problems on MFMA tiles; FP64 results equal the reference recurrence
    (plain first product, then fma in k order) bit for bit.  A problem of a
    small-path bucket that does not fit (m = 65) takes the 32x32 tile."""
    torch = gpu
    jb.tune_reset()
    load_small_table(str(tmp_path), "f64", SMALL_SHAPES)
    assert jb.tune_lookup("f64", 65, 63, 64)[0] == "32x32"  # bucket of (50, 40, 64), m too large
    probs = small_group(torch, ta, tb, torch.float64, 70 + ord(ta) + 2 * ord(tb))
    C0 = [to_host(p[5]) for p in probs]
    for alpha, beta in [(1.5, 0.5), (-1.0, 0.0), (0.0, 2.0)]:
        for i, p in enumerate(probs):
            p[5].copy_(to_dev(torch, C0[i]))
        descs = [GemmDesc(ta, tb, alpha, A, B, beta, C) for (m, n, k, A, B, C) in probs]
        kinds = jb.grouped_choices("f64", descs)
        assert [c[0] for c in kinds[:len(SMALL_SHAPES)]] == ["small"] * len(SMALL_SHAPES)
        if alpha != 0.0:  # alpha == 0 means no products (k taken as 0): every problem fits the small path
            assert all(c[0] != "small" for c in kinds[len(SMALL_SHAPES):]), kinds
        jb.gemm_grouped_("f64", descs)
        torch.cuda.synchronize()
        for i, (m, n, k, A, B, C) in enumerate(probs):
            ref = O.gemm_recurrence_chunked(ta, tb, alpha, to_host(A), to_host(B), beta, C0[i], kinds[i][2])
            assert np.array_equal(to_host(C), ref), (ta, tb, m, n, k, kinds[i], alpha, beta)
    jb.tune_reset()


@pytest.mark.parametrize("method", ["f32", "f32_bf16x3"])
def test_grouped_small_path_f32(gpu, method, tmp_path):
    """FP32 methods take the same small path (an FP32 fma chain) from a table entry."""
    torch = gpu
    jb.tune_reset()
    load_small_table(str(tmp_path), method, SMALL_SHAPES)
    assert all(jb.tune_lookup(method, m, n, k)[0] == "small" for m, n, k in SMALL_SHAPES)
    for ta, tb in [("N", "N"), ("T", "T")]:
        probs = small_group(torch, ta, tb, torch.float32, 90 + ord(ta))
        C0 = [to_host(p[5]) for p in probs]
        jb.gemm_grouped_(method, [GemmDesc(ta, tb, 1.5, A, B, 0.5, C) for (m, n, k, A, B, C) in probs])
        torch.cuda.synchronize()
        for i, (m, n, k, A, B, C) in enumerate(probs):
            Ah, Bh = to_host(A), to_host(B)
            ref = O.sgemm_dot2(ta, tb, 1.5, Ah, Bh, 0.5, C0[i])
            Ao = Ah if ta == "N" else Ah.T
            Bo = Bh if tb == "N" else Bh.T
            absab = 1.5 * (np.abs(Ao).astype(float) @ np.abs(Bo).astype(float)) + 0.5 * np.abs(C0[i])
            err = np.max(np.abs(to_host(C).astype(float) - ref) / absab)
            assert err <= 2 * k * U32, (method, ta, tb, m, n, k, err)
    jb.tune_reset()


def test_grouped_f64_bits_with_signed_zeros(gpu):
    """FP64 grouped results equal the oracle's (chunked) recurrence in every
    bit, the sign of zeros included (int64 views), on inputs full of +-0:
    MFMA tiles, the small path, split-K, and problems with k == 0 or alpha == 0
    (C = beta*C, +0 for beta == 0), in one launch."""
    torch = gpu
    rng = np.random.default_rng(31)
    shapes = [(100, 70, 37, 1.5, 0.5), (40, 30, 20, 1.0, 0.0), (16, 16, 2048, -1.0, 0.0), (130, 65, 47, 1.0, 1.0),
              (20, 10, 0, 2.0, -0.5), (25, 12, 9, 0.0, 0.0), (25, 12, 9, 0.0, -1.0)]
    probs = []
    for (m, n, k, alpha, beta) in shapes:
        A = np.asfortranarray(rng.choice([-2.0, -1.0, -0.0, 0.0, 1.0], (m, k), p=[.15, .15, .3, .3, .1]))
        B = np.asfortranarray(rng.choice([-1.0, -0.0, 0.0, 3.0], (k, n), p=[.2, .3, .3, .2]))
        C = np.asfortranarray(rng.choice([-0.0, 0.0, 2.0], (m, n), p=[.4, .4, .2]))
        probs.append((m, n, k, alpha, beta, A, B, C, dev_copy(torch, A), dev_copy(torch, B), dev_copy(torch, C)))
    descs = [GemmDesc("N", "N", p[3], p[8], p[9], p[4], p[10]) for p in probs]
    kc = kchunks("f64", descs)
    jb.gemm_grouped_("f64", descs)
    torch.cuda.synchronize()
    for i, (m, n, k, alpha, beta, A, B, C, Ad, Bd, Cd) in enumerate(probs):
        ref = O.gemm_recurrence_chunked("N", "N", alpha, A, B, beta, C, kc[i])
        assert np.array_equal(to_host(Cd).view(np.int64), ref.view(np.int64)), (m, n, k, alpha, beta, kc[i])


def test_grouped_plan_graph_replays(gpu):
    torch = gpu
    probs, descs = make_group(torch, 96, 53, torch.float64, alpha=1.0, beta=1.0)
    C0 = [p[7].clone() for p in probs]
    jb.gemm_grouped_("f64", descs)  # eager reference: C1 = A*B + C0
    torch.cuda.synchronize()
    C1 = [p[7].clone() for p in probs]
    plan = jb.GroupedPlan("f64", descs)
    for rep in range(3):
        for i, p in enumerate(probs):
            p[7].copy_(C0[i])
        junk = torch.randn(512, 512, device="cuda") @ torch.randn(512, 512, device="cuda")  # eager work between
        plan.launch()
        torch.cuda.synchronize()
        del junk
        for i, p in enumerate(probs):
            assert torch.equal(p[7], C1[i]), (rep, i)
    plan.destroy()


def test_grouped_bf16x3_c4_config_set(gpu):
    """Config: the 1024-problem set (m, n, k log-uniform in [16, 2048], seed 52)
    in Float32 through the 3xBF16 split, one persistent launch; every 8th
    problem within 2*K*2^-24 componentwise of the FP64 (Dot2) oracle."""
    torch = gpu
    probs, descs = make_group(torch, 1024, 52, torch.float32)
    C0 = [to_host(p[7]) for p in probs[::8]]
    jb.gemm_grouped_("f32_bf16x3", descs)
    torch.cuda.synchronize()
    worst = 0.0
    for idx, (ta, tb, m, n, k, A, B, C) in enumerate(probs[::8]):
        Ah, Bh = to_host(A), to_host(B)
        ref = O.sgemm_dot2(ta, tb, 1.0, Ah, Bh, 0.0, C0[idx])
        absab = np.abs(Ah).astype(float) @ np.abs(Bh).astype(float)
        err = np.max(np.abs(to_host(C).astype(float) - ref) / absab)
        worst = max(worst, err / k)
        assert err <= 2 * k * U32, (m, n, k, err)
    print(f"max componentwise error / k over 128 problems: {worst:.3e} (bar 2^-23 = {2 * U32:.3e})")


@pytest.mark.parametrize("method", ["f32", "f32_bf16x3"])
def test_grouped_f32_tolerance(gpu, method):
    torch = gpu
    probs, descs = make_group(torch, 256, 54, torch.float32, ops=[("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")],
                              alpha=1.5, beta=0.5)
    C0 = [to_host(p[7]) for p in probs[::8]]
    jb.gemm_grouped_(method, descs)
    torch.cuda.synchronize()
    for idx, (ta, tb, m, n, k, A, B, C) in enumerate(probs[::8]):
        Ah, Bh = to_host(A), to_host(B)
        ref = O.sgemm_dot2(ta, tb, 1.5, Ah, Bh, 0.5, C0[idx])
        Ao = Ah if ta == "N" else Ah.T
        Bo = Bh if tb == "N" else Bh.T
        absab = 1.5 * (np.abs(Ao).astype(float) @ np.abs(Bo).astype(float)) + 0.5 * np.abs(C0[idx])
        err = np.max(np.abs(to_host(C).astype(float) - ref) / absab)
        assert err <= 2 * k * U32, (method, m, n, k, err)


def test_grouped_edge_cases(gpu):
    torch = gpu
    A = dev_uniform(torch, 40, 30, 1, torch.float64)
    B = dev_uniform(torch, 30, 20, 2, torch.float64)
    C = dev_uniform(torch, 40, 20, 3, torch.float64)
    Cz = dev_uniform(torch, 40, 20, 4, torch.float64)
    C0, Cz0 = to_host(C), to_host(Cz)
    empty_A = torch.empty((0, 40), dtype=torch.float64, device="cuda").T  # 40 x 0
    empty_B = torch.empty((20, 0), dtype=torch.float64, device="cuda").T  # 0 x 20
    Ck = dev_uniform(torch, 40, 20, 5, torch.float64)
    Ck0 = to_host(Ck)
    big = dev_uniform(torch, 64, 64, 6, torch.float64)
    descs = [
        GemmDesc("N", "N", 0.0, A, B, 0.5, C),          # alpha = 0: beta*C
        GemmDesc("N", "N", 2.0, empty_A, empty_B, -1.0, Ck),  # k = 0
        GemmDesc("N", "N", 1.0, A, B, 0.0, Cz),         # beta = 0
        GemmDesc("N", "N", 1.0, big[:10, :7], big[20:27, 30:39], 1.0, big[40:50, 50:59]),  # strided views
    ]
    big0 = to_host(big)
    jb.gemm_grouped_("f64", descs)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(C), 0.5 * C0)
    assert np.array_equal(to_host(Ck), -Ck0)
    assert np.array_equal(to_host(Cz), O.gemm_recurrence("N", "N", 1.0, to_host(A), to_host(B), 0.0, Cz0))
    out = to_host(big)
    ref = O.gemm_recurrence("N", "N", 1.0, big0[:10, :7], big0[20:27, 30:39], 1.0, big0[40:50, 50:59])
    assert np.array_equal(out[40:50, 50:59], ref)
    mask = np.ones_like(out, dtype=bool)
    mask[40:50, 50:59] = False
    assert np.array_equal(out[mask], big0[mask])


def test_autotune_roundtrip(gpu):
    torch = gpu
    probs, descs = make_group(torch, 64, 55, torch.float64)
    jb.tune_reset()
    before = [jb.tune_lookup("f64", m, n, k) for (_, _, m, n, k, *_r) in probs]
    jb.autotune("f64", descs, reps=1)
    tuned = [jb.tune_lookup("f64", m, n, k) for (_, _, m, n, k, *_r) in probs]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tune.json")
        jb.tune_save(path)
        text = open(path).read()
        assert "jblas-tune-1" in text
        jb.tune_reset()
        assert [jb.tune_lookup("f64", m, n, k) for (_, _, m, n, k, *_r) in probs] == before
        jb.tune_load(path)
        assert [jb.tune_lookup("f64", m, n, k) for (_, _, m, n, k, *_r) in probs] == tuned
        with open(path, "w") as f:
            f.write("not a table")
        with pytest.raises(jb.JBError):
            jb.tune_load(path)
    # the tuned table still gives bit-exact results
    C0 = [to_host(p[7]) for p in probs]
    kc = kchunks("f64", descs)
    jb.gemm_grouped_("f64", descs)
    torch.cuda.synchronize()
    for idx, (ta, tb, m, n, k, A, B, C) in enumerate(probs[::4]):
        ref = O.gemm_recurrence_chunked("N", "N", 1.0, to_host(A), to_host(B), 0.0, C0[4 * idx], kc[4 * idx])
        assert np.array_equal(to_host(C), ref)
    jb.tune_reset()


def test_autotune_leaves_c_untouched(gpu):
    """The tuner's trials write to scratch outputs: a beta != 0 group's C is the
    same after tuning, and one launch afterwards computes alpha*A*B + beta*C once."""
    torch = gpu
    probs, descs = make_group(torch, 24, 57, torch.float64, alpha=1.5, beta=0.5)
    C0 = [to_host(p[7]) for p in probs]
    jb.tune_reset()
    jb.autotune("f64", descs, reps=2)
    torch.cuda.synchronize()
    for p, c0 in zip(probs, C0):
        assert np.array_equal(to_host(p[7]), c0)
    kc = kchunks("f64", descs)
    jb.gemm_grouped_("f64", descs)
    torch.cuda.synchronize()
    for idx, (ta, tb, m, n, k, A, B, C) in enumerate(probs):
        ref = O.gemm_recurrence_chunked("N", "N", 1.5, to_host(A), to_host(B), 0.5, C0[idx], kc[idx])
        assert np.array_equal(to_host(C), ref)
    jb.tune_reset()


def test_grouped_fault_injection(gpu):
    torch = gpu
    probs, descs = make_group(torch, 4, 56, torch.float64)
    os.environ["JB_FAULT_GRAPH"] = "1"
    try:
        with pytest.raises(jb.JBError) as ei:
            jb.GroupedPlan("f64", descs)
        assert ei.value.code == 4
    finally:
        del os.environ["JB_FAULT_GRAPH"]
    os.environ["JB_FAULT_ALLOC"] = "1"
    try:
        with pytest.raises(jb.JBError) as ei:
            jb.gemm_grouped_("f64", descs)
        assert ei.value.code == 2
    finally:
        del os.environ["JB_FAULT_ALLOC"]
    plan = jb.GroupedPlan("f64", descs)
    plan.launch()
    torch.cuda.synchronize()
    plan.destroy()
    with pytest.raises(ValueError):
        jb.gemm_grouped_("f64", [GemmDesc("N", "N", 1.0, probs[0][5], probs[1][6], 0.0, probs[0][7])])


# ---------------------------------------------------------------------------
# rrule / pullback
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("method", ["f32", "f32_bf16x3"])
def test_pullback_c5(gpu, method):
    """Config: F32 A 4096x2048, B 2048x4096 ~ N(0,1)/sqrt(2048), W ~ N(0,1),
    drawn in HBM by the device normal generator (no host copy; the checker
    reads the device data back); dA = W*B^T, dB = A^T*W checked on sampled
    rows/columns against the FP64 oracle."""
    torch = gpu
    M, K, N = 4096, 2048, 4096
    A = jb.fill_normal_(torch.empty((K, M), dtype=torch.float32, device="cuda"), 42, 1.0 / np.sqrt(2048)).T
    B = jb.fill_normal_(torch.empty((N, K), dtype=torch.float32, device="cuda"), 43, 1.0 / np.sqrt(2048)).T
    W = jb.fill_normal_(torch.empty((N, M), dtype=torch.float32, device="cuda"), 45).T
    dA = jb.device_matrix(M, K, torch.float32)
    dB = jb.device_matrix(K, N, torch.float32)
    from jblas_amd.rrule import pullback_

    pullback_("N", "N", 1.0, A, B, W, dA, dB, method=method)
    torch.cuda.synchronize()
    rows = np.random.default_rng(0).choice(M, 24, replace=False)
    cols = np.random.default_rng(1).choice(N, 24, replace=False)
    Ah, Bh, Wh = to_host(A), to_host(B), to_host(W)
    # dA rows: W[rows, :] * B^T  (k = N)
    ref_a = O.sgemm_dot2("N", "T", 1.0, Wh[rows, :], Bh, 0.0, np.zeros((24, K), np.float32))
    abs_a = np.abs(Wh[rows, :]).astype(float) @ np.abs(Bh.T).astype(float)
    got_a = to_host(dA)[rows, :].astype(float)
    assert np.max(np.abs(got_a - ref_a) / abs_a) <= 2 * N * U32
    # dB columns: A^T * W[:, cols]  (k = M)
    ref_b = O.sgemm_dot2("T", "N", 1.0, Ah, Wh[:, cols], 0.0, np.zeros((K, 24), np.float32))
    abs_b = np.abs(Ah.T).astype(float) @ np.abs(Wh[:, cols]).astype(float)
    got_b = to_host(dB)[:, cols].astype(float)
    assert np.max(np.abs(got_b - ref_b) / abs_b) <= 2 * M * U32


def test_rrule_fp64_finite_differences_and_cpu_gradient(gpu):
    """64^3 sub-problem in FP64: loss = sum(W .* (A*B)); FP64 central differences,
    h = 1e-6, and the oracle's recurrence for W*B^T / A^T*W (bit-exact)."""
    torch = gpu
    n = 64
    A = G.matrix(42, n, n, kind="normal") / np.sqrt(2048)
    B = G.matrix(43, n, n, kind="normal") / np.sqrt(2048)
    W = G.matrix(45, n, n, kind="normal")
    from jblas_amd.rrule import rrule_gemm

    Y, back = rrule_gemm("N", "N", 1.0, to_dev(torch, A), to_dev(torch, B), 0.0, to_dev(torch, np.zeros((n, n))))
    dA, dB, dalpha, dbeta, dC = back(to_dev(torch, W))
    torch.cuda.synchronize()
    dA, dB = to_host(dA), to_host(dB)
    assert np.array_equal(dA, O.gemm_recurrence("N", "T", 1.0, W, B, 0.0, np.zeros((n, n))))
    assert np.array_equal(dB, O.gemm_recurrence("T", "N", 1.0, A, W, 0.0, np.zeros((n, n))))

    def loss(Am, Bm):
        return float(np.sum(W * (Am @ Bm)))

    h = 1e-6
    rng = np.random.default_rng(3)
    L = abs(loss(A, B))
    for (i, j) in rng.integers(0, n, (40, 2)):
        E = np.zeros((n, n))
        E[i, j] = h
        fd_a = (loss(A + E, B) - loss(A - E, B)) / (2 * h)
        fd_b = (loss(A, B + E) - loss(A, B - E)) / (2 * h)
        tol = 1e-7 * max(1.0, L)
        assert abs(fd_a - dA[i, j]) <= tol and abs(fd_b - dB[i, j]) <= tol
    Yh = to_host(Y)
    assert abs(float(dalpha) - np.sum(W * Yh)) <= 1e-12 * np.sum(np.abs(W * Yh))


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_rrule_ops_closed_form(gpu, ta, tb):
    torch = gpu
    m, n, k = 50, 37, 29
    A = G.matrix(1, m if ta == "N" else k, k if ta == "N" else m)
    B = G.matrix(2, k if tb == "N" else n, n if tb == "N" else k)
    C = G.matrix(3, m, n)
    dY = G.matrix(4, m, n)
    from jblas_amd.rrule import rrule_gemm

    Y, back = rrule_gemm(ta, tb, 1.5, to_dev(torch, A), to_dev(torch, B), 0.5, to_dev(torch, C))
    dA, dB, dalpha, dbeta, dC = back(to_dev(torch, dY))
    torch.cuda.synchronize()
    Ao, Bo = (A if ta == "N" else A.T), (B if tb == "N" else B.T)
    ref_dAo = 1.5 * dY @ Bo.T
    ref_dBo = 1.5 * Ao.T @ dY
    ref_dA = ref_dAo if ta == "N" else ref_dAo.T
    ref_dB = ref_dBo if tb == "N" else ref_dBo.T
    assert np.allclose(to_host(dA), ref_dA, rtol=0, atol=1e-12)
    assert np.allclose(to_host(dB), ref_dB, rtol=0, atol=1e-12)
    assert np.allclose(to_host(Y), 1.5 * Ao @ Bo + 0.5 * C, rtol=0, atol=1e-12)
    assert abs(float(dalpha) - np.sum(dY * (Ao @ Bo))) < 1e-9
    assert abs(float(dbeta) - np.sum(dY * C)) < 1e-9
    assert np.array_equal(to_host(dC), 0.5 * dY)


def test_batched_backward_256(gpu):
    """Config: batched backward over 256 uneven shapes (C4 distribution), FP32 3xBF16."""
    torch = gpu
    from jblas_amd.rrule import pullback_grouped_

    shapes = G.loguniform_shapes(62, 256)
    items, host = [], []
    for i, (m, n, k) in enumerate(shapes):
        A = dev_uniform(torch, int(m), int(k), 5000 + 4 * i, torch.float32)
        B = dev_uniform(torch, int(k), int(n), 5001 + 4 * i, torch.float32)
        dY = dev_uniform(torch, int(m), int(n), 5002 + 4 * i, torch.float32)
        dA = jb.device_matrix(int(m), int(k), torch.float32)
        dB = jb.device_matrix(int(k), int(n), torch.float32)
        items.append(("N", "N", 1.0, A, B, dY, dA, dB))
    pullback_grouped_(items, method="f32_bf16x3")
    torch.cuda.synchronize()
    for (ta, tb, al, A, B, dY, dA, dB) in items[::16]:
        Ah, Bh, dYh = to_host(A), to_host(B), to_host(dY)
        k = Ah.shape[1]
        ref_a = O.sgemm_dot2("N", "T", 1.0, dYh, Bh, 0.0, np.zeros(Ah.shape, np.float32))
        abs_a = np.abs(dYh).astype(float) @ np.abs(Bh.T).astype(float)
        assert np.max(np.abs(to_host(dA) - ref_a) / abs_a) <= 2 * dYh.shape[1] * U32
        ref_b = O.sgemm_dot2("T", "N", 1.0, Ah, dYh, 0.0, np.zeros(Bh.shape, np.float32))
        abs_b = np.abs(Ah.T).astype(float) @ np.abs(dYh).astype(float)
        assert np.max(np.abs(to_host(dB) - ref_b) / abs_b) <= 2 * Ah.shape[0] * U32


def test_torch_autograd_binding(gpu):
    torch = gpu
    from jblas_amd.rrule import matmul

    A = torch.randn(70, 40, dtype=torch.float64, device="cuda", requires_grad=True)
    B = torch.randn(40, 30, dtype=torch.float64, device="cuda", requires_grad=True)
    W = torch.randn(70, 30, dtype=torch.float64, device="cuda")
    (matmul(A, B) * W).sum().backward()
    Ac, Bc = A.detach().cpu().requires_grad_(), B.detach().cpu().requires_grad_()
    (Ac @ Bc * W.cpu()).sum().backward()
    assert torch.allclose(A.grad.cpu(), Ac.grad, rtol=0, atol=1e-12)
    assert torch.allclose(B.grad.cpu(), Bc.grad, rtol=0, atol=1e-12)


@pytest.mark.parametrize("tile", ["128x128", "128x64", "64x128", "64x64", "32x32", "small"])
def test_grouped_f64_nonfinite_matches_recurrence(gpu, tile, tmp_path):
    """Inf, -Inf and NaN operand entries and overflowing products in a grouped
    FP64 launch, the tile forced through a tuning table, two splits on the MFMA
    tiles: bit-exact against the chunked recurrence where the result is a
    number or an infinity, NaN where it is NaN (payloads not compared)."""
    torch = gpu
    jb.tune_reset()
    m, n, k = (60, 50, 64) if tile == "small" else (150, 90, 120)
    cfg = {v: c for c, v in jb.grouped.TILES.items()}[tile]
    write_table(str(tmp_path / "nf.json"), "f64", [(m, n, k)], cfg, 1 if tile == "small" else 2)
    try:
        A = G.matrix(81, m, k)
        B = G.matrix(82, k, n)
        C0 = G.matrix(83, m, n)
        A[3, 7], A[40, 0], A[m - 1, k - 1] = np.inf, -np.inf, np.nan
        B[5, 11], B[k - 1, n - 1] = np.nan, -np.inf
        A[30, :] *= 1e300
        B[:, 20] *= 1e10
        Ad, Bd, Cd = dev_copy(torch, A), dev_copy(torch, B), dev_copy(torch, C0)
        desc = [GemmDesc("N", "N", 1.5, Ad, Bd, 0.5, Cd)]
        ch = jb.grouped_choices("f64", desc)[0]
        assert ch[0] == tile, ch
        jb.gemm_grouped_("f64", desc)
        torch.cuda.synchronize()
        got = to_host(Cd)
        ref = O.gemm_recurrence_chunked("N", "N", 1.5, A, B, 0.5, C0, ch[2])
        nan_ref, nan_got = np.isnan(ref), np.isnan(got)
        assert np.array_equal(nan_ref, nan_got), (tile, np.argwhere(nan_ref != nan_got)[:5])
        assert np.array_equal(got[~nan_ref].view(np.int64), ref[~nan_ref].view(np.int64)), tile
        assert np.isinf(ref).any() and nan_ref.any()
    finally:
        jb.tune_reset()
