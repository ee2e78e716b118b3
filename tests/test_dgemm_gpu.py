"""GPU parity of the native FP64/FP32 MFMA GEMM and the bit-exact recurrence
kernel, through the C ABI, against the oracle (oracle/oracle.c).

Bars: the jmul! recurrence and the FP64 MFMA path are compared BIT-FOR-BIT
with the oracle's restatement of the reference recurrence (the f64 MFMA is a
sequential fma chain in k order); against the Dot2 FP64 oracle the stated
tolerance is the componentwise bound
    |C - C*|_ij <= 2*K*u * (|alpha| |op(A)||op(B)| + |beta| |C0|)_ij,  u = 2^-53.
"""
import math
import os

import numpy as np
import pytest

import jblas_amd as jb
from jblas_amd import generators as G
from oracle import oracle as O

from conftest import to_dev, to_host

pytestmark = pytest.mark.gpu
U = 2.0 ** -53


def opnd(x, t):
    return x if t == "N" else np.asfortranarray(x.T)


def scaled_err(got, ref, absab):
    return float(np.max(np.abs(got - ref) / np.where(absab > 0, absab, 1.0)))


# ---------------------------------------------------------------------------
def test_device_properties(gpu):
    info = jb.device_info()
    print(info)
    assert info["compute_units"] > 0 and info["wavefront_size"] == 64
    assert "gfx950" in info["arch"]
    assert abs(info["peak_fp64_tflops"] - info["compute_units"] * info["clock_khz"] * 1e-9 * 128) < 1e-6


@pytest.mark.parametrize("m,n,k", [(16, 16, 4), (64, 64, 16), (128, 128, 16), (130, 70, 33), (257, 300, 129)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_mfma_layout_integer_kat(gpu, m, n, k, dtype):
    # small integers: every summation order is exact, so any lane-layout or index error shows
    rng = np.random.default_rng(m * 7 + n + k)
    A = np.asfortranarray(rng.integers(-8, 9, (m, k)).astype(dtype))
    B = np.asfortranarray(rng.integers(-8, 9, (k, n)).astype(dtype))
    ref = (A.astype(np.int64) @ B.astype(np.int64)).astype(dtype)
    C = np.zeros((m, n), dtype=dtype, order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C)
    assert np.array_equal(C, ref)
    for ta in "NT":
        for tb in "NT":
            Cd = to_dev(gpu, np.zeros((m, n), dtype=dtype))
            jb.gemm_(ta, tb, 1.0, to_dev(gpu, opnd(A, ta)), to_dev(gpu, opnd(B, tb)), 0.0, Cd)
            gpu.cuda.synchronize()
            assert np.array_equal(to_host(Cd), ref), (ta, tb)


def test_jmul_c1_bitexact(gpu):
    """Config: jmul! C = A*B, Float64, M=N=K=256, uniform[-1,1) seeds 42/43, C zero."""
    A = G.matrix(42, 256, 256)
    B = G.matrix(43, 256, 256)
    ref = O.jmul(A, B)
    D = np.zeros((256, 256), order="F")
    jb.jmul_(D, A, B)
    assert np.array_equal(D, ref)
    D2 = np.zeros((256, 256), order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, D2)  # FP64 MFMA path: same fma chain
    assert np.array_equal(D2, ref)
    Dd = to_dev(gpu, np.zeros((256, 256)))
    jb.jmul_(Dd, to_dev(gpu, A), to_dev(gpu, B))
    gpu.cuda.synchronize()
    assert np.array_equal(to_host(Dd), ref)
    # within the stated tolerance of the Dot2 oracle as well
    d2 = O.gemm_dot2("N", "N", 1.0, A, B, 0.0, np.zeros((256, 256)))
    assert scaled_err(D, d2, np.abs(A) @ np.abs(B)) <= 2 * 256 * U


@pytest.mark.parametrize("M,N,P", [(16, 32, 14), (32, 32, 28), (128, 128, 126), (32, 32, 6), (32, 32, 24),
                                   (13, 12, 12), (24, 18, 18), (800, 900, 840), (1, 1, 1), (3, 1000, 5)])
def test_jmul_reference_shapes(gpu, M, N, P):
    """The shapes test/runtests.jl exercises (plus fringe shapes the reference skips)."""
    A = G.matrix(1, M, N)
    X = G.matrix(2, N, P)
    D = np.full((M, P), np.nan, order="F")
    jb.jmul_(D, A, X)
    assert np.array_equal(D, O.jmul(A, X))


def test_jmul_f32_c1_and_reference_shapes(gpu):
    """jmul! in Float32 (the reference's jmul! is generic in T): the binary32
    recurrence, bit-exact against the oracle's Float32 restatement at 256^3
    (host and device pointers) and at the reference's shapes plus fringes."""
    A = G.matrix(42, 256, 256, dtype=np.float32)
    B = G.matrix(43, 256, 256, dtype=np.float32)
    ref = O.jmul(A, B)
    assert ref.dtype == np.float32
    D = np.zeros((256, 256), dtype=np.float32, order="F")
    jb.jmul_(D, A, B)
    assert np.array_equal(D.view(np.int32), ref.view(np.int32))
    Dd = to_dev(gpu, np.zeros((256, 256), dtype=np.float32))
    jb.jmul_(Dd, to_dev(gpu, A), to_dev(gpu, B))
    gpu.cuda.synchronize()
    assert np.array_equal(to_host(Dd).view(np.int32), ref.view(np.int32))
    for (M, N, P) in [(16, 32, 14), (32, 32, 28), (128, 128, 126), (13, 12, 12), (800, 900, 840), (1, 1, 1),
                      (3, 1000, 5), (65, 17, 130)]:
        A = G.matrix(1, M, N, dtype=np.float32)
        X = G.matrix(2, N, P, dtype=np.float32)
        D = np.full((M, P), np.nan, dtype=np.float32, order="F")
        jb.jmul_(D, A, X)
        assert np.array_equal(D.view(np.int32), O.jmul(A, X).view(np.int32)), (M, N, P)
    # signed zeros: the first step is a plain product, so -1 * 0 stays -0
    D = np.ones((1, 1), dtype=np.float32, order="F")
    jb.jmul_(D, np.array([[-1.0]], dtype=np.float32), np.array([[0.0]], dtype=np.float32))
    assert np.signbit(D[0, 0]) and D[0, 0] == 0


def test_jmul_sign_of_zero(gpu):
    D = np.ones((1, 1), order="F")
    jb.jmul_(D, np.array([[-1.0]]), np.array([[0.0]]))
    assert D[0, 0] == 0.0 and math.copysign(1.0, D[0, 0]) < 0  # acc_1 = fl(a*x), src/gemm.jl:86
    # a chain of -0 products stays -0 (fma(-1, 0, -0) = -0); +0 then -0 gives +0
    for A, sign in [(np.array([[-1.0, -2.0]]), -1.0), (np.array([[1.0, -1.0]]), 1.0), (np.array([[-1.0, 3.0]]), 1.0)]:
        X = np.zeros((2, 1), order="F")
        for f in (lambda D: jb.jmul_(D, A, X), lambda D: jb.gemm_("N", "N", 1.0, A, X, 0.0, D)):
            D = np.full((1, 1), np.nan, order="F")
            f(D)
            assert D[0, 0] == 0.0 and math.copysign(1.0, D[0, 0]) == sign, (A, sign)


def test_jmul_bits_equal_recurrence_kernel_with_signed_zeros(gpu):
    """jmul! (FP64 MFMA chains started at -0.0) against the sequential-fma VALU
    kernel on inputs full of +-0 and small integers: every bit equal, the sign
    of every zero included (array_equal would not see it)."""
    rng = np.random.default_rng(77)
    for M, N, P in [(100, 37, 70), (256, 256, 256), (65, 1, 33)]:
        A = np.asfortranarray(rng.choice([-2.0, -1.0, -0.0, 0.0, 1.0, 3.0], (M, N), p=[.1, .1, .3, .3, .1, .1]))
        X = np.asfortranarray(rng.choice([-1.0, -0.0, 0.0, 2.0], (N, P), p=[.2, .3, .3, .2]))
        D = np.full((M, P), np.nan, order="F")
        jb.jmul_(D, A, X)
        R = np.full((M, P), np.nan, order="F")
        jb.gemm_("N", "N", 1.0, A, X, 0.0, R, method="recurrence")
        assert np.array_equal(D.view(np.int64), R.view(np.int64)), (M, N, P)
        assert np.array_equal(D.view(np.int64), O.jmul(A, X).view(np.int64)), (M, N, P)


@pytest.mark.parametrize("M,N,P", [(1, 1, 1), (7, 1, 9), (13, 12, 12), (40, 256, 5), (64, 64, 64),
                                   (65, 33, 17), (200, 5, 300), (3, 1000, 5), (256, 256, 256)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fastmul_bitexact(gpu, M, N, P, dtype):
    """fastmul! (src/kernels.jl:202-208): one launch, the jmul! recurrence per
    element -- bit-identical to the oracle, host and device entry points.
    N == 1 and M not a multiple of 64 / P not a multiple of 16 are included."""
    A = G.matrix(31, M, N, dtype=dtype)
    X = G.matrix(32, N, P, dtype=dtype)
    ref = O.jmul(A, X)
    D = np.full((M, P), np.nan, dtype=dtype, order="F")
    jb.fastmul_(D, A, X)
    assert np.array_equal(D, ref)
    Dd = to_dev(gpu, np.full((M, P), np.nan, dtype=dtype))
    jb.fastmul_(Dd, to_dev(gpu, A), to_dev(gpu, X))
    gpu.cuda.synchronize()
    assert np.array_equal(to_host(Dd), ref)


def test_fastmul_sign_of_zero_and_neighbours(gpu):
    # acc_1 = fl(a*x) keeps -0; rows past M are never stored (no write beyond D).
    D = np.ones((1, 1), order="F")
    jb.fastmul_(D, np.array([[-1.0]]), np.array([[0.0]]))
    assert D[0, 0] == 0.0 and math.copysign(1.0, D[0, 0]) < 0
    M, N, P = 5, 3, 2
    buf = gpu.full((M * P + 64,), 7.0, dtype=gpu.float64, device="cuda")
    Dd = buf[: M * P].view(P, M).T  # column-major M x P at the start of a larger buffer
    A, X = G.matrix(33, M, N), G.matrix(34, N, P)
    jb.fastmul_(Dd, to_dev(gpu, A), to_dev(gpu, X))
    gpu.cuda.synchronize()
    assert np.array_equal(to_host(Dd), O.jmul(A, X))
    assert np.all(buf[M * P:].cpu().numpy() == 7.0)


@pytest.mark.parametrize("tile", ["128", "64", "64x32", "32"])
def test_native_tile_variants_bitexact(gpu, tile, monkeypatch):
    """Each MFMA tile variant (JB_GEMM_TILE pins it; the size rule picks it in
    normal use) is bit-identical to the recurrence, with fringes in m, n and k,
    every op and alpha/beta."""
    monkeypatch.setenv("JB_GEMM_TILE", tile)
    rng = np.random.default_rng(len(tile))
    for m, n, k in [(300, 200, 100), (64, 32, 16), (129, 65, 47)]:
        A0 = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
        B0 = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
        C0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        for ta in "NT":
            for tb in "NT":
                ref = O.gemm_recurrence("N", "N", 1.5, A0, B0, 0.5, C0)  # op(opnd(X, t)) == X
                Cd = to_dev(gpu, C0)
                jb.gemm_(ta, tb, 1.5, to_dev(gpu, opnd(A0, ta)), to_dev(gpu, opnd(B0, tb)), 0.5, Cd)
                gpu.cuda.synchronize()
                assert np.array_equal(to_host(Cd), ref), (tile, m, n, k, ta, tb)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_dgemm_c2_1024(gpu, ta, tb):
    """Config: C = 1.5*op(A)*op(B) + 0.5*C, M=N=K=1024, uniform[-1,1) seeds 42/43/44."""
    n = 1024
    A, B, C0 = G.matrix(42, n, n), G.matrix(43, n, n), G.matrix(44, n, n)
    Ad, Bd = to_dev(gpu, A), to_dev(gpu, B)
    ref = O.gemm_recurrence(ta, tb, 1.5, A, B, 0.5, C0)
    for method in ("native", "recurrence"):
        Cd = to_dev(gpu, C0)
        jb.gemm_(ta, tb, 1.5, Ad, Bd, 0.5, Cd, method=method)
        gpu.cuda.synchronize()
        got = to_host(Cd)
        assert np.array_equal(got, ref), method
    ii = np.random.default_rng(1).integers(0, n, 4096)
    jj = np.random.default_rng(2).integers(0, n, 4096)
    d2 = O.gemm_dot2_sample(ta, tb, 1.5, A, B, 0.5, C0[ii, jj], ii, jj)
    absab = 1.5 * (np.abs(opnd(A, ta)) @ np.abs(opnd(B, tb))) + 0.5 * np.abs(C0)
    assert scaled_err(got[ii, jj], d2, absab[ii, jj]) <= 2 * n * U


@pytest.mark.parametrize("n,ta,tb", [(n, ta, tb) for n in (4096, 8192) for ta in "NT" for tb in "NT"])
def test_dgemm_c2_large_sampled(gpu, n, ta, tb):
    """Config sizes 4096/8192: inputs generated in HBM (bit-identical to the host
    generator), sampled entries bit-checked against the recurrence and within
    tolerance of Dot2."""
    torch = gpu
    Ad = torch.empty((n, n), dtype=torch.float64, device="cuda")
    Bd = torch.empty((n, n), dtype=torch.float64, device="cuda")
    Cd = torch.empty((n, n), dtype=torch.float64, device="cuda")
    jb.fill_uniform_(Ad, 42)
    jb.fill_uniform_(Bd, 43)
    jb.fill_uniform_(Cd, 44)
    A, B, C0 = to_host(Ad.T), to_host(Bd.T), to_host(Cd.T)  # column-major views of the streams
    assert np.array_equal(A[:, 0], G.uniform(42, n))
    jb.gemm_(ta, tb, 1.5, Ad.T, Bd.T, 0.5, Cd.T)
    torch.cuda.synchronize()
    got = to_host(Cd.T)
    rng = np.random.default_rng(n)
    ii = np.concatenate([rng.integers(0, n, 2048), np.full(64, n - 1), np.arange(64)])
    jj = np.concatenate([rng.integers(0, n, 2048), np.arange(64), np.full(64, n - 1)])
    rec = O.gemm_recurrence_sample(ta, tb, 1.5, A, B, 0.5, C0[ii, jj], ii, jj)
    assert np.array_equal(got[ii, jj], rec)
    d2 = O.gemm_dot2_sample(ta, tb, 1.5, A, B, 0.5, C0[ii, jj], ii, jj)
    Ao, Bo = opnd(A, ta), opnd(B, tb)
    absab = 1.5 * np.einsum("sk,ks->s", np.abs(Ao[ii, :]), np.abs(Bo[:, jj])) + 0.5 * np.abs(C0[ii, jj])
    assert scaled_err(got[ii, jj], d2, absab) <= 2 * n * U
    # size-independent property: linearity in B, C(A, 2B) == 2*C(A, B) exactly (power of two)
    C2 = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    C3 = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    jb.gemm_(ta, tb, 1.0, Ad.T, Bd.T, 0.0, C2.T)
    jb.gemm_(ta, tb, 1.0, Ad.T, (2 * Bd).T, 0.0, C3.T)
    torch.cuda.synchronize()
    assert torch.equal(C3, 2 * C2)
    del Ad, Bd, Cd, C2, C3
    torch.cuda.empty_cache()


def test_kernel_initkernel_bits_with_signed_zeros(gpu):
    """kernel! (chains from D) and initkernel! (plain first product) on the
    FP64 MFMA kernel against the oracle, bitwise (int64 views), on inputs full
    of +-0 -- strided views, k not a multiple of the 16-k stage, and an N = 1
    chain whose only product is -0 on a D of -0 (stays -0) or +0 (stays +0)."""
    rng = np.random.default_rng(5)
    big = np.asfortranarray(rng.choice([-2.0, -1.0, -0.0, 0.0, 1.0], (300, 300), p=[.15, .15, .3, .3, .1]))
    Dbig = np.asfortranarray(rng.choice([-0.0, 0.0, 1.0], (301, 90), p=[.4, .4, .2]))
    for (ma, ka, kb) in [(100, 37, 70), (64, 16, 32), (33, 1, 5)]:
        A = big[10:10 + ma, 5:5 + ka]
        X = big[20:20 + ka, 100:100 + kb]
        D0 = Dbig[1:1 + ma, 3:3 + kb]
        Dw = np.asfortranarray(Dbig.copy())
        jb.kernel_(Dw[1:1 + ma, 3:3 + kb], A, X)
        ref = O.kernel_acc(D0, A, X)
        assert np.array_equal(Dw[1:1 + ma, 3:3 + kb].view(np.int64), ref.view(np.int64)), (ma, ka, kb)
        Di = np.asfortranarray(np.full((301, 90), np.nan))
        jb.initkernel_(Di[1:1 + ma, 3:3 + kb], A, X)
        assert np.array_equal(Di[1:1 + ma, 3:3 + kb].view(np.int64), O.jmul(A, X).view(np.int64)), (ma, ka, kb)
    for d0, sign in [(-0.0, -1.0), (0.0, 1.0)]:
        D = np.full((1, 1), d0, order="F")
        jb.kernel_(D, np.array([[-1.0]]), np.array([[0.0]]))  # d0 + (-0)
        assert D[0, 0] == 0.0 and math.copysign(1.0, D[0, 0]) == sign


def test_strided_subviews(gpu):
    """kernel!/initkernel! semantics: sub-blocks with leading dimensions != rows
    (src/kernels.jl:212-275), host and device."""
    big = G.matrix(7, 300, 300)
    A = big[10:110, 5:85]       # 100 x 80, lda 300
    B = big[20:100, 100:170]    # 80 x 70
    Cbig = G.matrix(8, 301, 90)
    C = Cbig[1:101, 3:73]       # 100 x 70, ldc 301
    ref_k = O.kernel_acc(C, A, B)  # kernel!: chains start from D
    Cw = np.asfortranarray(Cbig.copy())
    jb.kernel_(Cw[1:101, 3:73], A, B)  # D += A*X
    assert np.array_equal(Cw[1:101, 3:73], ref_k)
    assert np.array_equal(Cw[0, :], Cbig[0, :]) and np.array_equal(Cw[:, 80:], Cbig[:, 80:])
    Cd0 = to_dev(gpu, Cbig)
    jb.kernel_(Cd0[1:101, 3:73], to_dev(gpu, big)[10:110, 5:85], to_dev(gpu, big)[20:100, 100:170])
    gpu.cuda.synchronize()
    assert np.array_equal(to_host(Cd0)[1:101, 3:73], ref_k)
    Ci = np.asfortranarray(np.full((301, 90), np.nan))
    jb.initkernel_(Ci[1:101, 3:73], A, B)  # D = A*X, the jmul! recurrence
    assert np.array_equal(Ci[1:101, 3:73], O.jmul(A, B))
    assert np.all(np.isnan(Ci[0, :])) and np.all(np.isnan(Ci[:, 80:]))
    ref = O.gemm_recurrence("N", "N", 1.0, A, B, 1.0, C)
    bigd = to_dev(gpu, big)
    Cd = to_dev(gpu, Cbig)
    jb.gemm_("N", "N", 1.0, bigd[10:110, 5:85], bigd[20:100, 100:170], 1.0, Cd[1:101, 3:73])
    gpu.cuda.synchronize()
    out = to_host(Cd)
    assert np.array_equal(out[1:101, 3:73], ref)
    assert np.array_equal(out[101:, :], Cbig[101:, :])


def test_beta_zero_never_reads_c_alpha_zero_k_zero(gpu):
    A, B = G.matrix(3, 70, 50), G.matrix(4, 50, 60)
    C = np.full((70, 60), np.nan, order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C)
    assert not np.isnan(C).any()
    C0 = G.matrix(5, 70, 60)
    C = C0.copy(order="F")
    Anan = np.full((70, 50), np.nan, order="F")
    jb.gemm_("N", "N", 0.0, Anan, B, 0.5, C)  # alpha == 0: A not referenced
    assert np.array_equal(C, 0.5 * C0)
    C = C0.copy(order="F")
    jb.gemm_("N", "N", 2.0, np.zeros((70, 0), order="F"), np.zeros((0, 60), order="F"), -1.0, C)
    assert np.array_equal(C, -C0)


@pytest.mark.parametrize("name,dtype,extra", [("jb_dgemm", np.float64, ()), ("jb_dgemm_recurrence", np.float64, ()),
                                               ("jb_dgemm_ozaki", np.float64, (6,)),
                                               ("jb_dgemm_ozaki_i8", np.float64, (7,)),
                                               ("jb_sgemm", np.float32, ()), ("jb_sgemm_bf16x3", np.float32, (6,))])
def test_host_entries_alpha_zero_null_operands(gpu, name, dtype, extra):
    """alpha == 0 with m, n, k > 0: C = beta*C without reading A or B, so NULL
    operands are accepted and nothing is staged from them."""
    import ctypes

    m, n, k = 37, 29, 41
    C0 = G.matrix(6, m, n, dtype=dtype)
    C = C0.copy(order="F")
    L = jb.lib()
    rc = getattr(L, name)(b"N", b"N", m, n, k, 0.0, None, m, None, k, 0.5, C.ctypes.data_as(ctypes.c_void_p), m,
                          *extra)
    assert rc == 0, jb.last_error()
    assert np.array_equal(C, (dtype(0.5) * C0).astype(dtype))


def test_fill_bit_identical(gpu):
    torch = gpu
    t = torch.empty(100003, dtype=torch.float64, device="cuda")
    jb.fill_uniform_(t, 123)
    assert np.array_equal(t.cpu().numpy(), G.uniform(123, 100003))
    jb.fill_wide_(t, 9, -12, 12)
    assert np.array_equal(t.cpu().numpy(), G.wide(9, 100003))
    f = torch.empty(1001, dtype=torch.float32, device="cuda")
    jb.fill_uniform_(f, 5)
    assert np.array_equal(f.cpu().numpy(), G.uniform(5, 1001).astype(np.float32))
    jb.fill_wide_(f, 6, -12, 12)
    assert np.array_equal(f.cpu().numpy(), G.wide(6, 1001).astype(np.float32))


def test_fill_normal(gpu):
    """Device Box-Muller normals (mrandn's randn()) agree with the host
    generator to rounding (device log and cos) and have N(0, 1) moments."""
    torch = gpu
    n = 1 << 20
    t = torch.empty(n, dtype=torch.float64, device="cuda")
    jb.fill_normal_(t, 77)
    d, h = t.cpu().numpy(), G.normal(77, n)
    assert np.max(np.abs(d - h) / np.maximum(1.0, np.abs(h))) <= 1e-14
    assert abs(d.mean()) < 5 / np.sqrt(n) and abs(d.var() - 1.0) < 5 * np.sqrt(2.0 / n)
    f = torch.empty(4097, dtype=torch.float32, device="cuda")
    jb.fill_normal_(f, 78, 1.0 / np.sqrt(2048))
    hf = (G.normal(78, 4097) / np.sqrt(2048)).astype(np.float32)
    # f32: the same double rounded to float, or a neighbouring float where the
    # device's double differs across a rounding boundary
    assert np.max(np.abs(f.cpu().numpy().astype(np.float64) - hf) / np.maximum(2.0 ** -11, np.abs(hf))) <= 2.0 ** -23


def test_sgemm_native_tolerance(gpu):
    n = 512
    A = G.matrix(42, n, n, dtype=np.float32)
    B = G.matrix(43, n, n, dtype=np.float32)
    C0 = G.matrix(44, n, n, dtype=np.float32)
    ref = O.sgemm_dot2("N", "T", 1.5, A, opnd(B, "T"), 0.5, C0)
    C = C0.copy(order="F")
    jb.gemm_("N", "T", 1.5, A, opnd(B, "T"), 0.5, C)
    absab = 1.5 * (np.abs(A).astype(float) @ np.abs(B).astype(float)) + 0.5 * np.abs(C0)
    assert scaled_err(C.astype(float), ref, absab) <= 2 * n * 2.0 ** -24


def test_alloc_failure_is_reported(gpu):
    n = 2048  # larger host staging than any earlier host-pointer call in this session
    A, B = G.matrix(1, n, n), G.matrix(2, n, n)
    C = np.zeros((n, n), order="F")
    os.environ["JB_FAULT_ALLOC"] = "1"
    try:
        with pytest.raises(jb.JBError) as ei:
            jb.gemm_("N", "N", 1.0, A, B, 0.0, C)
        assert ei.value.code == 2
    finally:
        del os.environ["JB_FAULT_ALLOC"]
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C)
    assert np.isfinite(C).all()


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_native_nonfinite_matches_recurrence(gpu, ta, tb):
    """Inf, -Inf and NaN operand entries and products that overflow: every
    element of the FP64 MFMA result equals the reference recurrence's bit for
    bit where that is a number or an infinity, and is NaN where it is NaN
    (IEEE fma on both sides; NaN payloads are not compared), on full and
    fringe tiles of the 128-, 64- and 32-row variants."""
    m, n, k = 150, 90, 120
    A = G.matrix(71, m, k)
    B = G.matrix(72, k, n)
    C0 = G.matrix(73, m, n)
    A[3, 7], A[40, 0], A[149, 119] = np.inf, -np.inf, np.nan
    B[5, 11], B[119, 89] = np.nan, -np.inf
    A[60, :] *= 1e300  # row 60 of the products overflows against B's larger entries
    B[:, 20] *= 1e10
    ref = O.gemm_recurrence(ta, tb, 1.5, opnd(A, ta), opnd(B, tb), 0.5, C0)
    for tile in ("128", "64", "32"):
        os.environ["JB_GEMM_TILE"] = tile
        try:
            C = to_dev(gpu, C0)
            jb.gemm_(ta, tb, 1.5, to_dev(gpu, opnd(A, ta)), to_dev(gpu, opnd(B, tb)), 0.5, C, method="native")
            gpu.cuda.synchronize()
            got = to_host(C)
        finally:
            del os.environ["JB_GEMM_TILE"]
        nan_ref, nan_got = np.isnan(ref), np.isnan(got)
        assert np.array_equal(nan_ref, nan_got), (tile, ta, tb, np.argwhere(nan_ref != nan_got)[:5])
        assert np.array_equal(got[~nan_ref].view(np.int64), ref[~nan_ref].view(np.int64)), (tile, ta, tb)
        assert np.isinf(ref).any() and nan_ref.any()
