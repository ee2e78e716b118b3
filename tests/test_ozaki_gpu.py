"""GPU parity of the split-BF16 emulated DGEMM and the 3xBF16 SGEMM.

Emulated DGEMM: every FP32 level sum is an exact integer (kc*(l+1) <= 256) and
every FP64 step is the oracle's, so the GPU result must equal the oracle's
restatement (oracle.dgemm_ozaki) BIT-FOR-BIT, for every slice count, op and
alpha/beta.  At 8192^3 the check runs on a 64 x 64 block of sampled rows and
columns (an entry depends only on its row of op(A) and column of op(B)).
Accuracy against the Dot2 FP64 oracle is reported per slice count.

3xBF16 SGEMM: tolerance |C - C*|_ij <= 2*K*2^-24 (|alpha||A||B| + |beta||C0|)_ij
against the FP64 oracle (FP32-level accuracy, which is the point of the 3-slice split).
"""
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


@pytest.mark.parametrize("s", [1, 2, 3, 4, 5, 6, 7, 8])
def test_ozaki_bitexact_small(gpu, s):
    m, n, k = 150, 70, 700  # fringe tiles and several k-chunks
    A = G.matrix(42, m, k, kind="wide")
    B = G.matrix(43, k, n, kind="wide")
    C0 = G.matrix(44, m, n)
    ref = O.dgemm_ozaki("N", "N", 1.5, A, B, 0.5, C0, s)
    C = C0.copy(order="F")
    jb.gemm_("N", "N", 1.5, A, B, 0.5, C, method="ozaki", nslices=s)
    assert np.array_equal(C, ref)


def ozaki_worst_value(s):
    """The value whose s 9-bit digits are +256, -256, +256, ... (exact in float64 for s <= 6)."""
    return sum((256.0 if t % 2 == 0 else -256.0) * 2.0 ** (-8 - 9 * t) for t in range(s))


def test_ozaki_worst_case_digits_bitexact(gpu):
    """Every digit +-256 and every pair of level L of sign (-1)^L: the FP32 level sums
    of levels 0, 1 and 3 reach exactly 2^24 at their chunk ends (the exactness limit);
    rows of A and columns of B carry random signs."""
    s, m, n, k = 6, 130, 66, 1024
    x = ozaki_worst_value(s)
    rng = np.random.default_rng(17)
    A = np.asfortranarray(rng.choice([-x, x], (m, 1)) * np.ones((1, k)))
    B = np.asfortranarray(np.ones((k, 1)) * rng.choice([-x, x], (1, n)))
    ref = O.dgemm_ozaki("N", "N", 1.0, A, B, 0.0, np.zeros((m, n)), s)
    C = np.zeros((m, n), order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method="ozaki", nslices=s)
    assert np.array_equal(C, ref)


@pytest.mark.parametrize("s,k", [(6, 2 ** 18), (8, 60000)])
def test_ozaki_bitexact_long_k(gpu, s, k):
    """k*s beyond 2^18 with worst-case digits: the FP64 group sums round, so the flush
    order (per chunk end, increasing level; deferred two-stage levels at s = 6) must
    be the oracle's for the bits to agree."""
    m, n = 8, 8
    x = ozaki_worst_value(min(s, 6))
    rng = np.random.default_rng(18)
    A = np.asfortranarray(rng.choice([-x, x], (m, 1)) * np.ones((1, k)))
    B = np.asfortranarray(np.ones((k, 1)) * rng.choice([-x, x], (1, n)))
    ref = O.dgemm_ozaki("N", "N", 1.0, A, B, 0.0, np.zeros((m, n)), s)
    C = np.zeros((m, n), order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method="ozaki", nslices=s)
    assert np.array_equal(C, ref)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_ozaki_bitexact_ops_device(gpu, ta, tb):
    m, n, k = 260, 190, 333
    A = G.matrix(1, m, k, kind="wide")
    B = G.matrix(2, k, n, kind="wide")
    C0 = G.matrix(3, m, n)
    for s, alpha, beta in [(7, 1.5, 0.5), (4, -2.0, 0.0), (6, 1.0, 1.0)]:
        ref = O.dgemm_ozaki("N", "N", alpha, A, B, beta, C0, s)
        Cd = to_dev(gpu, C0)
        jb.gemm_(ta, tb, alpha, to_dev(gpu, opnd(A, ta)), to_dev(gpu, opnd(B, tb)), beta, Cd,
                 method="ozaki", nslices=s)
        gpu.cuda.synchronize()
        assert np.array_equal(to_host(Cd), ref), (s, alpha, beta)


def test_ozaki_exact_for_representable_inputs(gpu):
    rng = np.random.default_rng(9)
    m, n, k = 96, 80, 1500
    # 9 significant bits at the row/column scale: one BF16 digit (|d| <= 256) each
    A = np.asfortranarray(rng.integers(-255, 256, (m, k)) * 2.0 ** -8)
    B = np.asfortranarray(rng.integers(-255, 256, (k, n)) * 2.0 ** -20)
    A[:, 0], B[0, :] = 255 / 256, 255 * 2.0 ** -20
    C = np.zeros((m, n), order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method="ozaki", nslices=1)
    assert np.array_equal(C, A @ B)


def test_ozaki_nonfinite_is_an_error(gpu):
    A = G.matrix(1, 40, 30)
    B = G.matrix(2, 30, 20)
    A[5, 7] = np.inf
    with pytest.raises(jb.JBError) as ei:
        jb.gemm_("N", "N", 1.0, A, B, 0.0, np.zeros((40, 20), order="F"), method="ozaki", nslices=4)
    assert ei.value.code == 3
    B[3, 3] = np.nan
    A[5, 7] = 1.0
    with pytest.raises(jb.JBError):
        jb.gemm_("N", "N", 1.0, A, B, 0.0, np.zeros((40, 20), order="F"), method="ozaki", nslices=4)
    # and a clean call afterwards works
    B[3, 3] = 0.5
    C = np.zeros((40, 20), order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method="ozaki", nslices=4)
    assert np.array_equal(C, O.dgemm_ozaki("N", "N", 1.0, A, B, 0.0, np.zeros((40, 20)), 4))


def test_ozaki_c3_8192_sampled(gpu):
    """Config: M=N=K=8192, a_ij = u*2^e (e uniform in [-12,12]), s = 3..8 slices.
    Bit-exact vs the restatement on a 64x64 sample block; error vs Dot2 reported
    per slice count and compared with native FP64 MFMA DGEMM on the same entries."""
    torch = gpu
    n = 8192
    Ad = torch.empty((n, n), dtype=torch.float64, device="cuda")
    Bd = torch.empty((n, n), dtype=torch.float64, device="cuda")
    jb.fill_wide_(Ad, 42)
    jb.fill_wide_(Bd, 43)
    rng = np.random.default_rng(88)
    ii = np.sort(rng.choice(n, 64, replace=False))
    jj = np.sort(rng.choice(n, 64, replace=False))
    Afull, Bfull = to_host(Ad.T), to_host(Bd.T)
    Asub, Bsub = np.asfortranarray(Afull[ii, :]), np.asfortranarray(Bfull[:, jj])
    del Afull, Bfull
    d2 = O.gemm_dot2("N", "N", 1.0, Asub, Bsub, 0.0, np.zeros((64, 64)))
    absab = np.abs(Asub) @ np.abs(Bsub)
    Cd = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    jb.gemm_("N", "N", 1.0, Ad.T, Bd.T, 0.0, Cd.T, method="native")
    torch.cuda.synchronize()
    nat = to_host(Cd.T)[np.ix_(ii, jj)]
    nat_err = np.max(np.abs(nat - d2) / absab)
    nat_rel = np.max(np.abs(nat - d2) / np.abs(d2))
    print(f"\nnative: max err/(|A||B|) = {nat_err:.3e}  max rel = {nat_rel:.3e}")
    prev, errs = None, {}
    for s in (3, 4, 5, 6, 7, 8):
        jb.gemm_("N", "N", 1.0, Ad.T, Bd.T, 0.0, Cd.T, method="ozaki", nslices=s)
        torch.cuda.synchronize()
        got = to_host(Cd.T)[np.ix_(ii, jj)]
        ref = O.dgemm_ozaki("N", "N", 1.0, Asub, Bsub, 0.0, np.zeros((64, 64)), s)
        assert np.array_equal(got, ref), s
        err = np.max(np.abs(got - d2) / absab)
        rel = np.max(np.abs(got - d2) / np.abs(d2))
        print(f"s={s}: max err/(|A||B|) = {err:.3e}  max rel = {rel:.3e}  ratio to native {rel / nat_rel:.2f}")
        if prev is not None:
            assert err < prev
        prev = err
        errs[s] = (err, rel)
    # six 9-bit slices (21 products): componentwise at least as accurate as native
    # FP64 (0.45x on this sample), max relative error within 4x (2.0x)
    assert errs[6][0] < nat_err and errs[6][1] < 4 * nat_rel
    del Ad, Bd, Cd
    torch.cuda.empty_cache()


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_sgemm_bf16x3_tolerance(gpu, ta, tb):
    m, n, k = 300, 260, 1000
    A = G.matrix(42, m, k, dtype=np.float32)
    B = G.matrix(43, k, n, dtype=np.float32)
    C0 = G.matrix(44, m, n, dtype=np.float32)
    ref = O.sgemm_dot2("N", "N", 1.5, A, B, 0.5, C0)
    absab = 1.5 * (np.abs(A).astype(float) @ np.abs(B).astype(float)) + 0.5 * np.abs(C0)
    errs = {}
    for nprod in (3, 6, 9):
        Cd = to_dev(gpu, C0)
        jb.gemm_(ta, tb, 1.5, to_dev(gpu, opnd(A, ta)), to_dev(gpu, opnd(B, tb)), 0.5, Cd, method="bf16x3",
                 nprod=nprod)
        gpu.cuda.synchronize()
        errs[nprod] = np.max(np.abs(to_host(Cd).astype(float) - ref) / absab)
    print(errs)
    assert errs[6] <= 2 * k * 2.0 ** -24 and errs[9] <= 2 * k * 2.0 ** -24
    assert errs[3] > errs[6]


def test_sgemm_bf16x3_host_and_exact_small_ints(gpu):
    rng = np.random.default_rng(4)
    A = np.asfortranarray(rng.integers(-100, 100, (70, 90)).astype(np.float32))
    B = np.asfortranarray(rng.integers(-100, 100, (90, 50)).astype(np.float32))
    C = np.zeros((70, 50), dtype=np.float32, order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method="bf16x3", nprod=6)
    assert np.array_equal(C, (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32))
