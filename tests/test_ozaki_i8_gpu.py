"""GPU parity of the split-INT8 emulated DGEMM (method "ozaki_int8").

Every digit is an int8 and every level sum over a k-chunk of 8192 is an exact
int32 (|S_l| <= 8 * 8192 * 2^14 = 2^30), whatever the MFMA's internal order;
every FP64 step is the oracle's.  So the GPU result must equal the oracle's
restatement (oracle.dgemm_ozaki_i8) BIT-FOR-BIT, for every slice count, op,
alpha/beta and across k-chunks.  At 8192^3 the check runs on a 64 x 64 block
of sampled rows and columns; accuracy against the Dot2 FP64 oracle is compared
with native FP64 and with the BF16-slice method on the same entries.
"""
import numpy as np
import pytest

import jblas_amd as jb
from jblas_amd import generators as G
from oracle import oracle as O

from conftest import to_dev, to_host

pytestmark = pytest.mark.gpu


def opnd(x, t):
    return x if t == "N" else np.asfortranarray(x.T)


@pytest.mark.parametrize("s", [1, 2, 3, 4, 5, 6, 7, 8])
def test_ozaki_i8_bitexact_small(gpu, s):
    m, n, k = 150, 70, 700  # fringe tiles, k not a multiple of the 32-k stage
    A = G.matrix(42, m, k, kind="wide")
    B = G.matrix(43, k, n, kind="wide")
    C0 = G.matrix(44, m, n)
    ref = O.dgemm_ozaki_i8("N", "N", 1.5, A, B, 0.5, C0, s)
    C = C0.copy(order="F")
    jb.gemm_("N", "N", 1.5, A, B, 0.5, C, method="ozaki_int8", nslices=s)
    assert np.array_equal(C, ref)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_ozaki_i8_bitexact_ops_device(gpu, ta, tb):
    m, n, k = 260, 190, 333
    A = G.matrix(1, m, k, kind="wide")
    B = G.matrix(2, k, n, kind="wide")
    C0 = G.matrix(3, m, n)
    for s, alpha, beta in [(7, 1.5, 0.5), (4, -2.0, 0.0), (6, 1.0, 1.0)]:
        ref = O.dgemm_ozaki_i8("N", "N", alpha, A, B, beta, C0, s)
        Cd = to_dev(gpu, C0)
        jb.gemm_(ta, tb, alpha, to_dev(gpu, opnd(A, ta)), to_dev(gpu, opnd(B, tb)), beta, Cd,
                 method="ozaki_int8", nslices=s)
        gpu.cuda.synchronize()
        assert np.array_equal(to_host(Cd), ref), (s, alpha, beta)


def test_ozaki_i8_bitexact_across_kchunks(gpu):
    # k = 17000: three int32 chunks (8192, 8192, 616) flushed into the FP64 groups
    m, n, k = 40, 36, 17000
    A = G.matrix(5, m, k, kind="wide")
    B = G.matrix(6, k, n, kind="wide")
    ref = O.dgemm_ozaki_i8("N", "N", 1.0, A, B, 0.0, np.zeros((m, n)), 7)
    C = np.zeros((m, n), order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method="ozaki_int8", nslices=7)
    assert np.array_equal(C, ref)


def test_ozaki_i8_extreme_digits(gpu):
    # every lower digit -128 (products of 2^14 at every level, level sums up to
    # 4 * 8192 * 2^14 = 2^29 in one chunk) and top digits +-125: built from the
    # digits, exactly representable (6 digits = 47 bits)
    m, n, k = 64, 64, 8192
    rng = np.random.default_rng(12)
    low = sum(-128.0 * 2.0 ** (-7 - 8 * t) for t in range(1, 6))
    A = np.asfortranarray(rng.choice([-1.0, 1.0], (m, k)) * 125 / 128 + low)
    B = np.asfortranarray(rng.choice([-1.0, 1.0], (k, n)) * 125 / 128 + low)
    for s in (6, 8):
        e, d = O.ozaki_i8_split_a("N", A, s)
        assert set(np.unique(d[1:6])) == {-128} and e.max() == 0
        ref = O.dgemm_ozaki_i8("N", "N", 1.0, A, B, 0.0, np.zeros((m, n)), s)
        C = np.zeros((m, n), order="F")
        jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method="ozaki_int8", nslices=s)
        assert np.array_equal(C, ref), s


def test_ozaki_i8_exact_for_representable_inputs(gpu):
    rng = np.random.default_rng(9)
    m, n, k = 96, 80, 1500
    A = np.asfortranarray(rng.integers(-126, 127, (m, k)) * 2.0 ** -7)
    B = np.asfortranarray(rng.integers(-126, 127, (k, n)) * 2.0 ** -20)
    A[:, 0], B[0, :] = 126 / 128, 126 * 2.0 ** -20
    C = np.zeros((m, n), order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method="ozaki_int8", nslices=1)
    assert np.array_equal(C, A @ B)


def test_ozaki_i8_degenerate_sizes_and_alpha_zero(gpu):
    # k == 0 or alpha == 0: C = beta*C without reading A or B; m == 0 / n == 0: no-op
    C0 = G.matrix(7, 20, 10)
    C = C0.copy(order="F")
    jb.gemm_("N", "N", 1.0, np.zeros((20, 0), order="F"), np.zeros((0, 10), order="F"), 0.5, C,
             method="ozaki_int8", nslices=7)
    assert np.array_equal(C, 0.5 * C0)
    C = C0.copy(order="F")
    A = np.full((20, 30), np.nan, order="F")
    jb.gemm_("N", "N", 0.0, A, G.matrix(8, 30, 10), 2.0, C, method="ozaki_int8", nslices=7)
    assert np.array_equal(C, 2.0 * C0)
    jb.gemm_("N", "N", 1.0, np.zeros((0, 5), order="F"), np.zeros((5, 3), order="F"), 0.0,
             np.zeros((0, 3), order="F"), method="ozaki_int8", nslices=3)
    with pytest.raises(jb.JBError) as ei:
        jb.gemm_("N", "N", 1.0, G.matrix(1, 4, 4), G.matrix(2, 4, 4), 0.0, np.zeros((4, 4), order="F"),
                 method="ozaki_int8", nslices=9)
    assert ei.value.code == -14




def test_ozaki_i8_c3_8192_sampled(gpu):
    """Config: M=N=K=8192, a_ij = u*2^e (e uniform in [-12,12]), s = 5..8 INT8 slices.
    Bit-exact vs the restatement on a 64x64 sample block; error vs Dot2 compared
    with native FP64 and with the BF16-slice method (s = 6) on the same entries."""
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
    Cd = torch.zeros((n, n), dtype=torch.float64, device="cuda")

    def rel_err(method, s=7):
        jb.gemm_("N", "N", 1.0, Ad.T, Bd.T, 0.0, Cd.T, method=method, nslices=s)
        torch.cuda.synchronize()
        got = to_host(Cd.T)[np.ix_(ii, jj)]
        return got, float(np.max(np.abs(got - d2) / np.abs(d2)))

    _, nat = rel_err("native")
    _, bf6 = rel_err("ozaki", 6)
    print(f"\nnative max rel {nat:.3e}; BF16 s=6 {bf6:.3e}")
    prev = None
    for s in (5, 6, 7, 8):
        got, rel = rel_err("ozaki_int8", s)
        assert np.array_equal(got, O.dgemm_ozaki_i8("N", "N", 1.0, Asub, Bsub, 0.0, np.zeros((64, 64)), s)), s
        print(f"INT8 s={s}: max rel {rel:.3e}  ratio to native {rel / nat:.2f}")
        if prev is not None:
            assert rel < prev
        prev = rel
        if s == 7:
            assert rel <= 4 * nat  # within the emulated method's accuracy bar at 7 slices
    del Ad, Bd, Cd
    torch.cuda.empty_cache()
