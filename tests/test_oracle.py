"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake (a dropped term, a wrong
index, a transposed operand, fma-vs-plain first step) fails it; none of them
retypes the oracle's own formula.
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from jblas_amd import generators as G

ROOT_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bf16_round(x):
    """float32 array rounded to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)

U = 2.0 ** -53


# ---------------------------------------------------------------------------
# SplitMix64 generator
# ---------------------------------------------------------------------------
def test_splitmix64_published_vector():
    # SplitMix64 (Steele, Lea & Flood 2014; Vigna's reference C) started at state 0.
    assert O.splitmix64(0, 0) == 0xE220A8397B1DCDAF
    assert O.splitmix64(0, 1) == 0x6E789E6AA1B965F4
    assert O.splitmix64(0, 2) == 0x06C45D188009454F


def test_generator_numpy_matches_c():
    for seed in (0, 42, 2 ** 63 + 5):
        n = 10007
        assert np.array_equal(G.uniform(seed, n), O.fill_uniform(seed, n))
        assert np.array_equal(G.wide(seed, n), O.fill_wide(seed, n))
        assert np.array_equal(G.wide(seed, n, -3, 9), O.fill_wide(seed, n, -3, 9))


def test_generator_ranges():
    u = G.uniform(7, 100000)
    assert u.min() >= -1.0 and u.max() < 1.0 and abs(u.mean()) < 0.01
    w = G.wide(7, 100000)
    e = np.frexp(w[w != 0])[1]
    assert e.min() >= -12 - 53 and e.max() <= 12
    sh = G.loguniform_shapes(52, 1024)
    assert sh.min() >= 16 and sh.max() <= 2048 and np.all(sh % 8 == 0)


# ---------------------------------------------------------------------------
# jmul! restatement
# ---------------------------------------------------------------------------
SHAPES_REF = [(16, 32, 14), (32, 32, 28), (128, 128, 126), (32, 32, 6), (32, 32, 24),  # test/runtests.jl:103-179
              (13, 12, 12), (24, 18, 18), (41, 7, 11), (1, 1, 1), (40, 3, 5), (3, 5, 7)]


@pytest.mark.parametrize("M,N,P", SHAPES_REF)
@pytest.mark.parametrize("avx512", [True, False])
def test_jmul_tiled_equals_naive(M, N, P, avx512):
    A = G.matrix(1, M, N)
    X = G.matrix(2, N, P)
    assert np.array_equal(O.jmul(A, X, avx512=avx512), O.jmul_naive(A, X))


def test_jmul_first_step_is_plain_product():
    # acc_1 = fl(a*x) (src/gemm.jl:86): (-1)*0 = -0, while fma(-1, 0, +0) = +0.
    D = O.jmul(np.array([[-1.0]]), np.array([[0.0]]))
    assert D[0, 0] == 0.0 and math.copysign(1.0, D[0, 0]) < 0


def test_jmul_fma_not_split():
    # a*x + acc with a*x inexact: fma keeps the low part that mul-then-add drops.
    a = 1.0 + 2.0 ** -30
    A = np.array([[1.0, a]])
    X = np.array([[-(1.0 + 2.0 ** -29)], [a]])
    D = O.jmul(A, X)
    exact = Fraction(1) * Fraction(X[0, 0]) + Fraction(a) * Fraction(a)
    assert D[0, 0] == float(exact) == 2.0 ** -60


@pytest.mark.parametrize("M,N,P", [(45, 33, 17), (128, 64, 40)])
def test_jmul_integer_kat(M, N, P):
    rng = np.random.default_rng(3)
    A = rng.integers(-1000, 1000, (M, N))
    X = rng.integers(-1000, 1000, (N, P))
    D = O.jmul(A.astype(float), X.astype(float))
    assert np.array_equal(D, (A @ X).astype(float))


def test_jmul_identity_permutation_ones_rank1_pow2():
    A = G.matrix(5, 37, 29)
    assert np.array_equal(O.jmul(A, np.eye(29)), A)
    perm = np.random.default_rng(0).permutation(29)
    Pm = np.eye(29)[:, perm]
    assert np.array_equal(O.jmul(A, Pm), A[:, perm])
    assert np.array_equal(O.jmul(np.ones((9, 23)), np.ones((23, 7))), np.full((9, 7), 23.0))
    a, x = G.matrix(6, 31, 1), G.matrix(7, 1, 13)
    assert np.array_equal(O.jmul(a, x), a * x)
    Dg = np.diag(2.0 ** np.arange(-5, 24))
    assert np.array_equal(O.jmul(A, Dg), A * 2.0 ** np.arange(-5, 24)[None, :])


def test_jmul_float32_restatement():
    A = G.matrix(8, 50, 20, dtype=np.float32)
    X = G.matrix(9, 20, 11, dtype=np.float32)
    D = O.jmul(A, X)
    assert D.dtype == np.float32
    ref = A.astype(np.float64) @ X.astype(np.float64)
    assert np.max(np.abs(D - ref)) <= 20 * 2.0 ** -24 * np.max(np.abs(A) @ np.abs(X))


def _round_f32(x: Fraction) -> float:
    """Correct round-to-nearest-even of an exact rational to binary32."""
    if x == 0:
        return 0.0
    sign, x = (-1.0 if x < 0 else 1.0), abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    q = Fraction(2) ** (max(e, -126) - 23)
    n = x / q
    fl = n.numerator // n.denominator
    r = n - fl
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return sign * float(fl * q)


def test_jmul_float32_bruteforce():
    # Per-element chain evaluated with exact rationals and one binary32 rounding
    # per step: acc = rd(a0*x0), acc = rd(a_k*x_k + acc) (fused, src/gemm.jl:86,165).
    A = G.matrix(21, 6, 9, dtype=np.float32)
    X = G.matrix(22, 9, 5, dtype=np.float32)
    D = O.jmul(A, X)
    for i in range(6):
        for j in range(5):
            acc = _round_f32(Fraction(float(A[i, 0])) * Fraction(float(X[0, j])))
            for k in range(1, 9):
                acc = _round_f32(Fraction(float(A[i, k])) * Fraction(float(X[k, j])) + Fraction(acc))
            assert D[i, j] == np.float32(acc), (i, j)
    # an unfused multiply-add differs from the chain on this input
    a = np.float32(1 + 2.0 ** -12)
    D1 = O.jmul(np.array([[1.0, a]], dtype=np.float32), np.array([[-(1 + 2.0 ** -11)], [a]], dtype=np.float32))
    assert D1[0, 0] == np.float32(2.0 ** -24)


def test_kernel_acc_chain_starts_from_d():
    # kernel! (src/kernels.jl:225-233): acc = D; acc = fma(a_n, x_n, acc) for every n.
    # Brute force with exact rationals and one binary64 rounding per step:
    rng = np.random.default_rng(30)
    A = rng.uniform(-1, 1, (5, 7))
    X = rng.uniform(-1, 1, (7, 4))
    D0 = rng.uniform(-1, 1, (5, 4))
    D = O.kernel_acc(D0, A, X)
    for i in range(5):
        for j in range(4):
            acc = float(D0[i, j])
            for n in range(7):
                acc = float(Fraction(float(A[i, n])) * Fraction(float(X[n, j])) + Fraction(acc))
            assert D[i, j] == acc
    # unlike gemm!(beta = 1), whose chain starts with a plain product and adds D last
    A1, X1, D1 = np.array([[2.0 ** -53, 2.0 ** -53]]), np.array([[1.0], [1.0]]), np.ones((1, 1))
    assert O.kernel_acc(D1, A1, X1)[0, 0] == 1.0
    assert O.gemm_recurrence("N", "N", 1.0, A1, X1, 1.0, D1)[0, 0] == 1.0 + 2.0 ** -52
    # strided operands: only the addressed block changes; N = 0 leaves D unchanged
    big = G.matrix(31, 40, 40)
    out = O.kernel_acc(big, big[:, :0], np.zeros((0, 40)))
    assert np.array_equal(out, big)
    Ii = rng.integers(-50, 50, (6, 9)).astype(float)
    Xi = rng.integers(-50, 50, (9, 3)).astype(float)
    Di = rng.integers(-50, 50, (6, 3)).astype(float)
    assert np.array_equal(O.kernel_acc(Di, Ii, Xi), Di + Ii @ Xi)


def test_gemm_recurrence_matches_jmul_and_ops():
    A = G.matrix(10, 33, 21)
    B = G.matrix(11, 21, 17)
    C0 = G.matrix(12, 33, 17)
    D = O.gemm_recurrence("N", "N", 1.0, A, B, 0.0, np.full((33, 17), np.nan))
    assert np.array_equal(D, O.jmul(A, B))  # beta = 0 never reads C (NaN does not leak)
    for ta in "NT":
        for tb in "NT":
            Aop = A if ta == "N" else np.asfortranarray(A.T)
            Bop = B if tb == "N" else np.asfortranarray(B.T)
            got = O.gemm_recurrence(ta, tb, 1.5, Aop, Bop, 0.5, C0)
            ref = 1.5 * (A @ B) + 0.5 * C0
            assert np.max(np.abs(got - ref)) < 1e-13
            # the op must not change the per-element chain
            base = O.gemm_recurrence("N", "N", 1.5, A, B, 0.5, C0)
            assert np.array_equal(got, base)
    ii = np.array([0, 5, 32, 7])
    jj = np.array([0, 16, 3, 9])
    s = O.gemm_recurrence_sample("N", "N", 1.5, A, B, 0.5, C0[ii, jj], ii, jj)
    assert np.array_equal(s, O.gemm_recurrence("N", "N", 1.5, A, B, 0.5, C0)[ii, jj])


@pytest.mark.parametrize("kc", [1, 32, 64, 96, 200, 1000])
def test_gemm_recurrence_chunked_semantics(kc):
    # chunk q is the plain recurrence on the k-slice; chunks are summed in order
    A = G.matrix(13, 37, 200)
    B = G.matrix(14, 200, 23)
    C0 = G.matrix(15, 37, 23)
    got = O.gemm_recurrence_chunked("T", "N", 1.5, np.asfortranarray(A.T), B, 0.5, C0, kc)
    s = None
    for c0 in range(0, 200, kc):
        sl = slice(c0, min(200, c0 + kc))
        part = O.gemm_recurrence("N", "N", 1.0, np.asfortranarray(A[:, sl]), np.asfortranarray(B[sl, :]), 0.0,
                                 np.zeros((37, 23)))
        s = part if s is None else s + part
    ref = 1.5 * s  # c = alpha*s (one rounding)
    # then c = fma(beta, c0, c): one rounding of the exact beta*c0 + c
    fin = np.empty_like(ref)
    for i in range(37):
        for j in range(23):
            fin[i, j] = float(Fraction(0.5) * Fraction(C0[i, j]) + Fraction(ref[i, j]))
    assert np.array_equal(got, fin)
    if kc >= 200:
        assert np.array_equal(got, O.gemm_recurrence("N", "N", 1.5, A, B, 0.5, C0))


# ---------------------------------------------------------------------------
# Dot2 FP64 oracle
# ---------------------------------------------------------------------------
def _exact_gemm(ta, tb, alpha, A, B, beta, C):
    Aop = A if ta == "N" else A.T
    Bop = B if tb == "N" else B.T
    m, k = Aop.shape
    n = Bop.shape[1]
    out = np.empty((m, n))
    for i in range(m):
        for j in range(n):
            s = sum(Fraction(float(Aop[i, p])) * Fraction(float(Bop[p, j])) for p in range(k))
            v = Fraction(alpha) * s + (Fraction(beta) * Fraction(float(C[i, j])) if beta != 0 else 0)
            out[i, j] = float(v)  # correctly rounded
    return out


def test_dot2_known_cancellation():
    A = np.array([[1e16, 1.0, -1e16, 3.0]])
    B = np.array([[1.0], [1.0], [1.0], [1.0]])
    assert A @ B != 4.0  # plain summation loses it
    assert O.gemm_dot2("N", "N", 1.0, A, B, 0.0, np.zeros((1, 1)))[0, 0] == 4.0
    # x*y with x*y inexact summed against -fl(x*y): result is the exact rounding error
    x, y = 1.0 + 2.0 ** -27, 1.0 - 2.0 ** -28
    A = np.array([[x, -1.0]])
    B = np.array([[y], [x * y]])
    exact = Fraction(x) * Fraction(y) - Fraction(x * y)
    assert O.gemm_dot2("N", "N", 1.0, A, B, 0.0, np.zeros((1, 1)))[0, 0] == float(exact)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_dot2_vs_fraction(ta, tb):
    rng = np.random.default_rng(11)
    m, n, k = 5, 4, 9
    A = rng.standard_normal((m, k) if ta == "N" else (k, m)) * 2.0 ** rng.integers(-20, 20, (m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C = rng.standard_normal((m, n))
    got = O.gemm_dot2(ta, tb, 1.5, A, B, -0.75, C)
    ref = _exact_gemm(ta, tb, 1.5, A, B, -0.75, C)
    absab = 1.5 * (np.abs(A if ta == "N" else A.T) @ np.abs(B if tb == "N" else B.T)) + 0.75 * np.abs(C)
    assert np.all(np.abs(got - ref) <= U * np.abs(ref) + 4 * k * k * U * U * absab)
    assert np.mean(got == ref) > 0.9  # nearly always the correctly rounded value


def test_dot2_sample_and_sgemm():
    A = G.matrix(20, 40, 30)
    B = G.matrix(21, 30, 25)
    C0 = G.matrix(22, 40, 25)
    full = O.gemm_dot2("N", "T", 1.5, A, np.asfortranarray(B.T), 0.5, C0)
    ii, jj = np.array([0, 39, 17]), np.array([0, 24, 5])
    s = O.gemm_dot2_sample("N", "T", 1.5, A, np.asfortranarray(B.T), 0.5, C0[ii, jj], ii, jj)
    assert np.array_equal(s, full[ii, jj])
    Af, Bf, Cf = A.astype(np.float32), B.astype(np.float32), C0.astype(np.float32)
    got = O.sgemm_dot2("N", "N", 1.5, Af, Bf, 0.5, Cf)
    ref = O.gemm_dot2("N", "N", 1.5, Af.astype(float), Bf.astype(float), 0.5, Cf.astype(float))
    assert np.array_equal(got, ref)


def test_dot2_beats_plain_sum_at_scale():
    A = G.matrix(30, 64, 2048)
    B = G.matrix(31, 2048, 64)
    d2 = O.gemm_dot2("N", "N", 1.0, A, B, 0.0, np.zeros((64, 64)))
    # uniform entries are multiples of 2^-53: exact dot products in Python integers
    Ai = (A * 2.0 ** 53).astype(np.int64)
    Bi = (B * 2.0 ** 53).astype(np.int64)
    npd = A @ B
    worse = 0
    for i, j in [(0, 0), (3, 17), (63, 63), (10, 40), (31, 2), (50, 9), (7, 7), (22, 60)]:
        exact = Fraction(sum(int(a) * int(b) for a, b in zip(Ai[i, :], Bi[:, j])), 2 ** 106)
        assert d2[i, j] == float(exact)  # Dot2 returns the correctly rounded value here
        worse += npd[i, j] != float(exact)
    assert worse > 0  # plain summation does not


# ---------------------------------------------------------------------------
# Split-BF16 emulated DGEMM restatement
# ---------------------------------------------------------------------------
def test_ozaki_digits_reconstruct_and_bound():
    A = G.matrix(40, 17, 50, kind="wide")
    for s in (1, 3, 5, 7):
        e, d = O.ozaki_split_a("N", A, s)
        assert np.abs(d).max() <= 256
        # every digit is exact in BF16 (at most 8 significant bits)
        assert np.array_equal(bf16_round(d.astype(np.float32)), d.astype(np.float32))
        w = 2.0 ** (-8 - 9 * np.arange(s))
        rec = np.einsum("t,tik->ik", w, d.astype(np.float64)) * 2.0 ** e[:, None]
        err = np.abs(rec - A)
        assert np.all(err <= 2.0 ** (e[:, None] - 9 * s) * 1.0000001)
    # a value with 9 significant bits at the row scale is captured by one digit
    row = np.array([[0.75, -0.5, 255 / 256, 2.0 ** -3 * 0.5078125]])
    e, d = O.ozaki_split_a("N", row, 1)
    assert e[0] == 0 and list(d[0, 0]) == [192, -128, 255, 16]
    # the top digit reaches 256 (balanced rounding), the next one takes the residual
    e, d = O.ozaki_split_a("N", np.array([[1 - 2.0 ** -10, 0.5]]), 2)
    assert e[0] == 0 and list(d[:, 0, 0]) == [256, -128] and list(d[:, 0, 1]) == [128, 0]


def test_ozaki_exact_when_representable():
    rng = np.random.default_rng(5)
    m, n, k = 12, 9, 40
    # entries with <= 9 significant bits relative to the row/column scale
    A = rng.integers(-255, 256, (m, k)) * 2.0 ** -8
    B = rng.integers(-255, 256, (k, n)) * 2.0 ** -10
    A[:, 0], B[0, :] = 255 / 256, 255 / 256 / 4  # pin the row/column exponents (0 and -2)
    for s in (1, 2, 4, 7):
        got = O.dgemm_ozaki("N", "N", 1.0, A, B, 0.0, np.zeros((m, n)), s)
        assert np.array_equal(got, A @ B)  # integer-scaled inputs: A@B is exact in float64


def ozaki_worst_value(s):
    """The value whose s 9-bit digits are +256, -256, +256, ... (exact in float64 for s <= 6)."""
    return sum((256.0 if t % 2 == 0 else -256.0) * 2.0 ** (-8 - 9 * t) for t in range(s))


def test_ozaki_worst_case_digits():
    # every digit at the +-256 bound; every pair of level L has the sign (-1)^L, so the
    # level sums of a chunk reach (L+1)*kc*2^16 (2^24 for L = 0, 1, 3): still exact
    s, k = 6, 1024
    x = ozaki_worst_value(s)
    e, d = O.ozaki_split_a("N", np.full((3, 5), x), s)
    assert list(e) == [0, 0, 0] and [int(v) for v in d[:, 0, 0]] == [256, -256, 256, -256, 256, -256]
    got = O.dgemm_ozaki("N", "N", 1.0, np.full((2, k), x), np.full((k, 3), x), 0.0, np.zeros((2, 3)), s)
    # exact: k * sum over levels L < s of (L+1) pairs of 2^16 (-1)^L at weight 2^(-16-9L)
    exact = Fraction(k) * sum(Fraction((L + 1) * (-1) ** L, 2 ** (9 * L)) for L in range(s))
    assert np.all(np.abs(got - float(exact)) <= np.spacing(float(exact)))


def test_ozaki_level_chunks_are_the_longest_exact_ones():
    # level L has L+1 slice pairs of products |d_t d_u| <= 2^16: its FP32 sum over
    # kc k-values is exact while (L+1) * kc * 2^16 <= 2^24; kc is the largest power of two
    for L in range(12):
        kc = O.ozaki_level_chunk(L)
        assert kc & (kc - 1) == 0
        assert (L + 1) * kc * 2 ** 16 <= 2 ** 24 < (L + 1) * 2 * kc * 2 ** 16
    assert [O.ozaki_level_chunk(L) for L in range(6)] == [256, 128, 64, 64, 32, 32]
    # the worst case (every digit 256) over one chunk is exactly representable in float32
    L = 3
    total = (L + 1) * O.ozaki_level_chunk(L) * 256 * 256
    assert float(np.float32(total)) == total == 2 ** 24


def test_ozaki_error_shrinks_per_slice():
    A = G.matrix(42, 48, 256, kind="wide")
    B = G.matrix(43, 256, 40, kind="wide")
    ref = O.gemm_dot2("N", "N", 1.0, A, B, 0.0, np.zeros((48, 40)))
    scale = np.abs(A) @ np.abs(B)
    errs = []
    for s in range(2, 9):
        got = O.dgemm_ozaki("N", "N", 1.0, A, B, 0.0, np.zeros((48, 40)), s)
        errs.append(np.max(np.abs(got - ref) / scale))
    # each slice adds ~9 bits: the componentwise error drops by at least 2^-7 per
    # slice until the FP64 rounding of the result (~2^-53 here) takes over
    for a, b in zip(errs, errs[1:]):
        if a > 2.0 ** -50:
            assert b < a * 2.0 ** -7
    # truncation bound ~2^(-9s+c) relative to (|A||B|)_ij with c <= 10 for this input
    for s, e in zip(range(2, 9), errs):
        assert e < max(2.0 ** (-9 * s + 10), 2.0 ** -50)


def test_ozaki_ops_alpha_beta_consistent():
    A = G.matrix(44, 20, 33)
    B = G.matrix(45, 33, 14)
    C0 = G.matrix(46, 20, 14)
    base = O.dgemm_ozaki("N", "N", 1.5, A, B, 0.5, C0, 5)
    for ta in "NT":
        for tb in "NT":
            Aop = A if ta == "N" else np.asfortranarray(A.T)
            Bop = B if tb == "N" else np.asfortranarray(B.T)
            assert np.array_equal(O.dgemm_ozaki(ta, tb, 1.5, Aop, Bop, 0.5, C0, 5), base)
    with pytest.raises(ValueError):
        O.dgemm_ozaki("N", "N", 1.0, np.array([[np.inf]]), np.array([[1.0]]), 0.0, np.zeros((1, 1)), 3)


def _i8_rounded(a, e, s):
    """rint(a * 2^(8s-1-e)) as an exact Python integer (round half to even)."""
    q = Fraction(float(a)) * Fraction(2) ** (8 * s - 1 - int(e))
    fl = q.numerator // q.denominator
    r = q - fl
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return fl


def test_ozaki_i8_digits_are_signed_bytes_of_the_rounded_value():
    # rows whose max sits just below a power of two (the top digit would be 127.x),
    # exactly at one, tiny entries, mixed signs
    rows = np.array([[0.99999, -0.3, 1e-9, 0.5],
                     [-0.9883, 0.25, -2.0 ** -40, 0.1],
                     [0.98828125, -0.98828125, 0.0, 3e-5],
                     [1.0, -1.0, 0.75, -0.5],
                     [-7.9, 3.3, 1e-300, 6.0]])
    A = np.vstack([rows, G.matrix(41, 12, 4, kind="wide")])
    for s in (1, 2, 4, 7, 8):
        e, d = O.ozaki_i8_split_a("N", A, s)
        assert d.min() >= -128 and d.max() <= 127
        for i in range(A.shape[0]):
            # the frexp exponent, bumped by one when max*2^-E > 126.5/128
            f, ex = math.frexp(float(np.max(np.abs(A[i]))))
            assert e[i] == (ex + 1 if f > 126.5 / 128 else ex)
            for p in range(A.shape[1]):
                X = sum(int(d[t, i, p]) * 256 ** (s - 1 - t) for t in range(s))
                assert X == _i8_rounded(A[i, p], e[i], s), (s, i, p)
    # op('T') splits the columns of A^T the same way
    e1, d1 = O.ozaki_i8_split_a("N", A, 5)
    e2, d2 = O.ozaki_i8_split_a("T", np.asfortranarray(A.T), 5)
    assert np.array_equal(e1, e2) and np.array_equal(d1, d2)


def test_ozaki_i8_exact_when_representable():
    rng = np.random.default_rng(6)
    m, n, k = 10, 11, 70
    A = rng.integers(-100, 101, (m, k)) * 2.0 ** -7
    B = rng.integers(-100, 101, (k, n)) * 2.0 ** -10
    for s in (1, 2, 3):
        got = O.dgemm_ozaki_i8("N", "N", 1.0, A, B, 0.0, np.zeros((m, n)), s)
        assert np.array_equal(got, A @ B)


def test_ozaki_i8_error_shrinks_per_slice_and_matches_bf16_split():
    A = G.matrix(42, 48, 256, kind="wide")
    B = G.matrix(43, 256, 40, kind="wide")
    ref = O.gemm_dot2("N", "N", 1.0, A, B, 0.0, np.zeros((48, 40)))
    scale = np.abs(A) @ np.abs(B)
    errs, errs_bf = [], []
    for s in range(2, 9):
        got = O.dgemm_ozaki_i8("N", "N", 1.0, A, B, 0.0, np.zeros((48, 40)), s)
        errs.append(np.max(np.abs(got - ref) / scale))
        bf = O.dgemm_ozaki("N", "N", 1.0, A, B, 0.0, np.zeros((48, 40)), s)
        errs_bf.append(np.max(np.abs(bf - ref) / scale))
    for a, b in zip(errs, errs[1:]):
        assert b < a * 2.0 ** -5
    # an INT8 digit carries 8 bits, a BF16 digit 9: at s slices the INT8 split is
    # less accurate by up to ~2^s, never more accurate by more than rounding
    for s, e, eb in zip(range(2, 9), errs, errs_bf):
        assert e < max(2.0 ** (-8 * s + 11), 2.0 ** -50)
        assert eb / 2 <= e <= 2.0 ** (s + 2) * max(eb, 2.0 ** -52), (s, e, eb)


def test_ozaki_i8_chunked_level_sums():
    # k above one int32 chunk (8192): chunks flush into the FP64 groups; with
    # one-digit inputs every partial sum is an integer below 2^53: A@B exactly
    assert O.ozaki_i8_chunk() == 8192
    rng = np.random.default_rng(8)
    m, n, k = 3, 2, 20000
    A = rng.integers(-126, 127, (m, k)) * 2.0 ** -7
    B = rng.integers(-126, 127, (k, n)) * 2.0 ** -7
    A[:, 0] = 126 / 128  # row/column max below 126.5/128: exponent 0, one exact digit
    B[0, :] = 126 / 128
    got = O.dgemm_ozaki_i8("N", "N", 1.0, A, B, 0.0, np.zeros((m, n)), 1)
    ref = (A * 2 ** 7).astype(np.int64) @ (B * 2 ** 7).astype(np.int64)
    assert np.array_equal(got, ref * 2.0 ** -14)


def test_ozaki_i8_ops_alpha_beta_consistent():
    A = G.matrix(44, 20, 33)
    B = G.matrix(45, 33, 14)
    C0 = G.matrix(46, 20, 14)
    base = O.dgemm_ozaki_i8("N", "N", 1.5, A, B, 0.5, C0, 5)
    for ta in "NT":
        for tb in "NT":
            Aop = A if ta == "N" else np.asfortranarray(A.T)
            Bop = B if tb == "N" else np.asfortranarray(B.T)
            assert np.array_equal(O.dgemm_ozaki_i8(ta, tb, 1.5, Aop, Bop, 0.5, C0, 5), base)
    with pytest.raises(ValueError):
        O.dgemm_ozaki_i8("N", "N", 1.0, np.array([[np.nan]]), np.array([[1.0]]), 0.0, np.zeros((1, 1)), 3)


# ---------------------------------------------------------------------------
# Tile policy (src/kernel_structure.jl:76-99, src/memory_management.jl:78-177)
# ---------------------------------------------------------------------------
def test_pick_kernel_size_table():
    # (REGISTER_SIZE, REGISTER_COUNT, sizeof(T)) -> (elements/register, rows, cols), derived by hand
    assert O.pick_kernel_size(64, 32, 8) == (8, 40, 5)
    assert O.pick_kernel_size(64, 32, 4) == (16, 80, 5)
    assert O.pick_kernel_size(32, 16, 8) == (4, 12, 4)
    assert O.pick_kernel_size(32, 16, 4) == (8, 24, 4)


def test_blocking_structure_table():
    cs = (48 * 1024, 2 * 1024 * 1024, 105 * 1024 * 1024)
    assert O.blocking_structure(256, 256, 256, 8, cs, 64, 32) == (((40, 132, 5), (256, 256, 256), (256, 256, 256)), 1)
    assert O.blocking_structure(1024, 1024, 1024, 8, cs, 64, 32) == (((40, 132, 5), (400, 132, 395), (1024, 1024, 1024)), 2)
    assert O.blocking_structure(8, 8, 8, 8, cs, 64, 32) == (((8, 8, 5), (8, 8, 8), (8, 8, 8)), 0)


def test_oracle_under_address_and_ub_sanitizers():
    """Every entry point of the checker runs clean under AddressSanitizer and
    UndefinedBehaviorSanitizer (oracle/asan_check.c: exact-size buffers, fringe
    shapes, both transposes, strided leading dimensions, all slice counts)."""
    import subprocess

    here = os.path.join(ROOT_DIR, "oracle")
    subprocess.run(["make", "-s", "-C", here, "asan"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(here, "_build", "oracle_asan")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ran clean" in r.stdout
