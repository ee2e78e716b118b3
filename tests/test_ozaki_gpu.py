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


def clear_nonfinite():
    """Clears the condition that earlier calls (host-pointer ones included) raised."""
    try:
        jb.nonfinite_check()
    except jb.JBError as e:
        assert e.code == 3


def nonfinite_expected(restatement, A, B, s, rows=(), cols=()):
    """C = A*B from the restatement with the given rows of A / columns of B
    replaced by zeros, then those rows / columns NaN: an element of C depends
    only on its row of op(A) and column of op(B), split at their own scales."""
    A0, B0 = A.copy(order="F"), B.copy(order="F")
    A0[list(rows), :] = 0.0
    B0[:, list(cols)] = 0.0
    ref = restatement("N", "N", 1.0, A0, B0, 0.0, np.zeros((A.shape[0], B.shape[1])), s)
    ref[list(rows), :] = np.nan
    ref[:, list(cols)] = np.nan
    return ref


@pytest.mark.parametrize("method,restatement", [("ozaki", O.dgemm_ozaki), ("ozaki_int8", O.dgemm_ozaki_i8)])
def test_emulated_nonfinite_host_is_an_error(gpu, method, restatement):
    A = G.matrix(1, 40, 30)
    B = G.matrix(2, 30, 20)
    A[5, 7] = np.inf
    C = np.zeros((40, 20), order="F")
    with pytest.raises(jb.JBError) as ei:
        jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method=method, nslices=4)
    assert ei.value.code == 3
    # the rows fed by the Inf are NaN, every other element is the usual result
    assert np.array_equal(C, nonfinite_expected(restatement, A, B, 4, rows=[5]), equal_nan=True)
    B[3, 3] = np.nan
    A[5, 7] = 1.0
    C = np.zeros((40, 20), order="F")
    with pytest.raises(jb.JBError):
        jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method=method, nslices=4)
    assert np.array_equal(C, nonfinite_expected(restatement, A, B, 4, cols=[3]), equal_nan=True)
    # and a clean call afterwards works
    B[3, 3] = 0.5
    C = np.zeros((40, 20), order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method=method, nslices=4)
    assert np.array_equal(C, restatement("N", "N", 1.0, A, B, 0.0, np.zeros((40, 20)), 4))


@pytest.mark.parametrize("method,restatement", [("ozaki", O.dgemm_ozaki), ("ozaki_int8", O.dgemm_ozaki_i8)])
def test_emulated_dev_nonfinite_is_asynchronous(gpu, method, restatement):
    """The _dev form returns without synchronising; the NaN rows/columns are in C
    and jb_nonfinite_check reports the condition once."""
    torch = gpu
    clear_nonfinite()
    A = G.matrix(3, 70, 90, kind="wide")
    B = G.matrix(4, 90, 50, kind="wide")
    A[10, 2], A[69, 89], B[0, 49] = -np.inf, np.nan, np.inf
    Cd = to_dev(torch, np.zeros((70, 50)))
    jb.gemm_("N", "N", 1.0, to_dev(torch, A), to_dev(torch, B), 0.0, Cd, method=method, nslices=5)
    with pytest.raises(jb.JBError) as ei:
        jb.nonfinite_check()
    assert ei.value.code == 3
    jb.nonfinite_check()  # cleared
    assert np.array_equal(to_host(Cd), nonfinite_expected(restatement, A, B, 5, rows=[10, 69], cols=[49]),
                          equal_nan=True)
    A[10, 2], A[69, 89], B[0, 49] = 1.0, 1.0, 1.0
    Cd = to_dev(torch, np.zeros((70, 50)))
    jb.gemm_("N", "N", 1.0, to_dev(torch, A), to_dev(torch, B), 0.0, Cd, method=method, nslices=5)
    jb.nonfinite_check()
    assert np.array_equal(to_host(Cd), restatement("N", "N", 1.0, A, B, 0.0, np.zeros((70, 50)), 5))


@pytest.mark.parametrize("method", ["ozaki", "ozaki_int8", "bf16x3"])
def test_emulated_graph_capture_replays(gpu, method):
    """An emulated _dev call captured into a graph (after one eager call of the
    same size) replays bit-identically to the eager result, with eager work of
    the library between the replays."""
    torch = gpu
    f32 = method == "bf16x3"
    dt = np.float32 if f32 else np.float64
    kw = {"nprod": 6} if f32 else {"nslices": 6 if method == "ozaki" else 7}
    m, n, k = 300, 200, 520
    A = G.matrix(11, m, k, kind="uniform" if f32 else "wide", dtype=dt)
    B = G.matrix(12, k, n, kind="uniform" if f32 else "wide", dtype=dt)
    C0 = G.matrix(13, m, n, dtype=dt)
    Ad, Bd, Ce, Cg = to_dev(torch, A), to_dev(torch, B), to_dev(torch, C0), to_dev(torch, C0)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        jb.gemm_("N", "N", 1.5, Ad, Bd, 0.5, Ce, method=method, **kw)
    torch.cuda.synchronize()
    eager = to_host(Ce)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        jb.gemm_("N", "N", 1.5, Ad, Bd, 0.5, Cg, method=method, **kw)
    torch.cuda.synchronize()
    C0d = to_dev(torch, C0)
    for rep in range(3):
        Cg.copy_(C0d)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(to_host(Cg), eager), f"replay {rep} differs"
        # eager work of the library between the replays (same workspace)
        X = to_dev(torch, np.zeros((m, n), dtype=dt))
        jb.gemm_("N", "N", 1.0, Ad, Bd, 0.0, X, method=method, **kw)
        torch.cuda.synchronize()
    if not f32:
        ref = (O.dgemm_ozaki if method == "ozaki" else O.dgemm_ozaki_i8)("N", "N", 1.5, A, B, 0.5, C0, kw["nslices"])
        assert np.array_equal(eager, ref)
    del g


def test_sgemm_bf16x3_nonfinite_and_flt_max(gpu):
    """Entries up to FLT_MAX split exactly (the hi slice truncates where rounding
    to nearest would overflow BF16); an Inf or NaN operand makes every product
    it enters NaN and is reported."""
    torch = gpu
    m, n, k = 64, 48, 96
    A = G.matrix(21, m, k, dtype=np.float32)
    B = G.matrix(22, k, n, dtype=np.float32) * np.float32(2.0 ** -100)
    A[:, 0] = np.float32(3.4028235e38)   # FLT_MAX
    A[:, 1] = np.float32(-3.4e38)
    A[:, 2] = np.float32(3.3962e38)  # above the largest BF16 by more than half an ulp: nearest is Inf
    ref = A.astype(np.float64) @ B.astype(np.float64)
    absab = np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)
    Cd = to_dev(torch, np.zeros((m, n), dtype=np.float32))
    clear_nonfinite()
    jb.gemm_("N", "N", 1.0, to_dev(torch, A), to_dev(torch, B), 0.0, Cd, method="bf16x3", nprod=6)
    jb.nonfinite_check()  # nothing to report: every entry is finite
    got = to_host(Cd).astype(np.float64)
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - ref) / absab) <= 2 * k * 2.0 ** -24
    A[7, 5], B[9, 30] = np.inf, np.nan
    C = np.zeros((m, n), dtype=np.float32, order="F")
    with pytest.raises(jb.JBError) as ei:
        jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method="bf16x3", nprod=6)
    assert ei.value.code == 3
    bad = np.zeros((m, n), bool)
    bad[7, :] = True
    bad[:, 30] = True
    assert np.all(np.isnan(C[bad])) and np.all(np.isfinite(C[~bad]))


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


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_sgemm_bf16x3_wide_block_fringes(gpu, ta, tb):
    """A problem large enough for the 128x256 3xBF16 block on a 256-CU device
    (32 x 16 tiles) whose m and n are not multiples of the block (partial row
    blocks of the op(B) planes): the first and last 64 rows and columns within
    2*K*2^-24 componentwise of an FP64 product of the same FP32 operands."""
    m, n, k = 4000, 3900, 200
    A = G.matrix(45, m, k, dtype=np.float32)
    B = G.matrix(46, k, n, dtype=np.float32)
    C0 = G.matrix(47, m, n, dtype=np.float32)
    Cd = to_dev(gpu, C0)
    jb.gemm_(ta, tb, 1.5, to_dev(gpu, opnd(A, ta)), to_dev(gpu, opnd(B, tb)), 0.5, Cd, method="bf16x3", nprod=6)
    got = to_host(Cd).astype(float)
    rows = np.r_[0:64, m - 64:m]
    cols = np.r_[0:64, n - 64:n]
    A64, B64 = A[rows].astype(float), B[:, cols].astype(float)
    ref = 1.5 * (A64 @ B64) + 0.5 * C0[np.ix_(rows, cols)].astype(float)
    bound = 1.5 * (np.abs(A64) @ np.abs(B64)) + 0.5 * np.abs(C0[np.ix_(rows, cols)]).astype(float)
    err = np.max(np.abs(got[np.ix_(rows, cols)] - ref) / bound)
    assert err <= 2 * k * 2.0 ** -24, (ta, tb, err)


def test_sgemm_bf16x3_c3_8192(gpu):
    """Config: Float32 inputs via the 3xBF16 split at M=N=K=8192, a_ij = u*2^e
    (e uniform in [-12, 12]) generated in HBM (bit-identical to the host
    generator).  Checked on a 64x64 block of sampled rows x columns plus 16 whole
    rows and 16 whole columns against the FP64 Dot2 oracle (the float inputs are
    exact in FP64): componentwise |C - C*| <= 2*K*2^-24 (|A||B|) for 6 and 9
    products; max and mean relative error reported per product count."""
    torch = gpu
    n = 8192
    Ad = torch.empty((n, n), dtype=torch.float32, device="cuda")
    Bd = torch.empty((n, n), dtype=torch.float32, device="cuda")
    jb.fill_wide_(Ad, 42)
    jb.fill_wide_(Bd, 43)
    A, B = to_host(Ad.T).astype(np.float64), to_host(Bd.T).astype(np.float64)
    assert np.array_equal(A[:, 0], G.wide(42, n).astype(np.float32))
    rng = np.random.default_rng(8192)
    rows, cols = np.sort(rng.choice(n, 64, replace=False)), np.sort(rng.choice(n, 64, replace=False))
    wr, wc = rng.choice(n, 16, replace=False), rng.choice(n, 16, replace=False)
    ii = np.concatenate([np.repeat(rows, 64), np.repeat(wr, n), np.tile(np.arange(n), 16)])
    jj = np.concatenate([np.tile(cols, 64), np.tile(np.arange(n), 16), np.repeat(wc, n)])
    ref = O.gemm_dot2_sample("N", "N", 1.0, A, B, 0.0, np.zeros(ii.size), ii, jj)
    aA, aB = np.abs(A), np.abs(B)
    absab = np.concatenate([(aA[rows, :] @ aB[:, cols]).ravel(), (aA[wr, :] @ aB).ravel(), (aA @ aB[:, wc]).T.ravel()])
    del aA, aB
    big = np.abs(ref) > 1e-3 * np.max(np.abs(ref))
    stats = {}
    Cd = torch.empty((n, n), dtype=torch.float32, device="cuda")
    for nprod in (3, 6, 9):
        jb.gemm_("N", "N", 1.0, Ad.T, Bd.T, 0.0, Cd.T, method="bf16x3", nprod=nprod)
        torch.cuda.synchronize()
        got = Cd.T.cpu().numpy()[ii, jj].astype(np.float64)
        comp = np.max(np.abs(got - ref) / absab)
        rel = np.abs(got - ref)[big] / np.abs(ref[big])
        stats[nprod] = (comp, rel.max(), rel.mean())
    print({p: {"componentwise": v[0], "max_rel": v[1], "mean_rel": v[2]} for p, v in stats.items()},
          f"(relative errors over {int(big.sum())} of {ii.size} entries with |C*| > 1e-3 max|C*|)")
    assert stats[6][0] <= 2 * n * 2.0 ** -24 and stats[9][0] <= 2 * n * 2.0 ** -24
    assert stats[3][0] > stats[6][0] >= stats[9][0] * 0.5
    del Ad, Bd, Cd
    torch.cuda.empty_cache()


def test_sgemm_bf16x3_host_and_exact_small_ints(gpu):
    rng = np.random.default_rng(4)
    A = np.asfortranarray(rng.integers(-100, 100, (70, 90)).astype(np.float32))
    B = np.asfortranarray(rng.integers(-100, 100, (90, 50)).astype(np.float32))
    C = np.zeros((70, 50), dtype=np.float32, order="F")
    jb.gemm_("N", "N", 1.0, A, B, 0.0, C, method="bf16x3", nprod=6)
    assert np.array_equal(C, (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32))
