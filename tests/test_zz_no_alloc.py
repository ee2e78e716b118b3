"""The allocation-free emulated DGEMM (jb_dgemm_ozaki_noalloc_dev) through the C
ABI: bit-identical to jb_dgemm_ozaki_dev on two calls in a row, the second
smaller (its scratch keeps the first call's contents), and a size beyond its
scratch rejected before any HIP call."""
import ctypes

import numpy as np
import pytest

import jblas_amd as jb
from jblas_amd import generators as G
from oracle import oracle as O

from conftest import to_dev, to_host

pytestmark = pytest.mark.gpu

VP = ctypes.c_void_p


def call(name, ta, tb, m, n, k, alpha, A, B, beta, C, s, stream):
    return getattr(jb.lib(), name)(ta, tb, m, n, k, alpha, VP(A.data_ptr()), A.stride(1), VP(B.data_ptr()),
                                   B.stride(1), beta, VP(C.data_ptr()), C.stride(1), s, VP(stream))


@pytest.mark.parametrize("s", [6, 3])
def test_noalloc_matches_dev_entry(gpu, s):
    torch = gpu
    stream = torch.cuda.current_stream().cuda_stream
    cases = [(b"N", b"N", 900, 700, 1100, 1.5, 0.5, 61), (b"T", b"N", 300, 200, 260, -2.0, 0.0, 62)]
    for ta, tb, m, n, k, alpha, beta, seed in cases:  # the second call is smaller
        A = G.matrix(seed, m if ta == b"N" else k, k if ta == b"N" else m, kind="wide")
        B = G.matrix(seed + 100, k if tb == b"N" else n, n if tb == b"N" else k, kind="wide")
        C0 = G.matrix(seed + 200, m, n)
        Ad, Bd = to_dev(torch, A), to_dev(torch, B)
        Cref, Cna = to_dev(torch, C0), to_dev(torch, C0)
        assert call("jb_dgemm_ozaki_dev", ta, tb, m, n, k, alpha, Ad, Bd, beta, Cref, s, stream) == 0, jb.last_error()
        assert call("jb_dgemm_ozaki_noalloc_dev", ta, tb, m, n, k, alpha, Ad, Bd, beta, Cna, s, stream) == 0, \
            jb.last_error()
        torch.cuda.synchronize()
        got, ref = to_host(Cna), to_host(Cref)
        assert np.array_equal(got.view(np.int64), ref.view(np.int64)), (ta, tb, m, n, k)
        if m <= 300:
            assert np.array_equal(got, O.dgemm_ozaki(ta.decode(), tb.decode(), alpha, A, B, beta, C0, s))


def test_noalloc_rejects_beyond_scratch(gpu):
    torch = gpu
    stream = torch.cuda.current_stream().cuda_stream
    A = torch.zeros((2048, 2048), dtype=torch.float64, device="cuda")
    rc = call("jb_dgemm_ozaki_noalloc_dev", b"N", b"N", 2048, 2048, 2048, 1.0, A.T, A.T, 0.0, A.T, 6, stream)
    assert rc == 6 and "scratch" in jb.last_error()  # JB_ERR_UNSUPPORTED
    rc = call("jb_dgemm_ozaki_noalloc_dev", b"N", b"N", 16, 16, 16, 1.0, A.T, A.T, 0.0, A.T, 9, stream)
    assert rc == -14
