"""The C-ABI library: it loads, exports every symbol include/jblas_hip.h
declares, and validates arguments with BLAS numbering before any HIP call
(these tests run without a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import jblas_amd
from jblas_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "jblas_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the ctypes signature table covers the whole header
    assert sorted(_lib.SIGNATURES) == syms


def test_julia_binding_names_declared_symbols():
    """The Julia wrapper (shipped as source; no Julia in this image) only
    ccalls symbols the header declares, with the header's argument count."""
    src = os.path.join(ROOT, "jblas.jl_amd", "julia", "JBlasAMD", "src", "JBlasAMD.jl")
    text = open(src).read()
    called = re.findall(r"ccall\(\(:(jb_[a-z0-9_]+), libjblas\), \w+,\s*\(([^)]*)\)", text)
    assert len(called) >= 10
    header = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for name, argtypes in called:
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", header)
        assert m, name
        n_header = 0 if m.group(1).strip() in ("", "void") else m.group(1).count(",") + 1
        argtypes = argtypes.strip().rstrip(",")  # Julia 1-tuples carry a trailing comma
        n_julia = 0 if not argtypes else argtypes.count(",") + 1
        assert n_header == n_julia, (name, n_header, n_julia)


def test_version_and_error_string():
    assert _lib.lib().jb_version() >= 10000
    assert isinstance(_lib.last_error(), str)


def _dgemm(ta=b"N", tb=b"N", m=4, n=4, k=4, alpha=1.0, A=1, lda=4, B=1, ldb=4, beta=0.0, Cp=1, ldc=4):
    return _lib.lib().jb_dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cp, ldc)


@pytest.mark.parametrize("kw,code", [
    (dict(ta=b"X"), -1), (dict(tb=b"q"), -2), (dict(m=-1), -3), (dict(n=-2), -4), (dict(k=-3), -5),
    (dict(A=None), -7), (dict(lda=3), -8), (dict(B=None), -9), (dict(ldb=2), -10), (dict(Cp=None), -12),
    (dict(ldc=3), -13), (dict(ta=b"T", lda=3), -8), (dict(tb=b"T", ldb=3, n=5, ldc=4), -10),
])
def test_dgemm_argument_validation(kw, code):
    assert _dgemm(**kw) == code
    assert _lib.last_error() != ""


def test_quick_returns_need_no_device():
    # m == 0 or n == 0: nothing to do, success without touching a device
    assert _dgemm(m=0, lda=1, ldc=1) == 0
    assert _dgemm(n=0) == 0
    # alpha == 0 relaxes the A/B null checks (A and B are not referenced)
    assert _lib.lib().jb_dgemm(b"N", b"N", 0, 3, 3, 0.0, None, 1, None, 3, 1.0, None, 1) == 0


def test_other_entry_points_validate():
    L = _lib.lib()
    assert L.jb_dgemm_ozaki(b"N", b"N", 4, 4, 4, 1.0, 1, 4, 1, 4, 0.0, 1, 4, 0) == -14
    assert L.jb_dgemm_ozaki(b"N", b"N", 4, 4, 4, 1.0, 1, 4, 1, 4, 0.0, 1, 4, 9) == -14
    assert L.jb_sgemm_bf16x3(b"N", b"N", 4, 4, 4, 1.0, 1, 4, 1, 4, 0.0, 1, 4, 5) == -14
    assert L.jb_sgemm(b"N", b"Z", 4, 4, 4, 1.0, 1, 4, 1, 4, 0.0, 1, 4) == -2
    assert L.jb_jmul(4, 0, 4, 1, 1, 1) == -2
    assert L.jb_jmul(-1, 2, 4, 1, 1, 1) == -1
    assert L.jb_kernel(4, 3, 4, 1, 3, 1, 4, 1, 3) == -5      # ldd < M
    assert L.jb_kernel(4, 3, 4, 1, 4, 1, 2, 1, 3) == -7      # lda < M
    assert L.jb_kernel(4, 3, 4, 1, 4, 1, 4, 1, 2) == -9      # ldx < N
    assert L.jb_kernel(4, 0, 4, 1, 4, None, 4, None, 1) == 0  # N == 0: D unchanged, nothing read
    assert L.jb_initkernel(4, 0, 4, 1, 4, 1, 4, 1, 1) == -2  # initkernel! needs a product
    assert L.jb_kernel(4, 3, 4, None, 4, 1, 4, 1, 3) == -4
    # an output grid beyond one launch (2^23 tiles of 64 x 64) is refused before any HIP call
    assert L.jb_dgemm(b"N", b"N", 1 << 20, 1 << 20, 4, 1.0, 1, 1 << 20, 1, 4, 0.0, 1, 1 << 20) == 6
    assert L.jb_kernel(1 << 20, 4, 1 << 20, 1, 1 << 20, 1, 1 << 20, 1, 4) == 6
    assert L.jb_fastmul(4, 0, 4, 1, 1, 1) == -2
    assert L.jb_fastmul_f32(4, 3, -4, 1, 1, 1) == -3
    assert L.jb_fastmul(4, 3, 4, None, 1, 1) == -4
    assert L.jb_fill_uniform_dev(1, -5, None, None) == -2
    assert L.jb_fill_wide_dev(1, 5, 3, 2, 1, None) == -4


def test_host_layer_shape_errors():
    A, B, Cm = np.zeros((4, 3), order="F"), np.zeros((5, 2), order="F"), np.zeros((4, 2), order="F")
    with pytest.raises(jblas_amd.DimensionMismatch):
        jblas_amd.gemm_("N", "N", 1.0, A, B, 0.0, Cm)
    with pytest.raises(jblas_amd.DimensionMismatch):
        jblas_amd.jmul_(np.zeros((4, 4), order="F"), A, np.zeros((3, 5), order="F"))
    with pytest.raises(ValueError):
        jblas_amd.gemm_("N", "N", 1.0, A, np.zeros((3, 2), order="F"), 0.0, np.zeros((4, 2), order="C"))
    with pytest.raises(ValueError):
        jblas_amd.gemm_("N", "N", 1.0, A, np.zeros((3, 2), order="F"), 0.0, np.zeros((4, 2), order="F"),
                        method="bogus")


def test_no_device_is_a_clean_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    rc = _lib.lib().jb_init(0)
    assert rc == _lib.JB_ERR_HIP and "device" in _lib.last_error()


def test_grouped_static_choice_rule():
    """The planner's static choice (no tuning table; no GPU needed): problems
    with m, n, k <= 64 run whole in one work item (fastmul!'s form,
    src/kernels.jl:202-208); larger outputs take the 64x64 MFMA tile from
    max(m, n) >= 96, else 32x32; a small output with a long k splits k."""
    from jblas_amd import grouped as G
    G.tune_reset()
    for m, n, k in [(1, 1, 1), (64, 64, 64), (16, 40, 8), (64, 1, 64), (5, 7, 0)]:
        assert G.tune_lookup("f64", m, n, k) == ("small", 1), (m, n, k)
        assert G.tune_lookup("f32_bf16x3", m, n, k) == ("small", 1), (m, n, k)
    assert G.tune_lookup("f64", 65, 64, 64) == ("32x32", 1)
    assert G.tune_lookup("f64", 64, 64, 65) == ("32x32", 1)
    assert G.tune_lookup("f64", 96, 32, 64) == ("64x64", 1)
    assert G.tune_lookup("f64", 64, 64, 4096) == ("32x32", 1)  # 4 tiles: no split
    assert G.tune_lookup("f64", 16, 16, 4096) == ("32x32", 4)  # 1 tile: split while k / (2 s) >= 256 and tiles * s < 4
