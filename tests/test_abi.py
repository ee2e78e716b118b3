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


JL = os.path.join(ROOT, "jblas.jl_amd", "julia", "JBlasAMD")


def _julia_sources():
    files = [os.path.join(JL, "src", "JBlasAMD.jl")]
    files += sorted(os.path.join(JL, "ext", f) for f in os.listdir(os.path.join(JL, "ext")) if f.endswith(".jl"))
    return files


def test_julia_binding_names_declared_symbols():
    """The Julia wrapper and its extensions (shipped as source; no Julia in this
    image) only ccall symbols the header declares, with the header's argument
    count."""
    text = "\n".join(open(f).read() for f in _julia_sources())
    called = re.findall(r"ccall\(\(:(jb_[a-z0-9_]+), libjblas\), \w+,\s*\(([^)]*)\)", text)
    assert len(called) >= 30
    header = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for name, argtypes in called:
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", header)
        assert m, name
        n_header = 0 if m.group(1).strip() in ("", "void") else m.group(1).count(",") + 1
        argtypes = argtypes.strip().rstrip(",")  # Julia 1-tuples carry a trailing comma
        n_julia = 0 if not argtypes else argtypes.count(",") + 1
        assert n_header == n_julia, (name, n_header, n_julia)


def test_julia_reference_signatures():
    """The reference's own call forms have methods: jmul!(D::MMatrix{M,P,T},
    A::MMatrix{M,N,T}, X::MMatrix{N,P,T}, ::Val...) (src/gemm.jl:244-246),
    fastmul!(::MMatrix...) (src/kernels.jl:202-208), kernel!/initkernel!(pD, pA,
    pX, ::Kernel{Mk,Pk,stride_AD,stride_X,N}[, pf]) (src/kernels.jl:239-241,
    273-275) with Kernel(Mk, N, Pk, stride_AD, stride_X) (src/kernel_structure.jl:
    8-9), methods of jBLAS's own functions, and device-array (ROCMatrix) methods."""
    src = open(os.path.join(JL, "src", "JBlasAMD.jl")).read()
    ext = {f: open(os.path.join(JL, "ext", f)).read() for f in os.listdir(os.path.join(JL, "ext"))}
    sa, jb_, amd = ext["JBlasAMDStaticArraysExt.jl"], ext["JBlasAMDjBLASExt.jl"], ext["JBlasAMDAMDGPUExt.jl"]
    mm = r"\(D::MMatrix\{M,P,T\}, A::MMatrix\{M,N,T\}, X::MMatrix\{N,P,T\}"
    assert re.search(r"function JBlasAMD\.jmul!" + mm + r",\s*::Val = Val\(7\), ::Val = Val\(7\), ::Val = Val\(3\)", sa)
    assert re.search(r"function JBlasAMD\.fastmul!" + mm, sa)
    assert re.search(r"function jBLAS\.jmul!" + mm, jb_) and re.search(r"function jBLAS\.fastmul!" + mm, jb_)
    for fn in ("kernel!", "initkernel!"):
        assert re.search(r"function " + re.escape(fn) + r"\(pD::Ptr\{Float64\}, pA::Ptr\{Float64\}, pX::Ptr\{Float64\}, "
                         r"::Kernel\{M\u2096,P\u2096,stride_AD,stride_X,N\}", src)
        assert re.search(r"jBLAS\." + re.escape(fn) + r"\(pD::Ptr\{Float64\}.*\n\s*::jBLAS\.Kernel\{", jb_)
        assert re.search(r"function JBlasAMD\." + re.escape(fn) + r"\(D::ROCMatrix\{Float64\}", amd)
    assert "struct Kernel{M\u2096,P\u2096,stride_AD,stride_X,N} end" in src
    assert "Kernel(M\u2096, N, P\u2096, stride_AD, stride_X) = Kernel{M\u2096,P\u2096,stride_AD,stride_X,N}()" in src
    assert re.search(r"function JBlasAMD\.jmul!\(D::ROCMatrix\{Float64\}", amd)
    assert re.search(r"function JBlasAMD\.gemm!\(.*A::ROCMatrix\{Float64\}", amd)
    proj = open(os.path.join(JL, "Project.toml")).read()
    for name in ("JBlasAMDStaticArraysExt", "JBlasAMDjBLASExt", "JBlasAMDAMDGPUExt", "JBlasAMDChainRulesCoreExt"):
        assert name in proj and name + ".jl" in ext


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


_LOG_PROBE = """
import ctypes, sys
sys.path.insert(0, {pkg!r})
from jblas_amd import _lib
L = _lib.lib()
L.jb_dgemm(b"N", b"N", -3, 4, 4, ctypes.c_double(1.0), None, 1, None, 1, ctypes.c_double(0.0), None, 1)
"""


@pytest.mark.parametrize("mode", ["file", "off"])
def test_jb_log(tmp_path, mode):
    """JB_LOG: the library prints nothing unless it is set; set to a path,
    every error an entry point returns is appended there as one line."""
    import subprocess
    import sys
    log = tmp_path / "jb.log"
    env = dict(os.environ)
    env.pop("JB_LOG", None)
    if mode == "file":
        env["JB_LOG"] = str(log)
    code = _LOG_PROBE.format(pkg=os.path.join(ROOT, "jblas.jl_amd"))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "jblas:" not in r.stderr
    if mode == "file":
        lines = log.read_text().splitlines()
        assert "jblas: error -3: m < 0" in lines, lines  # the third argument (m) is invalid
        assert all(ln.startswith("jblas: ") for ln in lines), lines
    else:
        assert not log.exists()


def test_grouped_static_choice_rule():
    """The planner's static choice per problem (no tuning table; no GPU
    needed): problems with m, n, k <= 64 run whole in one work item (fastmul!'s
    form, src/kernels.jl:202-208); other FP64 problems take the tile with the
    least padded area divided by the tile's measured relative rate (128x128
    and 128x64: 1, 64x128: 0.97, 64x64: 0.9, 32x32: 0.65); 3xBF16 problems the
    64x128 tile; native FP32 problems the 64x64 tile from max(m, n) >= 96, else
    32x32; one split (a plan adds splits to fill the device)."""
    from jblas_amd import grouped as G
    G.tune_reset()
    for m, n, k in [(1, 1, 1), (64, 64, 64), (16, 40, 8), (64, 1, 64), (5, 7, 0)]:
        assert G.tune_lookup("f64", m, n, k) == ("small", 1), (m, n, k)
        assert G.tune_lookup("f32_bf16x3", m, n, k) == ("small", 1), (m, n, k)
    rate = {"128x128": 1.0, "128x64": 1.0, "64x128": 0.97, "64x64": 0.90, "32x32": 0.65}

    def rule(m, n):  # first minimum in the order listed
        def padded(t):
            bm, bn = map(int, t.split("x"))
            return -(-m // bm) * bm * (-(-n // bn) * bn) / rate[t]
        return min(rate, key=padded)

    for (m, n, k, want) in [(1024, 1024, 1024, "128x128"), (1000, 200, 64, "128x128"), (1024, 64, 512, "128x64"),
                            (64, 1024, 512, "64x128"), (65, 64, 64, "128x64"), (96, 32, 64, "32x32"),
                            (64, 64, 4096, "64x64"), (16, 16, 4096, "32x32"), (200, 72, 96, "128x128"),
                            (1000, 136, 300, "128x64"), (136, 1000, 300, "64x128")]:
        assert rule(m, n) == want, (m, n)
        assert G.tune_lookup("f64", m, n, k) == (want, 1), (m, n, k)
    assert G.tune_lookup("f32", 1024, 1024, 1024) == ("64x64", 1)  # no 128-row/column tiles in native FP32
    assert G.tune_lookup("f32", 65, 64, 64) == ("32x32", 1)
    for m, n, k in [(1024, 1024, 1024), (1024, 64, 512), (65, 64, 64), (16, 16, 4096), (2048, 16, 16)]:
        assert G.tune_lookup("f32_bf16x3", m, n, k) == ("64x128", 1), (m, n, k)
