"""Host-side mirror of jBLAS.jl's GEMM entry points over libjblas_hip.so.

Reference interface mirrored (paths relative to the jBLAS.jl repository):
  jmul!(D, A, X)  "calculates D = A * X"            src/gemm.jl:225-348
  kernel!(pD, pA, pX, Kernel(...))  D += A*X         src/kernels.jl:239-241
  initkernel!(pD, pA, pX, Kernel(...))  D = A*X      src/kernels.jl:273-275
and the BLAS-style gemm! the drop-in exposes (Julia's
LinearAlgebra.BLAS.gemm!(tA, tB, alpha, A, B, beta, C) argument order).

Operands are column-major matrices, as in Julia:
  * numpy arrays (host memory) -> the synchronous host entry points;
  * torch tensors on the GPU whose stride(0) == 1 (a column-major view,
    e.g. torch.empty(n, m, device="cuda").T) -> the *_dev entry points on the
    current torch stream.
Shape mismatches raise DimensionMismatch before any library call (the
reference rejects them by dispatch: MMatrix{M,P} / {M,N} / {N,P} must agree,
src/gemm.jl:244).  Library failures raise JBError with the BLAS-style code.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import JBError  # noqa: F401  (re-export)

try:  # torch is only needed for device-resident operands
    import torch
except Exception:  # pragma: no cover
    torch = None


class DimensionMismatch(ValueError):
    """Operand shapes do not agree (Julia's DimensionMismatch)."""


def _is_torch(x) -> bool:
    return torch is not None and isinstance(x, torch.Tensor)


class _Operand:
    __slots__ = ("obj", "ptr", "ld", "rows", "cols", "device", "dtype")

    def __init__(self, x, writable: bool, name: str):
        if _is_torch(x):
            if x.dim() != 2:
                raise DimensionMismatch(f"{name} must be 2-D")
            if not x.is_cuda:
                raise ValueError(f"{name}: torch tensors must live on the GPU (use numpy for host data)")
            rows, cols = x.shape
            if rows > 1 and x.stride(0) != 1:
                raise ValueError(f"{name}: device matrices must be column-major (stride(0) == 1)")
            self.obj, self.ptr, self.device = x, x.data_ptr(), True
            self.ld = max(1, x.stride(1) if cols > 1 else rows)
            self.dtype = {torch.float64: np.float64, torch.float32: np.float32}.get(x.dtype)
            if self.dtype is None:
                raise TypeError(f"{name}: dtype {x.dtype} not supported")
        else:
            a = np.asarray(x) if not isinstance(x, np.ndarray) else x
            if a.ndim != 2:
                raise DimensionMismatch(f"{name} must be 2-D")
            if a.dtype not in (np.float64, np.float32):
                raise TypeError(f"{name}: dtype {a.dtype} not supported")
            rows, cols = a.shape
            colmajor = rows <= 1 or a.strides[0] == a.itemsize
            if not colmajor:
                if writable:
                    raise ValueError(f"{name}: output must be column-major (order='F')")
                a = np.asfortranarray(a)
            self.obj, self.ptr, self.device = a, a.ctypes.data, False
            self.ld = max(1, a.strides[1] // a.itemsize if cols > 1 else rows)
            self.dtype = a.dtype.type
        self.rows, self.cols = rows, cols


def _stream_of(device: bool, stream):
    if not device:
        return None
    if stream is not None:
        return stream if isinstance(stream, int) else stream.cuda_stream
    return torch.cuda.current_stream().cuda_stream


def _trans(t: str) -> bytes:
    if t not in ("N", "T", "C", "n", "t", "c"):
        raise ValueError(f"trans must be 'N', 'T' or 'C', got {t!r}")
    return t.upper().encode()


def _check_shapes(tA, tB, a: _Operand, b: _Operand, c: _Operand):
    m, k = (a.rows, a.cols) if tA.upper() == "N" else (a.cols, a.rows)
    k2, n = (b.rows, b.cols) if tB.upper() == "N" else (b.cols, b.rows)
    if k != k2 or c.rows != m or c.cols != n:
        raise DimensionMismatch(f"op(A) is {m}x{k}, op(B) is {k2}x{n}, C is {c.rows}x{c.cols}")
    if not (a.device == b.device == c.device):
        raise ValueError("operands must all be host (numpy) or all device (torch) arrays")
    if not (a.dtype == b.dtype == c.dtype):
        raise TypeError("operands must share one element type")
    return m, n, k


METHODS_F64 = ("native", "recurrence", "ozaki", "ozaki_int8")
# slice counts whose componentwise error is at or below native FP64's (8192^3, wide-range inputs)
DEFAULT_NSLICES = {"ozaki": 6, "ozaki_int8": 7}
METHODS_F32 = ("native", "bf16x3")


def gemm_(tA: str, tB: str, alpha, A, B, beta, C, *, method: str = "native", nslices: int | None = None,
          nprod: int = 6, stream=None):
    """C = alpha*op(A)*op(B) + beta*C in place; returns C.

    method (Float64): "native" FP64 MFMA, "recurrence" (bit-exact jmul! order),
    "ozaki" (split-BF16 emulation with `nslices` slices of 9 bits, default 6),
    "ozaki_int8" (INT8 slices of 8 bits on int8 MFMA, default 7).
    method (Float32): "native" FP32 MFMA, "bf16x3" (3 BF16 slices, `nprod` products).
    """
    a, b, c = _Operand(A, False, "A"), _Operand(B, False, "B"), _Operand(C, True, "C")
    m, n, k = _check_shapes(tA, tB, a, b, c)
    ta, tb = _trans(tA), _trans(tB)
    dev = a.device
    s = _stream_of(dev, stream)
    common = (ta, tb, m, n, k)
    if c.dtype == np.float64:
        if method == "native":
            name = "jb_dgemm_dev" if dev else "jb_dgemm"
            args = common + (float(alpha), a.ptr, a.ld, b.ptr, b.ld, float(beta), c.ptr, c.ld)
        elif method == "recurrence":
            name = "jb_dgemm_recurrence_dev" if dev else "jb_dgemm_recurrence"
            args = common + (float(alpha), a.ptr, a.ld, b.ptr, b.ld, float(beta), c.ptr, c.ld)
        elif method in ("ozaki", "ozaki_int8"):
            name = "jb_dgemm_ozaki" + ("_i8" if method == "ozaki_int8" else "") + ("_dev" if dev else "")
            if nslices is None:
                nslices = DEFAULT_NSLICES[method]
            args = common + (float(alpha), a.ptr, a.ld, b.ptr, b.ld, float(beta), c.ptr, c.ld, int(nslices))
        else:
            raise ValueError(f"Float64 method must be one of {METHODS_F64}")
    else:
        if method == "native":
            name = "jb_sgemm_dev" if dev else "jb_sgemm"
            args = common + (float(alpha), a.ptr, a.ld, b.ptr, b.ld, float(beta), c.ptr, c.ld)
        elif method == "bf16x3":
            name = "jb_sgemm_bf16x3_dev" if dev else "jb_sgemm_bf16x3"
            args = common + (float(alpha), a.ptr, a.ld, b.ptr, b.ld, float(beta), c.ptr, c.ld, int(nprod))
        else:
            raise ValueError(f"Float32 method must be one of {METHODS_F32}")
    if dev:
        args = args + (s,)
    _lib.call(name, *args)
    return C


def nonfinite_check(*, stream=None) -> None:
    """Raises JBError (JB_ERR_NONFINITE) if an emulated call on the device
    (split-BF16, split-INT8, 3xBF16) met Inf or NaN in an operand since the
    previous check; synchronises `stream` (default: torch's current stream)."""
    _lib.call("jb_nonfinite_check", _stream_of(True, stream))


def jmul_(D, A, X, *, stream=None):
    """D = A*X (jmul!, src/gemm.jl:244-348): the reference's per-element fma
    recurrence, bit-identical, with the fringe rows/columns written too.
    Float64 (FP64 MFMA kernel, whose chains are that recurrence) or Float32
    (the recurrence in binary32 on the sequential-fma kernel; the reference's
    jmul! is generic in T, dense operands only)."""
    d, a, x = _Operand(D, True, "D"), _Operand(A, False, "A"), _Operand(X, False, "X")
    M, N = a.rows, a.cols
    if x.rows != N or d.rows != M or d.cols != x.cols:
        raise DimensionMismatch(f"D is {d.rows}x{d.cols}, A is {M}x{N}, X is {x.rows}x{x.cols}")
    if not (d.dtype == a.dtype == x.dtype and d.dtype in (np.float64, np.float32)):
        raise TypeError("jmul! drop-in takes Float64 or Float32 matrices of one type")
    P = x.cols
    dense = d.ld == max(1, M) and a.ld == max(1, M) and x.ld == max(1, N)
    suffix = "" if d.dtype == np.float64 else "_f32"
    if not dense:
        if suffix:
            raise ValueError("Float32 jmul! takes dense column-major matrices (leading dimensions M, N, M)")
        # Strided views go through the BLAS form with explicit leading dimensions.
        return gemm_("N", "N", 1.0, A, X, 0.0, D, method="recurrence", stream=stream)
    if d.device:
        _lib.call("jb_jmul" + suffix + "_dev", M, N, P, d.ptr, a.ptr, x.ptr, _stream_of(True, stream))
    else:
        _lib.call("jb_jmul" + suffix, M, N, P, d.ptr, a.ptr, x.ptr)
    return D


def fastmul_(D, A, X, *, stream=None):
    """D = A*X as one small-GEMM launch (fastmul!, src/kernels.jl:202-208):
    the same per-element recurrence as jmul_, Float64 or Float32, dense
    column-major operands (leading dimensions M, N, M)."""
    d, a, x = _Operand(D, True, "D"), _Operand(A, False, "A"), _Operand(X, False, "X")
    M, N = a.rows, a.cols
    if x.rows != N or d.rows != M or d.cols != x.cols:
        raise DimensionMismatch(f"D is {d.rows}x{d.cols}, A is {M}x{N}, X is {x.rows}x{x.cols}")
    if not (d.dtype == a.dtype == x.dtype) or d.dtype not in (np.float64, np.float32):
        raise TypeError("fastmul! takes three Float64 or three Float32 matrices")
    P = x.cols
    if d.ld != max(1, M) or a.ld != max(1, M) or x.ld != max(1, N):
        raise ValueError("fastmul! takes dense column-major matrices (leading dimensions M, N, M)")
    f64 = d.dtype == np.float64
    if d.device:
        _lib.call("jb_fastmul_dev" if f64 else "jb_fastmul_f32_dev", M, N, P, d.ptr, a.ptr, x.ptr,
                  _stream_of(True, stream))
    else:
        _lib.call("jb_fastmul" if f64 else "jb_fastmul_f32", M, N, P, d.ptr, a.ptr, x.ptr)
    return D


def _kernel_call(name, D, A, X, stream):
    d, a, x = _Operand(D, True, "D"), _Operand(A, False, "A"), _Operand(X, False, "X")
    M, N = a.rows, a.cols
    if x.rows != N or d.rows != M or d.cols != x.cols:
        raise DimensionMismatch(f"D is {d.rows}x{d.cols}, A is {M}x{N}, X is {x.rows}x{x.cols}")
    if not (d.dtype == a.dtype == x.dtype == np.float64):
        raise TypeError("kernel!/initkernel! take Float64 matrices")
    P = x.cols
    if d.device:
        _lib.call(name + "_dev", M, N, P, d.ptr, d.ld, a.ptr, a.ld, x.ptr, x.ld, _stream_of(True, stream))
    else:
        _lib.call(name, M, N, P, d.ptr, d.ld, a.ptr, a.ld, x.ptr, x.ld)
    return D


def kernel_(D, A, X, *, stream=None):
    """D += A*X on (sub)matrices with their own leading dimensions (kernel!,
    src/kernels.jl:212-241): each element's fma chain starts from D."""
    return _kernel_call("jb_kernel", D, A, X, stream)


def initkernel_(D, A, X, *, stream=None):
    """D = A*X on (sub)matrices (initkernel!, src/kernels.jl:242-275): plain
    first product, then fma, as jmul!."""
    return _kernel_call("jb_initkernel", D, A, X, stream)


# ---- seeded device generator (mrandn replacement, src/randmat.jl:11-14) ------
def fill_uniform_(out, seed: int, *, stream=None):
    """Fill a contiguous device tensor with SplitMix64 uniforms in [-1, 1)."""
    if not (_is_torch(out) and out.is_cuda and out.is_contiguous()):
        raise ValueError("out must be a contiguous GPU tensor")
    s = _stream_of(True, stream)
    if out.dtype == torch.float64:
        _lib.call("jb_fill_uniform_dev", seed, out.numel(), out.data_ptr(), s)
    elif out.dtype == torch.float32:
        _lib.call("jb_fill_uniform_f32_dev", seed, out.numel(), out.data_ptr(), s)
    else:
        raise TypeError("float64 or float32 only")
    return out


def fill_wide_(out, seed: int, emin: int = -12, emax: int = 12, *, stream=None):
    """Fill with u*2^e, u uniform in [-1,1), e uniform integer in [emin, emax]
    (float32: the float64 value rounded to nearest)."""
    if not (_is_torch(out) and out.is_cuda and out.is_contiguous()):
        raise ValueError("out must be a contiguous GPU tensor")
    s = _stream_of(True, stream)
    if out.dtype == torch.float64:
        _lib.call("jb_fill_wide_dev", seed, out.numel(), emin, emax, out.data_ptr(), s)
    elif out.dtype == torch.float32:
        _lib.call("jb_fill_wide_f32_dev", seed, out.numel(), emin, emax, out.data_ptr(), s)
    else:
        raise TypeError("float64 or float32 only")
    return out


def fill_normal_(out, seed: int, scale: float = 1.0, *, stream=None):
    """Fill with scale * N(0, 1) draws (Box-Muller on the SplitMix64 stream; the
    randn() of mrandn, src/randmat.jl:5-14).  Agrees with the host generator to
    a few ulp (device log and cos)."""
    if not (_is_torch(out) and out.is_cuda and out.is_contiguous()):
        raise ValueError("out must be a contiguous GPU tensor")
    s = _stream_of(True, stream)
    if out.dtype == torch.float64:
        _lib.call("jb_fill_normal_dev", seed, out.numel(), float(scale), out.data_ptr(), s)
    elif out.dtype == torch.float32:
        _lib.call("jb_fill_normal_f32_dev", seed, out.numel(), float(scale), out.data_ptr(), s)
    else:
        raise TypeError("float64 or float32 only")
    return out


def device_matrix(rows: int, cols: int, dtype=None, device="cuda"):
    """An uninitialised column-major device matrix (stride(0) == 1)."""
    dtype = dtype or torch.float64
    return torch.empty((cols, rows), dtype=dtype, device=device).T
