"""jblas_amd -- host layer of the MI355X GEMM engine behind jBLAS.jl's GEMM path.

The compute lives in libjblas_hip.so (hand-written gfx950 HIP kernels, C ABI in
include/jblas_hip.h); this package binds it with ctypes and mirrors the
reference's entry points (jmul!, kernel!, initkernel!) plus the BLAS-style
gemm!, grouped launches, the autotuner and the gemm! pullback.
"""
from ._lib import (JB_F32_BF16X3, JB_F32_NATIVE, JB_F64_NATIVE, JB_F64_OZAKI_BASE, JB_F64_RECURRENCE, JBError,
                   device_info, finalize, init, last_error, lib)
from .api import (DimensionMismatch, device_matrix, fill_uniform_, fill_wide_, gemm_, initkernel_, jmul_,
                  kernel_)

__all__ = [
    "JBError", "DimensionMismatch", "lib", "init", "finalize", "last_error", "device_info",
    "gemm_", "jmul_", "kernel_", "initkernel_", "fill_uniform_", "fill_wide_", "device_matrix",
    "JB_F64_NATIVE", "JB_F64_RECURRENCE", "JB_F32_NATIVE", "JB_F32_BF16X3", "JB_F64_OZAKI_BASE",
]
