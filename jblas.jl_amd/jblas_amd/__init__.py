"""jblas_amd -- host layer of the MI355X GEMM engine behind jBLAS.jl's GEMM path.

The compute lives in libjblas_hip.so (hand-written gfx950 HIP kernels, C ABI in
include/jblas_hip.h); this package binds it with ctypes and mirrors the
reference's entry points (jmul!, kernel!, initkernel!) plus the BLAS-style
gemm!, grouped launches, the autotuner and the gemm! pullback.
"""
from ._lib import (JB_F32_BF16X3, JB_F32_NATIVE, JB_F64_NATIVE, JB_F64_OZAKI_BASE, JB_F64_RECURRENCE, JBError,
                   device_info, finalize, init, kernel_time, kernel_timing, kernel_timing_reset, last_error, lib)
from .api import (DimensionMismatch, device_matrix, fill_normal_, fill_uniform_, fill_wide_, gemm_, initkernel_, jmul_, fastmul_,
                  kernel_, nonfinite_check)
from .grouped import (GemmDesc, GroupedPlan, autotune, gemm_grouped_, grouped_choices, split_chunk, tune_load,
                      tune_lookup, tune_reset, tune_save)

__all__ = [
    "JBError", "DimensionMismatch", "lib", "init", "finalize", "last_error", "device_info",
    "kernel_timing", "kernel_time", "kernel_timing_reset",
    "gemm_", "jmul_", "fastmul_", "kernel_", "initkernel_", "fill_uniform_", "fill_wide_", "fill_normal_",
    "device_matrix",
    "nonfinite_check",
    "GemmDesc", "GroupedPlan", "gemm_grouped_", "autotune", "tune_save", "tune_load", "tune_reset",
    "tune_lookup", "split_chunk", "grouped_choices",
    "JB_F64_NATIVE", "JB_F64_RECURRENCE", "JB_F32_NATIVE", "JB_F32_BF16X3", "JB_F64_OZAKI_BASE",
]


def __getattr__(name):
    # the autograd helpers import torch; load them on first use
    if name in ("rrule", "rrule_gemm", "pullback_", "pullback_grouped_", "matmul", "GemmFunction"):
        from . import rrule as _r

        return _r if name == "rrule" else getattr(_r, name)
    raise AttributeError(name)
