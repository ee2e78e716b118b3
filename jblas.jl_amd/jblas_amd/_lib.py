"""ctypes binding of libjblas_hip.so (C ABI: include/jblas_hip.h).

This is the Python counterpart of the Julia `ccall` binding: the reference's
only FFI precedent is `ccall((:sym, lib), Cvoid, (Ptr{Float64},...), D, A, X)`
(test/runtests.jl:97-101).  The library is loaded from the in-tree build; a
missing library is an error, never a silent fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("JB_LIBRARY", os.path.join(PKG_ROOT, "lib", "libjblas_hip.so"))

JB_OK = 0
JB_ERR_HIP = 1
JB_ERR_ALLOC = 2
JB_ERR_NONFINITE = 3
JB_ERR_GRAPH = 4
JB_ERR_STATE = 5
JB_ERR_UNSUPPORTED = 6

JB_F64_NATIVE = 0
JB_F64_RECURRENCE = 1
JB_F32_NATIVE = 2
JB_F32_BF16X3 = 3
JB_F64_OZAKI_BASE = 16

_ERRNAMES = {JB_ERR_HIP: "JB_ERR_HIP", JB_ERR_ALLOC: "JB_ERR_ALLOC", JB_ERR_NONFINITE: "JB_ERR_NONFINITE",
             JB_ERR_GRAPH: "JB_ERR_GRAPH", JB_ERR_STATE: "JB_ERR_STATE", JB_ERR_UNSUPPORTED: "JB_ERR_UNSUPPORTED"}


class JBError(RuntimeError):
    """A libjblas_hip call returned a non-zero status."""

    def __init__(self, func: str, code: int, msg: str):
        self.func, self.code, self.msg = func, code, msg
        kind = f"argument {-code} invalid" if code < 0 else _ERRNAMES.get(code, f"error {code}")
        super().__init__(f"{func}: {kind}: {msg}")


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int32),
                ("clock_khz", C.c_int32), ("wavefront_size", C.c_int32), ("lds_bytes_per_cu", C.c_int32),
                ("l2_bytes", C.c_int64), ("hbm_bytes", C.c_int64), ("peak_fp64_tflops", C.c_double),
                ("peak_bf16_tflops", C.c_double), ("peak_fp32_tflops", C.c_double)]


class GemmDescC(C.Structure):
    _fields_ = [("transA", C.c_char), ("transB", C.c_char), ("pad_", C.c_char * 6), ("m", C.c_int64),
                ("n", C.c_int64), ("k", C.c_int64), ("alpha", C.c_double), ("beta", C.c_double),
                ("A", C.c_void_p), ("lda", C.c_int64), ("B", C.c_void_p), ("ldb", C.c_int64),
                ("C", C.c_void_p), ("ldc", C.c_int64)]


_i64, _dbl, _flt, _p, _ch, _int = C.c_int64, C.c_double, C.c_float, C.c_void_p, C.c_char, C.c_int

# name -> (restype, argtypes); the complete exported surface of include/jblas_hip.h
SIGNATURES = {
    "jb_init": (_int, [_int]),
    "jb_finalize": (_int, []),
    "jb_last_error": (C.c_char_p, []),
    "jb_version": (_int, []),
    "jb_get_device_info": (_int, [C.POINTER(DeviceInfo)]),
    "jb_kernel_timing_enable": (_int, [_int]),
    "jb_kernel_timing_read": (_int, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(_i64)]),
    "jb_kernel_timing_reset": (_int, []),
    "jb_jmul": (_int, [_i64, _i64, _i64, _p, _p, _p]),
    "jb_jmul_dev": (_int, [_i64, _i64, _i64, _p, _p, _p, _p]),
    "jb_jmul_f32": (_int, [_i64, _i64, _i64, _p, _p, _p]),
    "jb_jmul_f32_dev": (_int, [_i64, _i64, _i64, _p, _p, _p, _p]),
    "jb_kernel": (_int, [_i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64]),
    "jb_kernel_dev": (_int, [_i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64, _p]),
    "jb_initkernel": (_int, [_i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64]),
    "jb_initkernel_dev": (_int, [_i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64, _p]),
    "jb_fastmul": (_int, [_i64, _i64, _i64, _p, _p, _p]),
    "jb_fastmul_dev": (_int, [_i64, _i64, _i64, _p, _p, _p, _p]),
    "jb_fastmul_f32": (_int, [_i64, _i64, _i64, _p, _p, _p]),
    "jb_fastmul_f32_dev": (_int, [_i64, _i64, _i64, _p, _p, _p, _p]),
    "jb_dgemm": (_int, [_ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64]),
    "jb_dgemm_dev": (_int, [_ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64, _p]),
    "jb_dgemm_recurrence": (_int, [_ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64]),
    "jb_dgemm_recurrence_dev": (_int, [_ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64, _p]),
    "jb_dgemm_ozaki": (_int, [_ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64, _int]),
    "jb_dgemm_ozaki_dev": (_int, [_ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64, _int, _p]),
    "jb_dgemm_ozaki_noalloc_dev": (_int, [_ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64, _int,
                                          _p]),
    "jb_dgemm_ozaki_i8": (_int, [_ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64, _int]),
    "jb_dgemm_ozaki_i8_dev": (_int, [_ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64, _int,
                                     _p]),
    "jb_sgemm": (_int, [_ch, _ch, _i64, _i64, _i64, _flt, _p, _i64, _p, _i64, _flt, _p, _i64]),
    "jb_sgemm_dev": (_int, [_ch, _ch, _i64, _i64, _i64, _flt, _p, _i64, _p, _i64, _flt, _p, _i64, _p]),
    "jb_sgemm_bf16x3": (_int, [_ch, _ch, _i64, _i64, _i64, _flt, _p, _i64, _p, _i64, _flt, _p, _i64, _int]),
    "jb_sgemm_bf16x3_dev": (_int, [_ch, _ch, _i64, _i64, _i64, _flt, _p, _i64, _p, _i64, _flt, _p, _i64, _int, _p]),
    "jb_nonfinite_check": (_int, [_p]),
    "jb_gemm_grouped_dev": (_int, [_int, _i64, _p, _p]),
    "jb_grouped_plan_create": (_int, [_int, _i64, _p, _p, C.POINTER(_p)]),
    "jb_grouped_plan_launch": (_int, [_p, _p]),
    "jb_grouped_plan_destroy": (_int, [_p]),
    "jb_autotune_grouped": (_int, [_int, _i64, _p, _int, _p]),
    "jb_autotune_save": (_int, [C.c_char_p]),
    "jb_autotune_load": (_int, [C.c_char_p]),
    "jb_autotune_lookup": (_int, [_int, _i64, _i64, _i64, C.POINTER(_int), C.POINTER(_int)]),
    "jb_autotune_reset": (_int, []),
    "jb_grouped_choices": (_int, [_int, _i64, C.c_void_p, C.POINTER(_int), C.POINTER(_int), C.POINTER(_i64)]),
    "jb_gemm_pullback_dev": (_int, [_int, _ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _p, _i64, _p, _i64,
                                    _p, _i64, _p]),
    "jb_gemm_pullback": (_int, [_int, _ch, _ch, _i64, _i64, _i64, _dbl, _p, _i64, _p, _i64, _p, _i64, _p, _i64,
                                _p, _i64]),
    "jb_fill_uniform_dev": (_int, [C.c_uint64, _i64, _p, _p]),
    "jb_fill_wide_dev": (_int, [C.c_uint64, _i64, _int, _int, _p, _p]),
    "jb_fill_uniform_f32_dev": (_int, [C.c_uint64, _i64, _p, _p]),
    "jb_fill_wide_f32_dev": (_int, [C.c_uint64, _i64, _int, _int, _p, _p]),
    "jb_fill_normal_dev": (_int, [C.c_uint64, _i64, _dbl, _p, _p]),
    "jb_fill_normal_f32_dev": (_int, [C.c_uint64, _i64, _dbl, _p, _p]),
}

_lock = threading.Lock()
_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"libjblas_hip.so not built ({LIB_PATH}); run __graft_entry__.build()")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def last_error() -> str:
    return (lib().jb_last_error() or b"").decode(errors="replace")


def check(func: str, rc: int) -> None:
    if rc != JB_OK:
        raise JBError(func, rc, last_error())


def call(name: str, *args) -> None:
    check(name, getattr(lib(), name)(*args))


def init(device: int = 0) -> None:
    call("jb_init", device)


def finalize() -> None:
    call("jb_finalize")


def kernel_timing(on: bool) -> None:
    call("jb_kernel_timing_enable", 1 if on else 0)


def kernel_time(kernel: str):
    """(total milliseconds, launches) of `kernel` since the last reset."""
    ms, n = C.c_double(), C.c_int64()
    call("jb_kernel_timing_read", kernel.encode(), C.byref(ms), C.byref(n))
    return ms.value, n.value


def kernel_timing_reset() -> None:
    call("jb_kernel_timing_reset")


def device_info() -> dict:
    info = DeviceInfo()
    check("jb_get_device_info", lib().jb_get_device_info(C.byref(info)))
    return {f: (getattr(info, f).decode() if isinstance(getattr(info, f), bytes) else getattr(info, f))
            for f, _ in DeviceInfo._fields_}
