"""ctypes front end of the CPU checker (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, as the checker.  The product never imports it.

Every function works on numpy arrays in column-major (Fortran) order, the
layout of Julia's Matrix and of the reference's MMatrix (src/gemm.jl:3-11 stores
with leading dimension M).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in (line + " ")
    except OSError:
        pass
    return False


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    """Load the ISA variant that matches this host (build it if missing)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    name = "liboracle_avx512.so" if host_has_avx512() else "liboracle_avx2.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    i64, dbl, p = C.c_int64, C.c_double, C.c_void_p
    L.orc_splitmix64.restype = C.c_uint64
    L.orc_splitmix64.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_fill_uniform.argtypes = [C.c_uint64, i64, p]
    L.orc_fill_wide.argtypes = [C.c_uint64, i64, C.c_int, C.c_int, p]
    L.orc_jmul_d.argtypes = [i64, i64, i64, p, i64, p, i64, p, i64, C.c_int, C.c_int]
    L.orc_jmul_s.argtypes = [i64, i64, i64, p, i64, p, i64, p, i64, C.c_int, C.c_int]
    L.orc_jmul_naive_d.argtypes = [i64, i64, i64, p, i64, p, i64, p, i64]
    L.orc_gemm_recurrence.argtypes = [C.c_char, C.c_char, i64, i64, i64, dbl, p, i64, p, i64, dbl, p, i64, C.c_int]
    L.orc_gemm_recurrence_sample.argtypes = [C.c_char, C.c_char, i64, dbl, p, i64, p, i64, dbl, p, i64, p, p, p,
                                             C.c_int]
    L.orc_gemm_recurrence_chunked.argtypes = [C.c_char, C.c_char, i64, i64, i64, dbl, p, i64, p, i64, dbl, p, i64,
                                              i64, C.c_int]
    L.orc_gemm_dot2.argtypes = [C.c_char, C.c_char, i64, i64, i64, dbl, p, i64, p, i64, dbl, p, i64, C.c_int]
    L.orc_gemm_dot2_sample.argtypes = [C.c_char, C.c_char, i64, dbl, p, i64, p, i64, dbl, p, i64, p, p, p, C.c_int]
    L.orc_sgemm_dot2.argtypes = [C.c_char, C.c_char, i64, i64, i64, dbl, p, i64, p, i64, dbl, p, i64, p, i64, C.c_int]
    L.orc_ozaki_split_a.argtypes = [C.c_char, i64, i64, p, i64, C.c_int, p, p]
    L.orc_ozaki_split_b.argtypes = [C.c_char, i64, i64, p, i64, C.c_int, p, p]
    L.orc_dgemm_ozaki.argtypes = [C.c_char, C.c_char, i64, i64, i64, dbl, p, i64, p, i64, dbl, p, i64, C.c_int, C.c_int]
    L.orc_ozaki_level_chunk.argtypes = [C.c_int]
    L.orc_ozaki_level_chunk.restype = C.c_int
    L.orc_ozaki_i8_split_a.argtypes = [C.c_char, i64, i64, p, i64, C.c_int, p, p]
    L.orc_ozaki_i8_split_b.argtypes = [C.c_char, i64, i64, p, i64, C.c_int, p, p]
    L.orc_dgemm_ozaki_i8.argtypes = [C.c_char, C.c_char, i64, i64, i64, dbl, p, i64, p, i64, dbl, p, i64, C.c_int,
                                     C.c_int]
    L.orc_ozaki_i8_chunk.argtypes = []
    L.orc_ozaki_i8_chunk.restype = C.c_int
    L.orc_threads_used.argtypes = [C.c_int]
    L.orc_threads_used.restype = C.c_int
    L.orc_kernel_acc_d.argtypes = [i64, i64, i64, p, i64, p, i64, p, i64]
    L.orc_kernel_acc_d.restype = None
    L.orc_pick_kernel_size.argtypes = [i64, i64, i64, p]
    L.orc_blocking_structure.argtypes = [i64, i64, i64, i64, p, i64, i64, p]
    for fn in ("orc_jmul_d", "orc_jmul_s", "orc_dgemm_ozaki", "orc_ozaki_split_a", "orc_ozaki_split_b",
               "orc_dgemm_ozaki_i8", "orc_ozaki_i8_split_a", "orc_ozaki_i8_split_b", "orc_blocking_structure"):
        getattr(L, fn).restype = C.c_int
    _LIB = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def threads_used(nthreads: int) -> int:
    """Threads a parallel region of the restatement actually runs with."""
    return int(lib().orc_threads_used(nthreads))


def nthreads_default() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def _f(a, dtype=np.float64):
    a = np.asarray(a, dtype=dtype)
    return np.asfortranarray(a)


def _ld(a: np.ndarray) -> int:
    return max(1, a.shape[0])


# --- generator ------------------------------------------------------------
def splitmix64(seed: int, i: int) -> int:
    return int(lib().orc_splitmix64(seed, i))


def fill_uniform(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().orc_fill_uniform(seed, n, _ptr(out))
    return out


def fill_wide(seed: int, n: int, emin: int = -12, emax: int = 12) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().orc_fill_wide(seed, n, emin, emax, _ptr(out))
    return out


# --- jmul! restatement ----------------------------------------------------
def jmul(A, X, avx512: bool | None = None, nthreads: int = 1) -> np.ndarray:
    """D = A*X with the reference recurrence (fringe included)."""
    dt = np.float32 if np.asarray(A).dtype == np.float32 else np.float64
    A, X = _f(A, dt), _f(X, dt)
    M, N = A.shape
    N2, P = X.shape
    assert N == N2
    D = np.zeros((M, P), dtype=dt, order="F")
    if avx512 is None:
        avx512 = host_has_avx512()
    fn = lib().orc_jmul_s if dt == np.float32 else lib().orc_jmul_d
    rc = fn(M, N, P, _ptr(A), _ld(A), _ptr(X), _ld(X), _ptr(D), _ld(D), int(avx512), nthreads)
    if rc != 0:
        raise ValueError(f"orc_jmul rc={rc}")
    return D


def jmul_naive(A, X) -> np.ndarray:
    A, X = _f(A), _f(X)
    M, N = A.shape
    P = X.shape[1]
    D = np.zeros((M, P), order="F")
    lib().orc_jmul_naive_d(M, N, P, _ptr(A), _ld(A), _ptr(X), _ld(X), _ptr(D), _ld(D))
    return D


def kernel_acc(D, A, X) -> np.ndarray:
    """kernel! restatement: returns D + A*X with each chain starting from D."""
    A, X = _f(A), _f(X)
    D = np.array(D, dtype=np.float64, order="F", copy=True)
    M, N = A.shape
    P = X.shape[1]
    lib().orc_kernel_acc_d(M, N, P, _ptr(A), _ld(A), _ptr(X), _ld(X), _ptr(D), _ld(D))
    return D


def _dims(ta, tb, A, B):
    m, k = (A.shape if ta == "N" else A.shape[::-1])
    k2, n = (B.shape if tb == "N" else B.shape[::-1])
    assert k == k2, (A.shape, B.shape, ta, tb)
    return m, n, k


def gemm_recurrence(ta, tb, alpha, A, B, beta, C, nthreads=None) -> np.ndarray:
    A, B = _f(A), _f(B)
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    m, n, k = _dims(ta, tb, A, B)
    lib().orc_gemm_recurrence(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B),
                              beta, _ptr(C), _ld(C), nthreads or nthreads_default())
    return C


def gemm_recurrence_chunked(ta, tb, alpha, A, B, beta, C, kc, nthreads=None) -> np.ndarray:
    """The recurrence with k cut into chunks of kc, chunk results summed in order."""
    A, B = _f(A), _f(B)
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    m, n, k = _dims(ta, tb, A, B)
    lib().orc_gemm_recurrence_chunked(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B),
                                      beta, _ptr(C), _ld(C), int(kc), nthreads or nthreads_default())
    return C


def gemm_recurrence_sample(ta, tb, alpha, A, B, beta, C0_vals, ii, jj, nthreads=None) -> np.ndarray:
    A, B = _f(A), _f(B)
    m, n, k = _dims(ta, tb, A, B)
    ii = np.ascontiguousarray(ii, dtype=np.int64)
    jj = np.ascontiguousarray(jj, dtype=np.int64)
    c0 = np.ascontiguousarray(C0_vals if C0_vals is not None else np.zeros(len(ii)), dtype=np.float64)
    out = np.empty(len(ii))
    lib().orc_gemm_recurrence_sample(ta.encode(), tb.encode(), k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta,
                                     _ptr(c0), len(ii), _ptr(ii), _ptr(jj), _ptr(out), nthreads or nthreads_default())
    return out


def gemm_dot2(ta, tb, alpha, A, B, beta, C, nthreads=None) -> np.ndarray:
    """FP64 oracle: alpha*op(A)*op(B) + beta*C, Dot2-compensated."""
    A, B = _f(A), _f(B)
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    m, n, k = _dims(ta, tb, A, B)
    lib().orc_gemm_dot2(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B),
                        beta, _ptr(C), _ld(C), nthreads or nthreads_default())
    return C


def gemm_dot2_sample(ta, tb, alpha, A, B, beta, C0_vals, ii, jj, nthreads=None) -> np.ndarray:
    A, B = _f(A), _f(B)
    m, n, k = _dims(ta, tb, A, B)
    ii = np.ascontiguousarray(ii, dtype=np.int64)
    jj = np.ascontiguousarray(jj, dtype=np.int64)
    c0 = np.ascontiguousarray(C0_vals if C0_vals is not None else np.zeros(len(ii)), dtype=np.float64)
    out = np.empty(len(ii))
    lib().orc_gemm_dot2_sample(ta.encode(), tb.encode(), k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta,
                               _ptr(c0), len(ii), _ptr(ii), _ptr(jj), _ptr(out), nthreads or nthreads_default())
    return out


def sgemm_dot2(ta, tb, alpha, A, B, beta, C, nthreads=None) -> np.ndarray:
    """FP64 oracle for float32 inputs (returns float64)."""
    A, B = _f(A, np.float32), _f(B, np.float32)
    C = _f(C, np.float32)
    m, n, k = _dims(ta, tb, A, B)
    out = np.zeros((m, n), order="F")
    lib().orc_sgemm_dot2(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta,
                         _ptr(C), _ld(C), _ptr(out), _ld(out), nthreads or nthreads_default())
    return out


def ozaki_split_a(ta, A, s):
    A = _f(A)
    m, k = (A.shape if ta == "N" else A.shape[::-1])
    e = np.zeros(m, dtype=np.int32)
    d = np.zeros((s, m, k), dtype=np.int16)
    rc = lib().orc_ozaki_split_a(ta.encode(), m, k, _ptr(A), _ld(A), s, _ptr(e), _ptr(d))
    if rc:
        raise ValueError("non-finite input")
    return e, d


def ozaki_split_b(tb, B, s):
    B = _f(B)
    k, n = (B.shape if tb == "N" else B.shape[::-1])
    e = np.zeros(n, dtype=np.int32)
    d = np.zeros((s, n, k), dtype=np.int16)
    rc = lib().orc_ozaki_split_b(tb.encode(), k, n, _ptr(B), _ld(B), s, _ptr(e), _ptr(d))
    if rc:
        raise ValueError("non-finite input")
    return e, d


def dgemm_ozaki(ta, tb, alpha, A, B, beta, C, s, nthreads=None) -> np.ndarray:
    """Exact restatement of the split-BF16 emulated DGEMM with s slices."""
    A, B = _f(A), _f(B)
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    m, n, k = _dims(ta, tb, A, B)
    rc = lib().orc_dgemm_ozaki(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta,
                               _ptr(C), _ld(C), s, nthreads or nthreads_default())
    if rc:
        raise ValueError(f"orc_dgemm_ozaki rc={rc}")
    return C


def ozaki_level_chunk(L: int) -> int:
    """k-chunk length over which the FP32 sum of 0-based level L (L+1 slice
    pairs) stays exact in the split-BF16 path."""
    return int(lib().orc_ozaki_level_chunk(L))


def ozaki_i8_split_a(ta, A, s):
    """INT8-slice split of op(A)'s rows: exponents (m,), digits (s, m, k) in [-128, 127]."""
    A = _f(A)
    m, k = (A.shape if ta == "N" else A.shape[::-1])
    e = np.zeros(m, dtype=np.int32)
    d = np.zeros((s, m, k), dtype=np.int16)
    if lib().orc_ozaki_i8_split_a(ta.encode(), m, k, _ptr(A), _ld(A), s, _ptr(e), _ptr(d)):
        raise ValueError("non-finite input")
    return e, d


def ozaki_i8_split_b(tb, B, s):
    B = _f(B)
    k, n = (B.shape if tb == "N" else B.shape[::-1])
    e = np.zeros(n, dtype=np.int32)
    d = np.zeros((s, n, k), dtype=np.int16)
    if lib().orc_ozaki_i8_split_b(tb.encode(), k, n, _ptr(B), _ld(B), s, _ptr(e), _ptr(d)):
        raise ValueError("non-finite input")
    return e, d


def dgemm_ozaki_i8(ta, tb, alpha, A, B, beta, C, s, nthreads=None) -> np.ndarray:
    """Exact restatement of the split-INT8 emulated DGEMM with s slices."""
    A, B = _f(A), _f(B)
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    m, n, k = _dims(ta, tb, A, B)
    rc = lib().orc_dgemm_ozaki_i8(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta,
                                  _ptr(C), _ld(C), s, nthreads or nthreads_default())
    if rc:
        raise ValueError(f"orc_dgemm_ozaki_i8 rc={rc}")
    return C


def ozaki_i8_chunk() -> int:
    """k-chunk length over which the INT8 path's int32 level sums stay exact."""
    return int(lib().orc_ozaki_i8_chunk())


def pick_kernel_size(register_size: int, register_count: int, t_size: int):
    out = np.zeros(3, dtype=np.int64)
    lib().orc_pick_kernel_size(register_size, register_count, t_size, _ptr(out))
    return tuple(int(x) for x in out)


def blocking_structure(M, N, P, t_size, cache_size, register_size, register_count):
    cs = np.asarray(cache_size, dtype=np.int64)
    out = np.zeros(9, dtype=np.int64)
    lvl = lib().orc_blocking_structure(M, N, P, t_size, _ptr(cs), register_size, register_count, _ptr(out))
    o = [int(x) for x in out]
    return ((o[0], o[1], o[2]), (o[3], o[4], o[5]), (o[6], o[7], o[8])), int(lvl)
