"""Grouped (batched) GEMM, hipGraph plans and the autotuner.

The reference has no batched path; its tile choice is static
(pick_kernel_size, src/kernel_structure.jl:76-99; blocking_structure,
src/memory_management.jl:78-140).  Here many GEMMs of uneven size run in one
persistent launch (optionally captured in a hipGraph), with the tile shape and
split-K factor of every size bucket taken from a measured table.

Operands are column-major GPU tensors (stride(0) == 1), as for gemm_.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Any, Sequence

from . import _lib
from .api import DimensionMismatch, _Operand, _stream_of, _trans

METHODS = {"f64": _lib.JB_F64_NATIVE, "f32": _lib.JB_F32_NATIVE, "f32_bf16x3": _lib.JB_F32_BF16X3}
# work-item configurations: MFMA tiles, the whole problem on the vector ALUs
# (m, n, k <= 64, the fastmul! form), and the FP64 tiles with 128 rows or
# columns (their own persistent kernels)
TILES = {0: "64x64", 1: "32x32", 2: "small", 3: "128x128", 4: "128x64", 5: "64x128"}


def method_id(method) -> int:
    if isinstance(method, int):
        return method
    try:
        return METHODS[method]
    except KeyError:
        raise ValueError(f"method must be one of {sorted(METHODS)}") from None


@dataclass
class GemmDesc:
    """One problem: C = alpha*op(A)*op(B) + beta*C (Julia gemm! argument meaning)."""
    tA: str
    tB: str
    alpha: float
    A: Any
    B: Any
    beta: float
    C: Any

    def to_c(self) -> _lib.GemmDescC:
        a, b, c = _Operand(self.A, False, "A"), _Operand(self.B, False, "B"), _Operand(self.C, True, "C")
        if not (a.device and b.device and c.device):
            raise ValueError("grouped GEMM operands must be GPU tensors")
        m, k = (a.rows, a.cols) if self.tA.upper() == "N" else (a.cols, a.rows)
        k2, n = (b.rows, b.cols) if self.tB.upper() == "N" else (b.cols, b.rows)
        if k != k2 or (c.rows, c.cols) != (m, n):
            raise DimensionMismatch(f"op(A) {m}x{k}, op(B) {k2}x{n}, C {c.rows}x{c.cols}")
        d = _lib.GemmDescC()
        d.transA, d.transB = _trans(self.tA), _trans(self.tB)
        d.m, d.n, d.k = m, n, k
        d.alpha, d.beta = float(self.alpha), float(self.beta)
        d.A, d.lda, d.B, d.ldb, d.C, d.ldc = a.ptr, a.ld, b.ptr, b.ld, c.ptr, c.ld
        return d


def _desc_array(descs: Sequence[GemmDesc]):
    arr = (_lib.GemmDescC * max(1, len(descs)))()
    for i, d in enumerate(descs):
        arr[i] = d.to_c()
    return arr


def gemm_grouped_(method, descs: Sequence[GemmDesc], *, stream=None) -> None:
    """All problems in one persistent launch (synchronous on the stream)."""
    arr = _desc_array(descs)
    _lib.call("jb_gemm_grouped_dev", method_id(method), len(descs), C.cast(arr, C.c_void_p),
              _stream_of(True, stream))


class GroupedPlan:
    """Descriptors uploaded once; the persistent launch captured in a hipGraph.
    `launch()` replays the graph on the stream.  Keep the operand tensors alive
    while the plan exists (the graph holds their addresses)."""

    def __init__(self, method, descs: Sequence[GemmDesc], *, stream=None):
        self._descs = list(descs)  # keep tensors referenced
        arr = _desc_array(self._descs)
        h = C.c_void_p()
        _lib.call("jb_grouped_plan_create", method_id(method), len(self._descs), C.cast(arr, C.c_void_p),
                  _stream_of(True, stream), C.byref(h))
        self._h = h

    def launch(self, *, stream=None) -> None:
        if self._h is None:
            raise RuntimeError("plan destroyed")
        _lib.call("jb_grouped_plan_launch", self._h, _stream_of(True, stream))

    def destroy(self) -> None:
        if self._h is not None:
            _lib.call("jb_grouped_plan_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---- autotuner ---------------------------------------------------------------
def autotune(method, descs: Sequence[GemmDesc], reps: int = 3, *, stream=None) -> None:
    """Tune (tile, split-K) per size bucket present in `descs`, timing the
    whole group for every candidate (coordinate descent over buckets)."""
    arr = _desc_array(descs)
    _lib.call("jb_autotune_grouped", method_id(method), len(descs), C.cast(arr, C.c_void_p), reps,
              _stream_of(True, stream))


def tune_save(path: str) -> None:
    _lib.call("jb_autotune_save", path.encode())


def tune_load(path: str) -> None:
    _lib.call("jb_autotune_load", path.encode())


def tune_reset() -> None:
    _lib.call("jb_autotune_reset")


def tune_lookup(method, m: int, n: int, k: int):
    """(tile name, split-K) the table or the static rule gives this shape on its
    own.  A plan may raise the split further to fill the device (grouped_choices
    reports the plan's final choice)."""
    t, s = C.c_int(), C.c_int()
    _lib.call("jb_autotune_lookup", method_id(method), m, n, k, C.byref(t), C.byref(s))
    return TILES[t.value], s.value


def split_chunk(k: int, splitk: int) -> int:
    """k per split used by the grouped launch (a multiple of 32, never an empty split)."""
    kc = -(-k // max(splitk, 1))
    return max(32, -(-kc // 32) * 32)


def grouped_choices(method, descs: Sequence[GemmDesc]):
    """Per problem, the (tile name, split-K, k per split) a plan of this group
    would use now (None for a problem with m or n == 0).  Split s sums
    k-values [s*kchunk, min(k, (s+1)*kchunk)); the splits' partial results are
    added in split order."""
    arr = _desc_array(descs)
    cnt = max(1, len(descs))
    t, s, kc = (C.c_int * cnt)(), (C.c_int * cnt)(), (C.c_int64 * cnt)()
    _lib.call("jb_grouped_choices", method_id(method), len(descs), C.cast(arr, C.c_void_p), t, s, kc)
    return [None if t[i] < 0 else (TILES[t[i]], s[i], kc[i]) for i in range(len(descs))]
