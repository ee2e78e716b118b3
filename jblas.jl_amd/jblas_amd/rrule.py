"""Reverse-mode rule of gemm! (the ChainRulesCore.rrule the Julia wrapper
defines, mirrored for Python callers).

For Y = alpha*op(A)*op(B) + beta*C and an upstream cotangent dY:
  dA    = alpha * dY * op(B)^T            (tA = 'N';  alpha * op(B) * dY^T for 'T')
  dB    = alpha * op(A)^T * dY            (tB = 'N';  alpha * dY^T * op(A) for 'T')
  dC    = beta * dY
  dalpha = <dY, op(A)*op(B)>,  dbeta = <dY, C>
dA and dB come from the library's pullback entry point: both adjoint GEMMs in
one grouped launch with the transposed-layout loaders.  The reference has no
differentiation (ChainRulesCore is absent from its Project.toml).
"""
from __future__ import annotations

from typing import Sequence

import torch

from . import _lib
from .api import _Operand, _stream_of, _trans, device_matrix, gemm_
from .grouped import GemmDesc, gemm_grouped_, method_id


def _method(dtype, method):
    if method is not None:
        return method_id(method)
    return _lib.JB_F64_NATIVE if dtype == torch.float64 else _lib.JB_F32_NATIVE


def _fwd_method(dtype, method):
    if dtype == torch.float64:
        return "native"
    return "bf16x3" if method_id(method or "f32") == _lib.JB_F32_BF16X3 else "native"


def pullback_(tA: str, tB: str, alpha: float, A, B, dY, dA=None, dB=None, *, method=None, stream=None):
    """dA, dB (either may be None) in one grouped launch; returns (dA, dB)."""
    a, b, dy = _Operand(A, False, "A"), _Operand(B, False, "B"), _Operand(dY, False, "dY")
    m, k = (a.rows, a.cols) if tA.upper() == "N" else (a.cols, a.rows)
    n = b.cols if tB.upper() == "N" else b.rows
    if (dy.rows, dy.cols) != (m, n):
        raise ValueError(f"dY must be {m}x{n}")
    da = _Operand(dA, True, "dA") if dA is not None else None
    db = _Operand(dB, True, "dB") if dB is not None else None
    _lib.call("jb_gemm_pullback_dev", _method(A.dtype, method), _trans(tA), _trans(tB), m, n, k, float(alpha),
              a.ptr, a.ld, b.ptr, b.ld, dy.ptr, dy.ld, da.ptr if da else None, da.ld if da else 1,
              db.ptr if db else None, db.ld if db else 1, _stream_of(True, stream))
    return dA, dB


def rrule_gemm(tA: str, tB: str, alpha: float, A, B, beta: float, C, *, method=None):
    """Returns (Y, pullback) with Y = alpha*op(A)*op(B) + beta*C (C untouched) and
    pullback(dY) -> (dA, dB, dalpha, dbeta, dC)."""
    m = A.shape[0] if tA.upper() == "N" else A.shape[1]
    n = B.shape[1] if tB.upper() == "N" else B.shape[0]
    P = device_matrix(m, n, A.dtype)
    gemm_(tA, tB, 1.0, A, B, 0.0, P, method=_fwd_method(A.dtype, method))
    Y = alpha * P + beta * C if beta != 0 else alpha * P
    Y = Y.T.contiguous().T  # column-major

    def pullback(dY):
        dA = device_matrix(A.shape[0], A.shape[1], A.dtype)
        dB = device_matrix(B.shape[0], B.shape[1], B.dtype)
        pullback_(tA, tB, alpha, A, B, dY, dA, dB, method=method)
        dalpha = torch.sum(dY * P)
        dbeta = torch.sum(dY * C) if beta != 0 else torch.zeros((), dtype=A.dtype, device=A.device)
        return dA, dB, dalpha, dbeta, beta * dY

    return Y, pullback


def pullback_grouped_(problems: Sequence[tuple], *, method=None, stream=None) -> None:
    """Batched backward: for each (tA, tB, alpha, A, B, dY, dA, dB) the two
    adjoint GEMMs; all 2*len(problems) GEMMs run in one grouped launch."""
    descs = []
    mth = None
    for (tA, tB, alpha, A, B, dY, dA, dB) in problems:
        mth = mth or _method(A.dtype, method)
        tA, tB = tA.upper(), tB.upper()
        if dA is not None:
            if tA == "N":
                descs.append(GemmDesc("N", "T" if tB == "N" else "N", alpha, dY, B, 0.0, dA))
            else:
                descs.append(GemmDesc("N" if tB == "N" else "T", "T", alpha, B, dY, 0.0, dA))
        if dB is not None:
            if tB == "N":
                descs.append(GemmDesc("T" if tA == "N" else "N", "N", alpha, A, dY, 0.0, dB))
            else:
                descs.append(GemmDesc("T", "N" if tA == "N" else "T", alpha, dY, A, 0.0, dB))
    if descs:
        gemm_grouped_(mth, descs, stream=stream)


class GemmFunction(torch.autograd.Function):
    """torch autograd binding: Y = A @ B through the library, backward through
    the pullback entry point (one grouped launch for dA and dB)."""

    @staticmethod
    def forward(ctx, A, B, method=None):
        ctx.save_for_backward(A, B)
        ctx.method = method
        Ac, Bc = A.T.contiguous().T, B.T.contiguous().T
        Y = device_matrix(A.shape[0], B.shape[1], A.dtype)
        gemm_("N", "N", 1.0, Ac, Bc, 0.0, Y, method=_fwd_method(A.dtype, method))
        return Y

    @staticmethod
    def backward(ctx, dY):
        A, B = ctx.saved_tensors
        Ac, Bc = A.T.contiguous().T, B.T.contiguous().T
        dYc = dY.T.contiguous().T
        dA = device_matrix(A.shape[0], A.shape[1], A.dtype) if ctx.needs_input_grad[0] else None
        dB = device_matrix(B.shape[0], B.shape[1], B.dtype) if ctx.needs_input_grad[1] else None
        pullback_("N", "N", 1.0, Ac, Bc, dYc, dA, dB, method=ctx.method)
        return dA, dB, None


def matmul(A, B, method=None):
    """Differentiable A @ B on the library's kernels."""
    return GemmFunction.apply(A, B, method)
