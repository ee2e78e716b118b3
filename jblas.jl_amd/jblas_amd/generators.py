"""Seeded host-side input generators (numpy), bit-identical to the device fill
kernels of libjblas_hip.so for the uniform and wide-range streams.

Replaces the reference's unseeded `mrandn`/`rand!` (src/randmat.jl:5-14).
Stream: SplitMix64, output i = mix(seed + (i+1)*0x9e3779b97f4a7c15).
Matrices are filled in column-major order from one stream per matrix.
"""
from __future__ import annotations

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
SEED_BASE = 42  # A = base, B = base+1, C = base+2, W = base+3, shapes = base+10


def splitmix64(seed: int, idx) -> np.ndarray:
    """Outputs `idx` (uint64 array) of the SplitMix64 sequence started at `seed`."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (idx + np.uint64(1)) * GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def _unit(x: np.ndarray) -> np.ndarray:
    return (x >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def uniform(seed: int, n: int) -> np.ndarray:
    """u_i = 2*((x_i >> 11) * 2^-53) - 1 in [-1, 1)."""
    return 2.0 * _unit(splitmix64(seed, np.arange(n, dtype=np.uint64))) - 1.0


def wide(seed: int, n: int, emin: int = -12, emax: int = 12) -> np.ndarray:
    """a_i = u_i * 2^e_i: u from stream index 2i, e = emin + floor((emax-emin+1) v) from 2i+1."""
    i = np.arange(n, dtype=np.uint64)
    u = 2.0 * _unit(splitmix64(seed, 2 * i)) - 1.0
    v = _unit(splitmix64(seed, 2 * i + np.uint64(1)))
    e = emin + np.floor((emax - emin + 1) * v).astype(np.int64)
    return np.ldexp(u, e)


def normal(seed: int, n: int) -> np.ndarray:
    """Box-Muller normals: z = sqrt(-2 ln u1) cos(2 pi u2), u1 in (0,1] from 2i, u2 from 2i+1."""
    i = np.arange(n, dtype=np.uint64)
    u1 = ((splitmix64(seed, 2 * i) >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = _unit(splitmix64(seed, 2 * i + np.uint64(1)))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def matrix(seed: int, rows: int, cols: int, kind: str = "uniform", dtype=np.float64, **kw) -> np.ndarray:
    """A rows x cols column-major matrix from one stream."""
    n = rows * cols
    if kind == "uniform":
        v = uniform(seed, n)
    elif kind == "wide":
        v = wide(seed, n, **kw)
    elif kind == "normal":
        v = normal(seed, n)
    else:
        raise ValueError(kind)
    return np.asfortranarray(v.reshape(cols, rows).T.astype(dtype))


def loguniform_shapes(seed: int, count: int, lo: int = 16, hi: int = 2048, mult: int = 8) -> np.ndarray:
    """count x 3 array of (m, n, k): clamp(mult*round(exp(U(ln lo, ln hi))/mult), lo, hi)."""
    u = _unit(splitmix64(seed, np.arange(3 * count, dtype=np.uint64)))
    x = np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * u)
    v = np.clip(mult * np.round(x / mult), lo, hi).astype(np.int64)
    return v.reshape(count, 3)
