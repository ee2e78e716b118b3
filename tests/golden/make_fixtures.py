"""Writes tests/golden/fixtures.json: input seeds and expected outputs for the
GEMM path at small sizes.

The reference ships no golden vectors (its test/runtests.jl only benchmarks),
so these are made here.  Inputs come from the SplitMix64 generator
(jblas_amd.generators, the same stream the device generator reproduces).

* jmul! / gemm! at 16^3 and 64^3: computed in pure Python, independently of
  the C oracle, by the reference recurrence (src/gemm.jl:86 plain first
  product, :165 fma k-steps) on exact rationals -- each product and each fma
  rounded once to the nearest double (Fraction -> float is correctly rounded).
* jmul! at 256^3 and the split-BF16 / split-INT8 emulated DGEMM at 48x40x70:
  from the oracle restatement (regression pins; the oracle is itself pinned
  by tests/test_oracle.py).

Outputs are stored as a SHA-256 of the column-major float64 bytes plus sampled
entries as float.hex strings.  Run from the repository root:
    python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import sys
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "jblas.jl_amd")]
from jblas_amd import generators as G  # noqa: E402


def rnd(q: Fraction) -> float:
    return float(q)  # correctly rounded to nearest even


def brute_recurrence(ta, tb, alpha, A, B, beta, C):
    """C = alpha*op(A)*op(B) + beta*C with the reference's per-element chain."""
    Ao = A if ta == "N" else A.T
    Bo = B if tb == "N" else B.T
    m, k = Ao.shape
    n = Bo.shape[1]
    Af = [[Fraction(float(Ao[i, p])) for p in range(k)] for i in range(m)]
    Bf = [[Fraction(float(Bo[p, j])) for j in range(n)] for p in range(k)]
    out = np.empty((m, n), order="F")
    for j in range(n):
        for i in range(m):
            acc = rnd(Af[i][0] * Bf[0][j])                      # plain first product
            for p in range(1, k):
                acc = rnd(Af[i][p] * Bf[p][j] + Fraction(acc))  # fma: one rounding
            c = rnd(Fraction(alpha) * Fraction(acc))
            if beta != 0.0:
                c = rnd(Fraction(beta) * Fraction(float(C[i, j])) + Fraction(c))
            out[i, j] = c
    return out


def digest(X) -> str:
    return hashlib.sha256(np.asfortranarray(X, dtype=np.float64).tobytes(order="F")).hexdigest()


def samples(X, seed, count=64):
    rng = np.random.default_rng(seed)
    ii = rng.integers(0, X.shape[0], count)
    jj = rng.integers(0, X.shape[1], count)
    return [[int(i), int(j), float(X[i, j]).hex()] for i, j in zip(ii, jj)]


def main():
    from oracle import oracle as O
    cases = []
    for n in (16, 64):
        A, B = G.matrix(42, n, n), G.matrix(43, n, n)
        D = brute_recurrence("N", "N", 1.0, A, B, 0.0, None)
        cases.append({"op": "jmul", "m": n, "n": n, "k": n, "seeds": [42, 43], "kind": "uniform",
                      "source": "brute force (exact rationals)", "sha256": digest(D), "samples": samples(D, n)})
        C0 = G.matrix(44, n, n)
        for ta, tb in [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")]:
            Ag = A if ta == "N" else np.asfortranarray(A.T)
            Bg = B if tb == "N" else np.asfortranarray(B.T)
            R = brute_recurrence(ta, tb, 1.5, Ag, Bg, 0.5, C0)
            cases.append({"op": "gemm", "ta": ta, "tb": tb, "m": n, "n": n, "k": n, "alpha": 1.5, "beta": 0.5,
                          "seeds": [42, 43, 44], "kind": "uniform", "stored_as_op": True,
                          "source": "brute force (exact rationals)", "sha256": digest(R), "samples": samples(R, n + 1)})
    A, B = G.matrix(42, 256, 256), G.matrix(43, 256, 256)
    D = O.jmul(A, B)
    cases.append({"op": "jmul", "m": 256, "n": 256, "k": 256, "seeds": [42, 43], "kind": "uniform",
                  "source": "oracle", "sha256": digest(D), "samples": samples(D, 256)})
    m, n, k = 48, 40, 70
    A, B, C0 = G.matrix(5, m, k, kind="wide"), G.matrix(6, k, n, kind="wide"), G.matrix(7, m, n)
    for method, fn, counts in [("ozaki", O.dgemm_ozaki, (3, 6, 7)), ("ozaki_int8", O.dgemm_ozaki_i8, (3, 7))]:
        for s in counts:
            R = fn("N", "N", 1.5, A, B, 0.5, C0, s)
            cases.append({"op": method, "nslices": s, "m": m, "n": n, "k": k, "alpha": 1.5, "beta": 0.5,
                          "seeds": [5, 6, 7], "kind": "wide", "source": "oracle", "sha256": digest(R),
                          "samples": samples(R, s)})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")
    with open(path, "w") as f:
        json.dump({"format": "jblas-golden-1", "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {path}")


if __name__ == "__main__":
    main()
