"""Golden fixtures (tests/golden/fixtures.json, written by
tests/golden/make_fixtures.py): seeds and expected outputs of the GEMM path at
small sizes -- jmul!/gemm! at 16^3 and 64^3 computed independently of the C
oracle by an exact-rational brute force of the reference recurrence
(src/gemm.jl:86, :165), jmul! at 256^3 and the emulated DGEMM from the oracle.

CPU: the oracle reproduces every fixture (SHA-256 of the output bytes and the
sampled entries).  GPU: the HIP paths reproduce them bit for bit."""
import hashlib
import json
import os

import numpy as np
import pytest

import jblas_amd as jb
from jblas_amd import generators as G
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))["cases"]


def inputs(c):
    """The case's operands as stored (op(A) is m x k, op(B) is k x n)."""
    m, n, k = c["m"], c["n"], c["k"]
    if c["op"] == "jmul":
        return G.matrix(c["seeds"][0], m, k), G.matrix(c["seeds"][1], k, n), None
    if c["op"] == "gemm":
        A, B, C0 = G.matrix(c["seeds"][0], m, k), G.matrix(c["seeds"][1], k, n), G.matrix(c["seeds"][2], m, n)
        A = A if c["ta"] == "N" else np.asfortranarray(A.T)
        B = B if c["tb"] == "N" else np.asfortranarray(B.T)
        return A, B, C0
    return (G.matrix(c["seeds"][0], m, k, kind="wide"), G.matrix(c["seeds"][1], k, n, kind="wide"),
            G.matrix(c["seeds"][2], m, n))


def check(c, X):
    assert X.shape == (c["m"], c["n"])
    for i, j, h in c["samples"]:
        assert float(X[i, j]).hex() == h, (c["op"], c["m"], i, j)
    assert hashlib.sha256(np.asfortranarray(X, dtype=np.float64).tobytes(order="F")).hexdigest() == c["sha256"], \
        (c["op"], c["m"])


def label(c):
    return f"{c['op']}-{c['m']}-{c.get('ta', '')}{c.get('tb', '')}{c.get('nslices', '')}-{c['source'].split()[0]}"


@pytest.mark.parametrize("c", CASES, ids=label)
def test_oracle_reproduces_fixture(c):
    A, B, C0 = inputs(c)
    if c["op"] == "jmul":
        X = O.jmul(A, B)
    elif c["op"] == "gemm":
        X = O.gemm_recurrence(c["ta"], c["tb"], c["alpha"], A, B, c["beta"], C0)
    elif c["op"] == "ozaki":
        X = O.dgemm_ozaki("N", "N", c["alpha"], A, B, c["beta"], C0, c["nslices"])
    else:
        X = O.dgemm_ozaki_i8("N", "N", c["alpha"], A, B, c["beta"], C0, c["nslices"])
    check(c, X)


@pytest.mark.gpu
@pytest.mark.parametrize("c", CASES, ids=label)
def test_gpu_reproduces_fixture(gpu, c):
    A, B, C0 = inputs(c)
    if c["op"] == "jmul":
        X = np.full((c["m"], c["n"]), np.nan, order="F")
        jb.jmul_(X, A, B)
    elif c["op"] == "gemm":
        X = C0.copy(order="F")
        jb.gemm_(c["ta"], c["tb"], c["alpha"], A, B, c["beta"], X, method="native")
    else:
        X = C0.copy(order="F")
        jb.gemm_("N", "N", c["alpha"], A, B, c["beta"], X, method=c["op"], nslices=c["nslices"])
    check(c, X)
