"""Guard bands: no kernel writes outside its output.

C (or D) is a column-major view inside a larger device buffer filled with a
canary bit pattern: guard regions before the first and after the last
element, and a gap of rows between the columns (leading dimension > rows).
After each call every canary word must be unchanged (compared bitwise), for
fringe shapes that leave partial tiles in m, n and k, on every kernel family:
the FP64/FP32 MFMA GEMM (each tile variant), the sequential-fma kernels,
kernel!/initkernel!, the split-BF16 / split-INT8 / 3xBF16 slice-pair kernels
(including the allocation-free form) and the grouped persistent kernels
(MFMA tiles up to 128x128, split-K slabs and the small-problem path).
"""
import ctypes

import numpy as np
import pytest

import jblas_amd as jb
from jblas_amd import generators as G
from jblas_amd.grouped import GemmDesc

from conftest import to_dev

pytestmark = pytest.mark.gpu

PAD = 4096  # guard elements before and after
GAP = 37    # guard rows between columns
CANARY64 = 0x7FF4DEADBEEF0001  # bit patterns (signalling NaNs), written and compared as integers
CANARY32 = 0x7FA0BEEF


class Guarded:
    """m x n column-major view with ld = m + GAP inside a canary-filled buffer."""

    def __init__(self, torch, m, n, dtype, init=None):
        self.torch, self.m, self.n, self.ld = torch, m, n, m + GAP
        canary = CANARY64 if dtype == np.float64 else CANARY32
        self.itype = np.int64 if dtype == np.float64 else np.int32
        total = 2 * PAD + self.ld * n
        self.buf = torch.from_numpy(np.full(total, canary, dtype=self.itype).view(dtype)).cuda()
        self.view = self.buf[PAD:PAD + self.ld * n].view(n, self.ld).T[:m]
        if init is not None:
            self.view.copy_(torch.from_numpy(np.asfortranarray(init)).cuda())
        self.before = self.buf.cpu().numpy().view(self.itype).copy()
        mask = np.ones(total, dtype=bool)
        for j in range(n):
            mask[PAD + j * self.ld: PAD + j * self.ld + m] = False
        self.mask = mask  # True: canary words

    def check(self, what):
        self.torch.cuda.synchronize()
        after = self.buf.cpu().numpy().view(self.itype)
        bad = np.nonzero(after[self.mask] != self.before[self.mask])[0]
        assert bad.size == 0, f"{what}: {bad.size} canary words overwritten"


SHAPES = [(130, 70, 45), (1, 1, 1), (63, 129, 200), (257, 3, 17)]


@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("method,dtype", [("native", np.float64), ("recurrence", np.float64),
                                          ("ozaki", np.float64), ("ozaki_int8", np.float64),
                                          ("native", np.float32), ("bf16x3", np.float32)])
def test_gemm_guard_bands(gpu, m, n, k, method, dtype):
    torch = gpu
    for ta in "NT":
        for tb in "NT":
            A = G.matrix(3, m if ta == "N" else k, k if ta == "N" else m, dtype=dtype)
            B = G.matrix(4, k if tb == "N" else n, n if tb == "N" else k, dtype=dtype)
            for beta in (0.0, 0.5):
                g = Guarded(torch, m, n, dtype, init=G.matrix(5, m, n, dtype=dtype))
                jb.gemm_(ta, tb, 1.5, to_dev(torch, A), to_dev(torch, B), beta, g.view, method=method)
                g.check(f"{method} {dtype.__name__} {ta}{tb} {m}x{n}x{k} beta={beta}")


@pytest.mark.parametrize("tile", ["128", "64", "64x32", "32"])
def test_mfma_tile_variants_guard_bands(gpu, tile, monkeypatch):
    torch = gpu
    monkeypatch.setenv("JB_GEMM_TILE", tile)
    for m, n, k in [(129, 65, 47), (200, 300, 33)]:
        for dtype in (np.float64, np.float32):
            g = Guarded(torch, m, n, dtype, init=G.matrix(6, m, n, dtype=dtype))
            jb.gemm_("N", "T", 1.0, to_dev(torch, G.matrix(7, m, k, dtype=dtype)),
                     to_dev(torch, G.matrix(8, n, k, dtype=dtype)), 1.0, g.view)
            g.check(f"tile {tile} {dtype.__name__} {m}x{n}x{k}")


def test_jmul_kernel_initkernel_guard_bands(gpu):
    torch = gpu
    for M, N, P in [(41, 13, 7), (130, 70, 66)]:
        A, X = G.matrix(9, M, N), G.matrix(10, N, P)
        for fn in (jb.kernel_, jb.initkernel_, jb.jmul_):
            g = Guarded(torch, M, P, np.float64, init=G.matrix(11, M, P))
            fn(g.view, to_dev(torch, A), to_dev(torch, X))
            g.check(f"{fn.__name__} {M}x{N}x{P}")


def test_noalloc_guard_bands(gpu):
    torch = gpu
    stream = torch.cuda.current_stream().cuda_stream
    m, n, k = 130, 70, 45
    A, B = to_dev(torch, G.matrix(12, m, k, kind="wide")), to_dev(torch, G.matrix(13, k, n, kind="wide"))
    g = Guarded(torch, m, n, np.float64, init=G.matrix(14, m, n))
    VP = ctypes.c_void_p
    rc = jb.lib().jb_dgemm_ozaki_noalloc_dev(b"N", b"N", m, n, k, 1.0, VP(A.data_ptr()), A.stride(1),
                                             VP(B.data_ptr()), B.stride(1), 0.5, VP(g.view.data_ptr()), g.ld, 6,
                                             VP(stream))
    assert rc == 0, jb.last_error()
    g.check("noalloc ozaki")


@pytest.mark.parametrize("method,dtype", [("f64", np.float64), ("f32", np.float32), ("f32_bf16x3", np.float32)])
def test_grouped_guard_bands(gpu, method, dtype, tmp_path):
    """Mixed shapes in one launch: small-path problems (m, n, k <= 64), MFMA
    tiles with fringes, and tiles and split-K forced through a tuning table
    (the 128-row/column FP64 and 3xBF16 tiles included)."""
    torch = gpu
    shapes = [(17, 9, 5), (64, 64, 64), (130, 70, 45), (33, 200, 700), (1, 1, 1), (96, 40, 1500)]
    guards, descs = [], []
    for i, (m, n, k) in enumerate(shapes):
        ta, tb = ("N", "T") if i % 2 else ("T", "N")
        A = to_dev(torch, G.matrix(20 + i, m if ta == "N" else k, k if ta == "N" else m, dtype=dtype))
        B = to_dev(torch, G.matrix(40 + i, k if tb == "N" else n, n if tb == "N" else k, dtype=dtype))
        g = Guarded(torch, m, n, dtype, init=G.matrix(60 + i, m, n, dtype=dtype))
        guards.append((g, A, B))
        descs.append(GemmDesc(ta, tb, 1.5, A, B, 0.5, g.view))
    jb.tune_reset()
    jb.gemm_grouped_(method, descs)
    for i, (g, _, _) in enumerate(guards):
        g.check(f"grouped {method} problem {i} {shapes[i]}")
    # the same group with tile and split-K forced on every bucket (tuning
    # table): 32x32 tiles with 4 splits; for FP64 and 3xBF16 also the 128x128
    # tile with one split and with 2, and the 128x64 and 64x128 tiles
    path = tmp_path / "tune.json"
    dev = jb.device_info()
    mid = jb.grouped.method_id(method)
    # the table's bucket key: method, floor(log2(floor(sqrt(m n)))), floor(log2 k)
    keys = [f"m{mid}|mn{int(np.sqrt(float(m) * n)).bit_length() - 1}|k{k.bit_length() - 1}" for (m, n, k) in shapes]
    forced = [(1, 4)] + ([(3, 1), (3, 2), (4, 2), (5, 1)] if method in ("f64", "f32_bf16x3") else [])
    try:
        for cfg, splitk in forced:
            entries = ",\n".join(f'    {{"key": "{dev["name"]}|{dev["arch"]}|{key}", "cfg": {cfg}, '
                                 f'"splitk": {splitk}, "ms": 1.0}}' for key in keys)
            path.write_text('{\n  "format": "jblas-tune-1",\n  "entries": [\n' + entries + "\n  ]\n}\n")
            for g, _, _ in guards:
                g.before = g.buf.cpu().numpy().view(g.itype).copy()
            jb.tune_reset()
            jb.tune_load(str(path))
            jb.gemm_grouped_(method, descs)
            for i, (g, _, _) in enumerate(guards):
                g.check(f"grouped {method} tile {cfg} split-K {splitk} problem {i} {shapes[i]}")
    finally:
        jb.tune_reset()
