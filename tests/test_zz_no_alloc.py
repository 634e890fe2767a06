"""vt_modularity_noalloc, the allocation-free form of vt_modularity (distComputeModularity,
louvain.cpp:2468-2532) on device arrays: called through the C ABI on buffers the test
allocates, its result has the same bits as vt_modularity's on the same values, for the
fixed-order tree and for the left fold, on several calls in a row that share the one
device scratch array (a large input, then smaller ones, then an empty one), and the
left fold also equals the oracle's.  Bad arguments are refused on the host."""
import ctypes as C

import numpy as np
import pytest

import vite_amd
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    """Device buffers from the HIP runtime the library itself loaded."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    def get_f64(self, p):
        out = np.zeros(1, np.float64)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, 8, D2H) == 0
        return out[0]

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.fixture
def hip(ctx):  # ctx first: the library (and with it the HIP runtime) is loaded and the device current
    h = _Hip()
    yield h
    h.free()


def _inputs(seed, n):
    rng = np.random.default_rng(seed)
    cw = rng.uniform(0.0, 10.0, n)
    cdeg = rng.uniform(0.0, 50.0, n)
    cw[rng.integers(0, max(n, 1), n // 7)] = 0.0  # communities with no internal weight
    constant = 1.0 / max(cdeg.sum(), 1.0)
    return cw, cdeg, constant


def _noalloc(ctx, hip, cw, cdeg, constant, sequential):
    d_cw, d_cdeg, d_q = hip.put(cw), hip.put(cdeg), hip.put(np.full(1, np.nan))
    flags = vite_amd.VT_FLAG_SEQUENTIAL_SUMS if sequential else 0
    ctx.check(vite_amd._lib.vt_modularity_noalloc(ctx.h, len(cw), d_cw, d_cdeg, constant, flags, d_q))
    return hip.get_f64(d_q)


@pytest.mark.parametrize("sequential", [False, True])
def test_noalloc_modularity_equals_vt_modularity(ctx, hip, sequential):
    # sizes: many tree blocks (the 1024-block cap), fewer, one block, empty
    for seed, n in [(1, 3_000_000), (2, 70_000), (3, 700), (4, 0)]:
        cw, cdeg, c = _inputs(seed, n)
        want = ctx.modularity(cw, cdeg, c, sequential=sequential)
        got = _noalloc(ctx, hip, cw, cdeg, c, sequential)
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), (n, got, want)
        if sequential:
            assert got == po.modularity(cw, cdeg, c), n
        else:
            assert abs(got - po.modularity(cw, cdeg, c)) <= 1e-12 * max(1.0, abs(got)), n


def test_noalloc_modularity_refuses_bad_arguments(ctx, hip):
    d = hip.put(np.ones(4))
    f = vite_amd._lib.vt_modularity_noalloc
    assert f(ctx.h, -1, d, d, 1.0, 0, d) == vite_amd.VT_ERR_ARG
    assert f(ctx.h, 4, None, d, 1.0, 0, d) == vite_amd.VT_ERR_ARG
    assert f(ctx.h, 4, d, d, 1.0, 0, None) == vite_amd.VT_ERR_ARG
    assert "bad arguments" in vite_amd._lib.vt_last_error(ctx.h).decode()
