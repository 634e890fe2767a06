"""The batched loop's device scan and stable radix sort, whose element count is read from
device memory (vt_dev_scan, vt_dev_sort_pairs): checked against numpy's exclusive cumsum
and stable argsort on the same inputs.  Values carry each pair's input position, so equal
keys must come out in input order.  Counts below the grids' capacity exercise the blocks
that exit early; tile boundaries (4096 pairs), both digit widths (8 bits up to 24 key bits, 9 bits
above) and every pass count are covered."""
import ctypes as C

import numpy as np
import pytest

import vite_amd

pytestmark = pytest.mark.gpu

_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)


def _sort(ctx, keys, vals, bits, capacity):
    n = len(keys)
    ko, vo = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    ctx.check(vite_amd._lib.vt_dev_sort_pairs(ctx.h, n, capacity, bits, keys.ctypes.data_as(_u32p),
                                              vals.ctypes.data_as(_u32p), ko.ctypes.data_as(_u32p),
                                              vo.ctypes.data_as(_u32p)))
    return ko, vo


@pytest.mark.parametrize("n", [0, 1, 2, 2047, 2048, 2049, 4095, 4096, 4097, 100_003, 3_000_000])
@pytest.mark.parametrize("bits", [8, 17, 24, 25, 26, 27, 32])
def test_dev_sort_is_stable_sort(ctx, n, bits):
    rng = np.random.default_rng(n * 131 + bits)
    hi = 1 << bits
    # few distinct keys (long runs of equal keys) for half the cases, else the full range
    keys = rng.integers(0, min(hi, 37) if n % 2 else hi, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    for capacity in (n, 3 * n + 5):
        ko, vo = _sort(ctx, keys, vals, bits, capacity)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(ko, keys[order]), (n, bits, capacity)
        assert np.array_equal(vo, vals[order]), (n, bits, capacity)


def test_dev_sort_uses_only_the_low_bits(ctx):
    """Keys are ordered by bits [0, bits) only; higher bits ride along, ties stay in order."""
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1 << 32, 50_000, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(len(keys), dtype=np.uint32)
    ko, vo = _sort(ctx, keys, vals, 12, len(keys))
    order = np.argsort(keys & np.uint32(0xFFF), kind="stable")
    assert np.array_equal(ko, keys[order]) and np.array_equal(vo, vals[order])


@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 1_000_001, 5_000_000])
def test_dev_scan_is_exclusive_cumsum(ctx, n):
    rng = np.random.default_rng(n + 7)
    a = (rng.random(n) < 0.3).astype(np.int32) if n % 2 else rng.integers(0, 5, n, dtype=np.int32)
    for capacity in (n, 2 * n + 3):
        out = np.full(n, -1, np.int32)
        total = C.c_int64(-1)
        ctx.check(vite_amd._lib.vt_dev_scan(ctx.h, n, capacity, a.ctypes.data_as(_i32p), out.ctypes.data_as(_i32p),
                                            C.byref(total)))
        want = np.concatenate([[0], np.cumsum(a, dtype=np.int64)[:-1]]) if n else np.zeros(0, np.int64)
        assert np.array_equal(out.astype(np.int64), want), (n, capacity)
        assert total.value == int(a.sum(dtype=np.int64)), (n, capacity)


def test_dev_primitives_refuse_bad_arguments(ctx):
    k = np.zeros(4, np.uint32)
    f = vite_amd._lib.vt_dev_sort_pairs
    assert f(ctx.h, 4, 3, 8, k.ctypes.data_as(_u32p), k.ctypes.data_as(_u32p), k.ctypes.data_as(_u32p),
             k.ctypes.data_as(_u32p)) == vite_amd.VT_ERR_ARG          # capacity below n
    assert f(ctx.h, 4, 4, 33, k.ctypes.data_as(_u32p), k.ctypes.data_as(_u32p), k.ctypes.data_as(_u32p),
             k.ctypes.data_as(_u32p)) == vite_amd.VT_ERR_ARG          # bits out of range
    a = np.zeros(4, np.int32)
    assert vite_amd._lib.vt_dev_scan(ctx.h, -1, 4, a.ctypes.data_as(_i32p), a.ctypes.data_as(_i32p),
                                     None) == vite_amd.VT_ERR_ARG
