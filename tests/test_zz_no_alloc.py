"""gcmf_radix_sort_u64_noalloc, the allocation-free sort, through the C ABI:
bit-identical to gcmf_radix_sort_u64 (keys and values) on two calls in a
row, the second smaller (the shared static scratch keeps the first call's
contents), and sizes beyond GCMF_NOALLOC_SORT_MAX are rejected."""
import ctypes as C

import numpy as np
import pytest

import gcmf_amd as G

pytestmark = pytest.mark.gpu


def sort_both(gctx, keys, bits):
    n = len(keys)
    vals = np.arange(n, dtype=np.int32)[::-1].copy()
    dk, dv = gctx.upload(keys), gctx.upload(vals)
    out = [(G.DeviceBuffer(gctx, 8 * n), G.DeviceBuffer(gctx, 4 * n)) for _ in range(2)]
    try:
        gctx.radix_sort(dk, dv, out[0][0], out[0][1], n, bits)
        gctx.radix_sort_noalloc(dk, dv, out[1][0], out[1][1], n, bits)
        gctx.synchronize()
        res = [(gctx.download(k, np.uint64, n), gctx.download(v, np.int32, n)) for k, v in out]
    finally:
        for b in [dk, dv] + [x for pr in out for x in pr]:
            b.free()
    return res


@pytest.mark.parametrize("bits", [64, 20])
def test_noalloc_sort_matches_sort_twice(gctx, bits):
    rs = np.random.RandomState(bits)
    for n in (G.NOALLOC_SORT_MAX - 3, 100003):
        # keys of at most `bits` significant bits, as the sort's contract asks
        keys = rs.randint(0, 2 ** min(bits, 63), size=n, dtype=np.int64).astype(np.uint64)
        if bits == 64:
            keys |= np.uint64(1 << 63) * (rs.rand(n) < 0.5).astype(np.uint64)
        keys[::7] = keys[3]  # ties: the sort is stable
        (k0, v0), (k1, v1) = sort_both(gctx, keys, bits)
        np.testing.assert_array_equal(k1, k0)
        np.testing.assert_array_equal(v1, v0)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(k0, keys[order])
        np.testing.assert_array_equal(v0, (np.arange(n, dtype=np.int32)[::-1])[order])


def test_noalloc_sort_rejects_oversize(gctx):
    # the size check comes before any launch; real buffers all the same
    L = G.hip_lib()
    bufs = [G.DeviceBuffer(gctx, 64) for _ in range(4)]
    try:
        rc = L.gcmf_radix_sort_u64_noalloc(gctx.h, *[b.ptr for b in bufs],
                                           G.NOALLOC_SORT_MAX + 1, 64)
        assert rc == -3  # GCMF_ERR_ARG
        assert b"2^21" in L.gcmf_last_error(gctx.h)
    finally:
        for b in bufs:
            b.free()
