"""GPU parity of the spatial join (join.hip, prim.hip) against the oracle.

Every integer output is compared bit-exactly: the endpoint sort permutation
and ranks (closed form: stable sort by key), per-polygon candidate counts,
the ordered candidate pair list, CMBR values and PiP types/flags, CMF counts
and edge lists, segment-intersection flags and the joined pair list.
"""
import numpy as np
import pytest

import gcmf_amd as G
import oracle_py as O

pytestmark = pytest.mark.gpu


def run_join(gctx, base, over):
    db, do = G.DeviceLayer(gctx, base), G.DeviceLayer(gctx, over)
    st = gctx.join(db, do)
    out = {"st": st}
    K, N = st["candidates"], base.npoly + over.npoly
    for key, n in [("pairs", 2 * K), ("counts", N), ("perm", 2 * N), ("rank", 2 * N),
                   ("cmbr", 4 * K), ("pip_type", K), ("pip_flag", K), ("join_flag", K),
                   ("b_count", K), ("o_count", K), ("b_edges", st["b_edges"]),
                   ("o_edges", st["o_edges"]), ("seg_flag", K), ("joined", 2 * st["joined"])]:
        out[key] = gctx.join_get(key, n)
    db.free()
    do.free()
    return out


def check_join(gctx, base, over):
    g = run_join(gctx, base, over)
    st = g["st"]
    perm, rank = O.endpoint_sort(base.mbr, over.mbr)
    np.testing.assert_array_equal(g["perm"], perm)
    np.testing.assert_array_equal(g["rank"], rank)
    # closed form: stable argsort of the (mod 1e16) keys
    keys = np.empty(2 * (base.npoly + over.npoly), dtype=np.uint64)
    m = np.concatenate([base.mbr.reshape(-1, 4), over.mbr.reshape(-1, 4)])
    keys[0::2] = m[:, 0] % np.uint64(10 ** 16)
    keys[1::2] = m[:, 2] % np.uint64(10 ** 16)
    np.testing.assert_array_equal(g["perm"], np.argsort(keys, kind="stable"))
    counts, pairs = O.mbr_sweep(base.mbr, over.mbr)
    np.testing.assert_array_equal(g["counts"], counts)
    assert st["candidates"] == len(pairs)
    np.testing.assert_array_equal(g["pairs"].reshape(-1, 2), pairs)
    r = O.join_refine(pairs, base.as_dict(), over.as_dict())
    np.testing.assert_array_equal(g["cmbr"].reshape(-1, 4), r["cmbr"])
    np.testing.assert_array_equal(g["pip_type"], r["pip_type"])
    np.testing.assert_array_equal(g["pip_flag"], r["pip_flag"])
    np.testing.assert_array_equal(g["join_flag"], r["join_flag"])
    np.testing.assert_array_equal(g["b_count"], r["b_count"])
    np.testing.assert_array_equal(g["o_count"], r["o_count"])
    np.testing.assert_array_equal(g["b_edges"], r["b_edges"])
    np.testing.assert_array_equal(g["o_edges"], r["o_edges"])
    np.testing.assert_array_equal(g["seg_flag"], r["seg_flag"])
    assert st["pip_candidates"] == int(r["pip_candidate"].sum())
    assert st["within"] == r["within"]
    assert st["survivors"] == r["E"]
    assert st["intersecting"] == r["hits"]
    joined = pairs[(r["pip_flag"] == 1) | (r["seg_flag"] == 1)]
    np.testing.assert_array_equal(g["joined"].reshape(-1, 2), joined)
    return st


def test_join_small_dense(gctx):
    base, over = G.gen_join_layers(1500, 1200, seed=3, rmin=5e-3, rmax=5e-2, nmin=8, nmax=64)
    st = check_join(gctx, base, over)
    assert st["candidates"] > 1000 and st["within"] > 0 and st["intersecting"] > 0


def test_join_c2_full_size(gctx):
    # 20,000 x 20,000 polygons, radii log-uniform [1e-4, 5e-3], 16..512 vertices
    base, over = G.gen_join_layers(20000, 20000, seed=0)
    st = check_join(gctx, base, over)
    assert st["candidates"] > 10000


def test_join_many_keys(gctx):
    # 2 x 150k small polygons: 600k endpoint keys, multi-tile sort and scans;
    # above strip_min, so the MBR sweep runs over y strips
    base, over = G.gen_join_layers(150000, 150000, seed=11, rmin=1e-4, rmax=1e-3, nmin=8,
                                   nmax=16)
    assert base.npoly + over.npoly >= gctx.tuning()["strip_min"]
    check_join(gctx, base, over)


@pytest.mark.parametrize("group", [16, 64])
def test_join_cmf_group_widths(gctx, group):
    """The CMF edge filter with 16 lanes per candidate pair (four pairs per
    wave, edge lists ranked by masked ballots) and with a wave per pair, on
    small and on larger polygons: the same outputs as the oracle."""
    default = gctx.tuning()
    try:
        gctx.set_tuning(**dict(default, cmf_group=group))
        for base, over in (G.gen_join_layers(1500, 1200, seed=3, rmin=5e-3, rmax=5e-2, nmin=8,
                                             nmax=64),
                           G.gen_join_layers(3000, 3000, seed=5, rmin=2e-3, rmax=2e-2, nmin=16,
                                             nmax=300)):
            check_join(gctx, base, over)
    finally:
        gctx.set_tuning(**default)


def test_join_c2_strip_sweep(gctx):
    """The join workload's layers with the y-strip MBR sweep forced
    (strip_min 0): every stage's output still equals the oracle's."""
    base, over = G.gen_join_layers(20000, 20000, seed=0)
    default = gctx.tuning()
    try:
        gctx.set_tuning(**dict(default, strip_min=0))
        st = check_join(gctx, base, over)
    finally:
        gctx.set_tuning(**default)
    assert st["candidates"] > 10000


def test_join_then_clip_pipeline(gctx):
    base, over = G.gen_join_layers(3000, 3000, seed=5, rmin=2e-3, rmax=2e-2, nmin=16, nmax=128)
    db, do = G.DeviceLayer(gctx, base), G.DeviceLayer(gctx, over)
    js, cs = gctx.join_clip(db, do, G.CLIP_NORMALIZE)
    status, res = gctx.clip_results(cs)
    joined = gctx.join_get("joined", 2 * js["joined"]).reshape(-1, 2)
    db.free()
    do.free()
    assert cs["pairs"] == js["joined"] and cs["failed"] == 0
    for k in range(0, len(joined), max(1, len(joined) // 200)):
        b, o = joined[k]
        P, Q = base.ring(b), over.ring(o)
        if len(P) < len(Q):
            P, Q = Q, P
        s_ref, comps, _ = O.clip_cpu(P, Q, 5)
        if s_ref == 2:
            assert status[k] == 2
            continue
        assert status[k] == 0
        assert [len(c) for c in res[k]] == [len(c) for c in comps]
        for a, c in zip(res[k], comps):
            np.testing.assert_array_equal(a, c)


@pytest.mark.parametrize("n,bits", [(1, 8), (5000, 13), (300000, 32), (1 << 20, 64)])
def test_radix_sort_entry_point(gctx, n, bits):
    """gcmf_radix_sort_u64 equals numpy's stable argsort on the low `bits`
    bits (the closed form of a stable LSD radix sort)."""
    rs = np.random.RandomState(n)
    keys = rs.randint(0, 1 << 62, size=n, dtype=np.int64).astype(np.uint64)
    keys[: n // 3] = keys[n // 3: 2 * (n // 3)]  # ties
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    keys &= mask
    vals = np.arange(n, dtype=np.int32)
    bufs = [gctx.upload(keys), gctx.upload(vals), gctx.upload(np.zeros(n, np.uint64)),
            gctx.upload(np.zeros(n, np.int32))]
    gctx.radix_sort(bufs[0], bufs[1], bufs[2], bufs[3], n, bits)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(gctx.download(bufs[3], np.int32, n), order)
    np.testing.assert_array_equal(gctx.download(bufs[2], np.uint64, n), keys[order])
    for b in bufs:
        b.free()


class _At:
    """A device address `off` bytes into a buffer, for entry points that
    take caller-owned pointers."""

    def __init__(self, buf, off):
        import ctypes as C
        self.ptr = C.c_void_p(buf.ptr.value + off)


def test_radix_sort_unaligned_buffers(gctx):
    """Keys and outputs one element past a 16-byte boundary: the histogram
    reads such keys one by one instead of in 16-byte pairs; the sort still
    equals the stable argsort."""
    n, bits = 3 * 4096 + 5, 64
    rs = np.random.RandomState(9)
    keys = rs.randint(0, 1 << 62, size=n, dtype=np.int64).astype(np.uint64)
    keys[: n // 4] = keys[n // 4: 2 * (n // 4)]  # ties
    vals = np.arange(n, dtype=np.int32)
    pad_k = np.concatenate([np.zeros(1, np.uint64), keys])
    pad_v = np.concatenate([np.zeros(1, np.int32), vals])
    bk, bv = gctx.upload(pad_k), gctx.upload(pad_v)
    ok, ov = gctx.upload(np.zeros(n + 1, np.uint64)), gctx.upload(np.zeros(n + 1, np.int32))
    assert bk.ptr.value % 16 == 0 and ok.ptr.value % 16 == 0
    gctx.radix_sort(_At(bk, 8), _At(bv, 4), _At(ok, 8), _At(ov, 4), n, bits)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(gctx.download(ov, np.int32, n + 1)[1:], order)
    np.testing.assert_array_equal(gctx.download(ok, np.uint64, n + 1)[1:], keys[order])
    for b in (bk, bv, ok, ov):
        b.free()


# sizes around the single-pass scan's 4096-element tiles and its 64-tile
# look-back windows
@pytest.mark.parametrize("n", [0, 1, 4096, 4097, 64 * 4096 + 5, 3_000_000, 20_000_003])
def test_scan_and_compact_entry_points(gctx, n):
    """gcmf_scan_i32_i64 is the exclusive cumulative sum with the total last;
    gcmf_compact_flags lists the nonzero flags in order."""
    rs = np.random.RandomState(7)
    a = rs.randint(-3, 1000, size=n).astype(np.int32)
    d_in, d_out = gctx.upload(a), gctx.upload(np.zeros(n + 1, np.int64))
    gctx.scan(d_in, d_out, n)
    ref = np.concatenate([[0], np.cumsum(a.astype(np.int64))])
    np.testing.assert_array_equal(gctx.download(d_out, np.int64, n + 1), ref)
    flags = (rs.rand(n) < 0.3).astype(np.int32) * rs.randint(1, 5, size=n).astype(np.int32)
    d_f, d_idx = gctx.upload(flags), gctx.upload(np.zeros(max(n, 1), np.int32))
    cnt = gctx.compact(d_f, d_idx, n)
    want = np.nonzero(flags)[0]
    assert cnt == len(want)
    np.testing.assert_array_equal(gctx.download(d_idx, np.int32, cnt), want)
    for b in (d_in, d_out, d_f, d_idx):
        b.free()


@pytest.mark.parametrize("n", [(1 << 20) - 3, (1 << 21) + 4099])
def test_compact_dense_and_empty_tiles(gctx, n):
    """Compaction over tiles of 4096 flags that are all set (every group of
    1024 full), all clear, set only at a tile's first and last flag, or
    random, with negative flag values counting as set and a partial last
    tile that is all set; single-pass (up to 2^20 flags) and two-pass forms."""
    rs = np.random.RandomState(3)
    tiles = (n + 4095) // 4096
    flags = np.zeros(n, np.int32)
    for t in range(tiles):
        lo, hi = 4096 * t, min(4096 * (t + 1), n)
        kind = t % 4 if t < tiles - 1 else 0
        if kind == 0:
            flags[lo:hi] = rs.choice(np.array([-7, 1, 2**31 - 1], np.int32), size=hi - lo)
        elif kind == 2:
            flags[lo] = flags[hi - 1] = -1
        elif kind == 3:
            flags[lo:hi] = (rs.rand(hi - lo) < 0.5).astype(np.int32)
    d_f, d_idx = gctx.upload(flags), gctx.upload(np.zeros(n, np.int32))
    want = np.nonzero(flags)[0]
    for _ in range(2):  # the second call reuses the workspace
        cnt = gctx.compact(d_f, d_idx, n)
        assert cnt == len(want)
        np.testing.assert_array_equal(gctx.download(d_idx, np.int32, cnt), want)
    d_f.free()
    d_idx.free()


def test_scan_extreme_values(gctx):
    """int32 extremes of both signs over 2^22 + 7 elements: the running sum
    leaves the int32 range many times and every tile's carry is 64-bit."""
    n = (1 << 22) + 7
    rs = np.random.RandomState(11)
    a = np.where(rs.rand(n) < 0.7, np.int32(2**31 - 1), np.int32(-2**31)).astype(np.int32)
    d_in, d_out = gctx.upload(a), gctx.upload(np.zeros(n + 1, np.int64))
    for _ in range(2):  # the second call reuses the tile descriptors
        gctx.scan(d_in, d_out, n)
        ref = np.concatenate([[0], np.cumsum(a.astype(np.int64))])
        np.testing.assert_array_equal(gctx.download(d_out, np.int64, n + 1), ref)
    d_in.free()
    d_out.free()
