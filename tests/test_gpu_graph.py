"""HIP-graph capture of the batched clip and of join-and-clip
(gcmf_clip_capture, gcmf_join_clip_capture, gcmf_graph_launch).

A replay runs every kernel of the batch again: it gives the eager call's
results bit for bit, replay after replay, with eager work of another context
and synchronisations between the replays (the batch includes pairs
labelled by 1024-thread blocks on the side stream, so the captured graph
holds the fork and join between the two streams).  After the layers'
coordinates are scaled by 2 in place -- the normalised predicates, and so
every size, are unchanged -- a replay returns exactly the doubled vertices,
which shows that it reads the current inputs; scaled back, the original
results again.  While a graph exists the calls that would reuse its
workspace are refused."""
import ctypes as C

import numpy as np
import pytest

import gcmf_amd as G
import oracle_py as O
from test_gpu_clip import check_results, random_cases, snapped_cases, fig_cases, NORM

pytestmark = pytest.mark.gpu


def _cases():
    cases = fig_cases() + random_cases(40, seed0=501, nmax=400) + snapped_cases(700, n=3)
    cases += [("wavy%d" % s,) + tuple(G.gen_single_pair(1, 90 + s, 6000, 1500)) for s in range(2)]
    return cases


def _results(ctx, st):
    status, res = ctx.clip_results(st)
    return status, [[c.copy() for c in r] for r in res]


def _same(a, b, scale=1.0):
    assert np.array_equal(a[0], b[0])
    for x, y in zip(a[1], b[1]):
        assert [len(c) for c in x] == [len(c) for c in y]
        for p, q in zip(x, y):
            np.testing.assert_array_equal(p, scale * q)


def test_clip_graph_replays(gctx):
    cases = _cases()
    M = len(cases)
    base = G.Layer.from_rings([c[1] for c in cases])
    over = G.Layer.from_rings([c[2] for c in cases])
    ctx = G.Context(0)
    db, do = G.DeviceLayer(ctx, base), G.DeviceLayer(ctx, over)
    pairs = np.stack([np.arange(M), np.arange(M)], axis=1).astype(np.int32)
    dp = ctx.upload(pairs)
    try:
        st0 = ctx.clip(db, do, pairs)
        ref = _results(ctx, st0)
        assert set(ctx.clip_get("label_path", M).tolist()) >= {G.LP_SEGMENTS, G.LP_SEG_CLEANUP}
        st = ctx.clip_capture(db, do, dp, M)
        assert st == st0
        _same(_results(ctx, st), ref)
        # refused while the graph holds the workspace
        with pytest.raises(G.GcmfError):
            ctx.clip(db, do, pairs)
        gb, go = G.DeviceLayer(gctx, base), G.DeviceLayer(gctx, over)
        for rep in range(3):
            # eager work elsewhere between the replays
            gctx.clip(gb, go, pairs[:50])
            gctx.synchronize()
            assert ctx.graph_launch()
            _same(_results(ctx, st), ref)
        gb.free()
        go.free()
        check_results(ctx, st, cases, NORM)
        # the replay reads the current coordinates: scale both layers by 2
        for L, d in ((base, db), (over, do)):
            x2 = np.ascontiguousarray(L.coords * 2.0)
            ctx._check(G.hip_lib().gcmf_memcpy_h2d(ctx.h, d.bufs[0].ptr, x2.ctypes.data,
                                                   x2.nbytes), "h2d")
        assert ctx.graph_launch()
        _same(_results(ctx, st), ref, scale=2.0)
        for L, d in ((base, db), (over, do)):
            ctx._check(G.hip_lib().gcmf_memcpy_h2d(ctx.h, d.bufs[0].ptr, L.coords.ctypes.data,
                                                   L.coords.nbytes), "h2d")
        # replays without waiting, then one that waits
        for _ in range(3):
            ctx.graph_launch(wait=False)
        assert ctx.graph_launch()
        _same(_results(ctx, st), ref)
        ctx.graph_destroy()
        st1 = ctx.clip(db, do, pairs)
        assert st1 == st0
        _same(_results(ctx, st1), ref)
    finally:
        dp.free()
        db.free()
        do.free()
        ctx.close()


@pytest.mark.parametrize("strip_min", [None, 0])
def test_join_clip_graph_replays(strip_min):
    """Join and clip captured together, with the plain MBR sweep and with the
    y-strip sweep (whose strip geometry comes from a sample read back)."""
    base, over = G.gen_join_layers(3000, 3000, seed=5, rmin=2e-3, rmax=2e-2, nmin=16, nmax=128)
    ctx = G.Context(0)
    if strip_min is not None:
        ctx.set_tuning(**dict(ctx.tuning(), strip_min=strip_min))
    db, do = G.DeviceLayer(ctx, base), G.DeviceLayer(ctx, over)
    try:
        js0, cs0 = ctx.join_clip(db, do)
        joined0 = ctx.join_get("joined", 2 * js0["joined"])
        ref = _results(ctx, cs0)
        js, cs = ctx.join_clip_capture(db, do)
        assert js == js0 and cs == cs0
        for _ in range(3):
            assert ctx.graph_launch()
            np.testing.assert_array_equal(ctx.join_get("joined", 2 * js0["joined"]), joined0)
            _same(_results(ctx, cs), ref)
        # the joined pairs against the oracle
        _, cand = O.mbr_sweep(base.mbr, over.mbr)
        r = O.join_refine(cand, base.as_dict(), over.as_dict())
        np.testing.assert_array_equal(joined0.reshape(-1, 2),
                                      cand[(r["pip_flag"] == 1) | (r["seg_flag"] == 1)])
        ctx.graph_destroy()
    finally:
        db.free()
        do.free()
        ctx.close()
