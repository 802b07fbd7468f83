"""Pins the CPU oracle of the spatial join: CoordToMBR keys (known answers
and a Decimal closed form), the endpoint sort (stable argsort), the MBR
candidate predicate and order (brute force over all pairs),
and the refine stages (numpy restatements of the cited formulas)."""
from decimal import Decimal

import numpy as np
import pytest

import gcmf_amd as G
import oracle_py as O

KEY_KAT = [("0.123456789012345", 9000000000000000 + 123456789012),
           ("12.5", 9000000000000000 + 12500000000000),
           ("3", 9000000000000000 + 3000000000000),
           ("0.000000000001999", 9000000000000000 + 1),
           ("-0.5", 9000000000000000 - 500000000000)]


@pytest.mark.parametrize("tok,key", KEY_KAT)
def test_coord_to_mbr_kat(tok, key):
    assert O.lib().orc_coord_to_mbr(tok.encode()) == key
    assert G.coord_to_mbr(tok) == key


def test_coord_to_mbr_closed_form():
    rs = np.random.RandomState(0)
    for v in rs.uniform(-2, 2, 500):
        tok = "%.15f" % v
        d = Decimal(tok)
        trunc = int(d.scaleb(12).to_integral_value(rounding="ROUND_DOWN"))
        assert G.coord_to_mbr(tok) == (9000000000000000 + trunc) % 2 ** 64


def small_layers(seed, nb=300, no=250):
    return G.gen_join_layers(nb, no, seed=seed, rmin=5e-3, rmax=8e-2, nmin=8, nmax=40)


def test_endpoint_sort_is_stable_argsort():
    b, o = small_layers(1)
    perm, rank = O.endpoint_sort(b.mbr, o.mbr)
    m = np.concatenate([b.mbr.reshape(-1, 4), o.mbr.reshape(-1, 4)])
    keys = np.empty(2 * len(m), dtype=np.uint64)
    keys[0::2] = m[:, 0] % np.uint64(10 ** 16)
    keys[1::2] = m[:, 2] % np.uint64(10 ** 16)
    np.testing.assert_array_equal(perm, np.argsort(keys, kind="stable"))
    np.testing.assert_array_equal(rank[perm], np.arange(len(perm)))


def brute_candidates(b, o):
    """All (base, overlay) pairs by the sweep predicate, in sweep order."""
    B, Ob = b.npoly, o.npoly
    bm = b.mbr.reshape(-1, 4).astype(object)
    om = o.mbr.reshape(-1, 4).astype(object)
    mod = 10 ** 16
    out = []
    for i in range(B):
        lo, hi = (bm[i][0] % mod, 2 * i), (bm[i][2] % mod, 2 * i + 1)
        cand = []
        for j in range(Ob):
            x = (om[j][0] % mod, 2 * (B + j))
            if lo < x < hi and not (bm[i][1] > om[j][3]) and not (om[j][1] > bm[i][3]):
                cand.append((x, (i, j)))
        out += [c for _, c in sorted(cand)]
    for j in range(Ob):
        lo, hi = (om[j][0] % mod, 2 * (B + j)), (om[j][2] % mod, 2 * (B + j) + 1)
        cand = []
        for i in range(B):
            x = (bm[i][0] % mod, 2 * i)
            if lo < x < hi and not (om[j][1] > bm[i][3]) and not (bm[i][1] > om[j][3]):
                cand.append((x, (i, j)))
        out += [c for _, c in sorted(cand)]
    return np.array(out, dtype=np.int32).reshape(-1, 2)


@pytest.mark.parametrize("seed", [1, 2])
def test_sweep_against_brute_force(seed):
    b, o = small_layers(seed, 120, 100)
    counts, pairs = O.mbr_sweep(b.mbr, o.mbr)
    ref = brute_candidates(b, o)
    np.testing.assert_array_equal(pairs, ref)
    # each overlapping MBR pair (closed on y, half-open on x by rank) once
    assert len({tuple(p) for p in pairs}) == len(pairs)


def test_sweep_with_key_ties():
    # identical x keys: ties are broken by endpoint index
    rings = [np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]]) for _ in range(3)]
    b = G.Layer.from_rings(rings)
    o = G.Layer.from_rings(rings[:2])
    counts, pairs = O.mbr_sweep(b.mbr, o.mbr)
    np.testing.assert_array_equal(pairs, brute_candidates(b, o))


def test_refine_stages_numpy():
    b, o = small_layers(3, 200, 200)
    counts, pairs = O.mbr_sweep(b.mbr, o.mbr)
    r = O.join_refine(pairs, b.as_dict(), o.as_dict())
    bm, om = b.mbr2.reshape(-1, 4), o.mbr2.reshape(-1, 4)
    for k, (i, j) in enumerate(pairs):
        c = r["cmbr"][k]
        assert c[0] == max(om[j][0], bm[i][0]) and c[2] == min(om[j][2], bm[i][2])
        assert c[1] == max(om[j][1], bm[i][1]) and c[3] == min(om[j][3], bm[i][3])
        t = 1 if np.array_equal(bm[i], c) else (2 if np.array_equal(om[j], c) else 0)
        assert r["pip_type"][k] == t
        if t:
            inner = b.ring(i)[0] if t == 1 else o.ring(j)[0]
            lay = o if t == 1 else b
            idx = j if t == 1 else i
            end = lay.vpsnum[idx]
            ring = lay.coords[2 * (end - lay.vnum[idx]):2 * end].reshape(-1, 2)
            px, py = inner
            c_ = 0
            n = len(ring)
            for a in range(n):
                X1, Y1 = ring[a]
                X2, Y2 = ring[a - 1]
                if (Y1 > py) == (Y2 > py):
                    continue
                if px < (X2 - X1) * (py - Y1) / (Y2 - Y1) + X1:
                    c_ ^= 1
            assert r["pip_flag"][k] == c_
    # CMF counts: edges (closing edge included) whose bbox meets the CMBR
    bw = ow = 0
    for k, (i, j) in enumerate(pairs):
        if r["join_flag_after_pip"][k] != 1:
            continue
        c = r["cmbr"][k]
        cnt = []
        for lay, idx in ((b, i), (o, j)):
            end = lay.vpsnum[idx]
            v = lay.coords[2 * (end - lay.vnum[idx]):2 * end].reshape(-1, 2)
            w = np.roll(v, -1, axis=0)
            keep = ~(((c[0] > v[:, 0]) & (c[0] > w[:, 0])) | ((c[2] < v[:, 0]) & (c[2] < w[:, 0])))
            keep &= ~(((c[1] > v[:, 1]) & (c[1] > w[:, 1])) | ((c[3] < v[:, 1]) & (c[3] < w[:, 1])))
            cnt.append(np.nonzero(keep)[0])
        if len(cnt[0]) == 0 or len(cnt[1]) == 0:
            assert r["join_flag"][k] == 0
            continue
        assert r["b_count"][k] == len(cnt[0]) and r["o_count"][k] == len(cnt[1])
        np.testing.assert_array_equal(r["b_edges"][bw:bw + len(cnt[0])], cnt[0])
        np.testing.assert_array_equal(r["o_edges"][ow:ow + len(cnt[1])], cnt[1])
        bw += len(cnt[0])
        ow += len(cnt[1])
    assert r["hits"] > 0 and r["within"] > 0
