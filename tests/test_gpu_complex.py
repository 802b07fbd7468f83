"""GPU parity of the complex-polygon clip (gcmf_clip_complex, gh_complex.inl):
several rings per side (outer rings, holes, islands) and the union, against
the oracle's orc_clip_multi (itself bit-identical to the reference's CPU
code, tests/test_oracle_multi.py), bit for bit: status, component sizes and
vertices.  The reference's Fig19 pair reproduces GH-CUDA/results/Fig19-R.poly
at its printed precision in the raw frame."""
import numpy as np
import pytest

import gcmf_amd as G
import oracle_py as O
from conftest import REFDATA, load_poly_text, printed
from test_oracle_multi import random_complex_pairs

pytestmark = pytest.mark.gpu


def fig(name):
    return [load_poly_text("%s/%s-%s.poly.gz" % (REFDATA, name, s)) for s in "PQR"]


def run_complex(gctx, cases, flags):
    """cases: list of (P rings, Q rings); one layer per side, each pair a run
    of consecutive rings."""
    base = G.Layer.from_rings([r for P, _ in cases for r in P])
    over = G.Layer.from_rings([r for _, Q in cases for r in Q])
    parts, b, o = [], 0, 0
    for P, Q in cases:
        parts.append((b, len(P), o, len(Q)))
        b += len(P)
        o += len(Q)
    db, do = G.DeviceLayer(gctx, base), G.DeviceLayer(gctx, over)
    try:
        st = gctx.clip_complex(db, do, np.array(parts, np.int32), flags)
        return st, gctx.clip_results(st)
    finally:
        db.free()
        do.free()


def check(cases, status, res, mode):
    for k, (P, Q) in enumerate(cases):
        s_ref, comps, _ = O.clip_multi(P, Q, mode)
        assert status[k] == s_ref, (k, status[k], s_ref)
        assert [len(c) for c in res[k]] == [len(c) for c in comps], k
        for a, b in zip(res[k], comps):
            np.testing.assert_array_equal(a, b, err_msg=str(k))


@pytest.mark.parametrize("union", [False, True])
@pytest.mark.parametrize("frame", [0, 1])
def test_complex_matches_oracle(gctx, union, frame):
    cases = [tuple(fig("Fig18")[:2]), tuple(fig("Fig19")[:2])] + random_complex_pairs(30, 7)
    flags = (G.CLIP_NORMALIZE if frame else 0) | (G.CLIP_UNION if union else 0)
    mode = (5 if frame else 0) | (8 if union else 0)
    st, (status, res) = run_complex(gctx, cases, flags)
    assert st["pairs"] == len(cases) and st["failed"] == 0
    check(cases, status, res, mode)


def test_complex_fig19_fixture(gctx):
    P, Q, R = fig("Fig19")
    st, (status, res) = run_complex(gctx, [(P, Q)], 0)  # raw frame, the reference's
    assert status[0] == 0 and [len(c) for c in res[0]] == [1240]
    np.testing.assert_array_equal(printed(res[0][0]), R[0])


def test_complex_single_rings_equal_clip(gctx):
    """One ring a side: the complex path gives the batched clip's results
    (P kept as the base ring: pairs with |base| >= |overlay|)."""
    cases = []
    for s in range(20):
        P, Q = G.gen_single_pair(0, 300 + s, 150 + 5 * s, 90)
        cases.append(([P], [Q]))
    st, (status, res) = run_complex(gctx, cases, G.CLIP_NORMALIZE)
    check(cases, status, res, 5)


def test_complex_union_special_cases(gctx):
    P, Q = G.gen_single_pair(0, 5, 100, 80)
    cases = [([P], [Q + np.array([10.0, 10.0])]),  # disjoint MBRs: both rings
             ([P], [0.2 * P]),                       # nested: the outer ring
             ([P], [P.copy()])]                      # identical: P
    st, (status, res) = run_complex(gctx, cases, G.CLIP_NORMALIZE | G.CLIP_UNION)
    assert list(status) == [0, 0, 0]
    assert len(res[0]) == 2 and len(res[1]) == 1 and len(res[2]) == 1
    check(cases, status, res, 5 | 8)
    # the intersection of the disjoint pair is empty (status 2)
    st, (status, res) = run_complex(gctx, cases[:1], G.CLIP_NORMALIZE)
    assert status[0] == 2 and res[0] == []


def test_complex_rejects_bad_parts(gctx):
    P, Q = G.gen_single_pair(0, 5, 40, 30)
    base, over = G.Layer.from_rings([P]), G.Layer.from_rings([Q])
    db, do = G.DeviceLayer(gctx, base), G.DeviceLayer(gctx, over)
    try:
        with pytest.raises(G.GcmfError):
            gctx.clip_complex(db, do, np.array([[0, 2, 0, 1]], np.int32))  # 2 rings of 1
        st = gctx.clip_complex(db, do, np.array([[0, 1, 0, 1]], np.int32))
        assert st["ok"] == 1
    finally:
        db.free()
        do.free()
