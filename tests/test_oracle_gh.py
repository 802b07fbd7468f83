"""Pins the CPU oracle of the Greiner-Hormann clip before it is trusted.

Checks, each chosen so that a plausible slip in the restatement fails one:
  * the reference's committed fixtures (GH-CUDA/data + GH-CUDA/results,
    copied as data under tests/golden/fig): topology, component and vertex
    counts, coordinates to the fixtures' printed precision;
  * the reference's own CPU code (oracle/_ref, compiled from
    GH-CUDA/lib/polyclip.h behind a driver) -- bit-identical results;
  * known-answer tables for every intersection type and initial label,
    built by hand from the definitions (integer coordinates, exact in FP64);
  * an independent boundary-integral area and the textbook
    Sutherland-Hodgman clip against convex polygons;
  * closed-form special cases: P n P = P, disjoint = empty, nested = inner.
"""
import os

import numpy as np
import pytest

import gcmf_amd as G
import oracle_py as O
from conftest import GOLDEN, load_poly

FIGS = [("Fig1", "Fig1-P", "Fig1-Q", "Fig1-R"), ("Fig14", "Fig14-P", "Fig14-Q", "Fig14-R"),
        ("Fig14cw", "Fig14-P-clockWise", "Fig14-Q", "Fig14-R"),
        ("Fig15", "Fig15-P", "Fig15-Q", "Fig15-R"),
        ("Fig17", "Fig17-P-Clockwise", "Fig17-Q", "Fig17-R")]


def fig(p, q, r):
    f = lambda n: load_poly(os.path.join(GOLDEN, "fig", n + ".poly"))
    return f(p)[0], f(q)[0], f(r)


def canon(comp, digits=6):
    """A component as a rotation- and direction-invariant tuple of vertices
    rounded to the fixtures' 6 significant digits (a clockwise input ring
    traces its result the other way round)."""
    r = [tuple(float("%.*g" % (digits, v)) for v in p) for p in comp]
    best = None
    for seq in (r, r[::-1]):
        i = min(range(len(seq)), key=lambda j: seq[j])
        t = tuple(seq[i:] + seq[:i])
        best = t if best is None or t < best else best
    return best


@pytest.mark.parametrize("name,p,q,r", FIGS)
@pytest.mark.parametrize("mode", [0, 1])
def test_fixture_results(name, p, q, r, mode):
    P, Q, R = fig(p, q, r)
    for route in (O.clip_cpu, O.clip_gpu_route):
        st, comps, _ = route(P, Q, mode)
        assert st == 0
        assert sorted(len(c) for c in comps) == sorted(len(c) for c in R)
        assert sorted(canon(c) for c in comps) == sorted(canon(c) for c in R)


def test_fixture_fig1_exact_order():
    P, Q, R = fig("Fig1-P", "Fig1-Q", "Fig1-R")
    st, comps, _ = O.clip_cpu(P, Q, 1)
    np.testing.assert_allclose(comps[0], R[0], rtol=2e-6)


def rings_for_ref():
    out = [fig(p, q, r)[:2] for _, p, q, r in FIGS]
    for s in range(40):
        rs = np.random.RandomState(s)
        out.append(G.gen_single_pair(0, 2000 + s, int(rs.randint(8, 300)),
                                     int(rs.randint(8, 300))))
    out.append(G.gen_single_pair(0, 0, 2000, 1500))
    return out


def test_against_reference_cpu_code():
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent when built)")
    for P, Q in rings_for_ref():
        ref = O.ref_clip(P, Q)
        st, comps, _ = O.clip_cpu(P, Q, 0)
        assert st == 0
        assert len(ref) == len(comps)
        for a, b in zip(ref, comps):
            np.testing.assert_array_equal(a, b)


def test_routes_agree_random():
    for s in range(60):
        rs = np.random.RandomState(s)
        a, b = int(rs.randint(8, 300)), int(rs.randint(8, 300))
        P, Q = G.gen_single_pair(0, 3000 + s, max(a, b), min(a, b))
        for mode in (1, 3):
            s1, c1, _ = O.clip_cpu(P, Q, mode)
            s2, c2, _ = O.clip_gpu_route(P, Q, mode)
            assert s1 == s2 == 0
            assert len(c1) == len(c2)
            for x, y in zip(c1, c2):
                np.testing.assert_array_equal(x, y)


# --- known answers for the intersection classification (polyclip.h:48-58)
TYPE_KAT = [
    # P1, P2, Q1, Q2, expected type
    ((0, 0), (4, 0), (2, -2), (2, 2), 1),    # X: proper crossing
    ((0, 0), (4, 0), (0, -2), (0, 2), 2),    # T on Q: P1 in Q's interior
    ((0, 0), (4, 0), (2, 0), (2, 2), 3),     # T on P: Q1 in P's interior
    ((0, 0), (4, 0), (0, 0), (0, 2), 4),     # V: P1 == Q1
    ((0, 0), (4, 0), (2, 0), (-2, 0), 5),    # X-overlap
    ((0, 0), (4, 0), (6, 0), (-2, 0), 6),    # T-overlap Q: P1 in Q, Q1 outside P
    ((0, 0), (4, 0), (2, 0), (6, 0), 7),     # T-overlap P: Q1 in P, P1 outside Q
    ((0, 0), (4, 0), (0, 0), (-2, 0), 8),    # V-overlap
    ((0, 0), (4, 0), (0, 1), (4, 1), 0),     # parallel, apart
    ((0, 0), (4, 0), (5, -1), (5, 1), 0),    # lines cross outside P
    ((0, 0), (4, 0), (4, -2), (4, 2), 0),    # crossing at P2 belongs to the next edge
]


@pytest.mark.parametrize("P1,P2,Q1,Q2,t", TYPE_KAT)
def test_intersection_type_kat(P1, P2, Q1, Q2, t):
    got, a, b = O.intersect_type(P1, P2, Q1, Q2)
    assert got == t
    if t == 1:
        assert a == 0.5 and b == 0.5


def test_initial_label_table():
    # getInitialLabel (gppolyclip_mDIV512.cu:237-252) from the definitions:
    # 0 LEFT, 1 RIGHT, 2 IS_P_m, 3 IS_P_p
    want = {(0, 1): 1, (1, 0): 1, (0, 0): 2, (1, 1): 2, (1, 3): 3, (3, 1): 3, (0, 3): 4,
            (3, 0): 4, (2, 3): 5, (3, 2): 5, (2, 1): 6, (1, 2): 6, (2, 0): 7, (0, 2): 7}
    for m in range(4):
        for p in range(4):
            assert O.lib().orc_initial_label(m, p) == want.get((m, p), -102)


# --- area checks
def sutherland_hodgman(subject, clip):
    """Textbook clip of any polygon by a convex CCW polygon."""
    out = [tuple(p) for p in subject]
    n = len(clip)
    for i in range(n):
        a, b = clip[i], clip[(i + 1) % n]
        inside = lambda p: (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0]) >= 0
        def inter(p, q):
            d1 = (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0])
            d2 = (b[0] - a[0]) * (q[1] - a[1]) - (b[1] - a[1]) * (q[0] - a[0])
            t = d1 / (d1 - d2)
            return (p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1]))
        inp, out = out, []
        for j in range(len(inp)):
            p, q = inp[j - 1], inp[j]
            if inside(q):
                if not inside(p):
                    out.append(inter(p, q))
                out.append(q)
            elif inside(p):
                out.append(inter(p, q))
        if not out:
            break
    return np.array(out)


def test_area_against_sutherland_hodgman():
    rs = np.random.RandomState(7)
    for s in range(20):
        P, _ = G.gen_single_pair(0, 400 + s, int(rs.randint(20, 200)), 8)
        k = int(rs.randint(3, 12))
        th = np.sort(rs.uniform(0, 2 * np.pi, k))
        Q = np.stack([0.3 + 0.8 * np.cos(th), 0.2 + 0.8 * np.sin(th)], axis=1)  # convex, CCW
        sh = sutherland_hodgman(P, Q)
        a_sh = O.comps_area([sh]) if len(sh) else 0.0
        st, comps, _ = O.clip_cpu(P, Q, 3)
        a_gh = O.comps_area(comps)
        assert abs(a_gh - a_sh) <= 1e-12 * max(a_sh, 1e-300)
        assert abs(O.area_boundary(P, Q) - a_sh) <= 1e-12 * max(a_sh, 1e-300)


def test_traced_area_matches_boundary_integral():
    for s in range(40):
        rs = np.random.RandomState(100 + s)
        P, Q = G.gen_single_pair(0, 5000 + s, int(rs.randint(50, 600)), int(rs.randint(50, 600)))
        st, comps, _ = O.clip_cpu(P, Q, 3)
        A = O.area_boundary(P, Q)
        assert abs(O.comps_area(comps, 0.15, 0.1) - A) <= 1e-12 * A


def test_special_cases():
    P, Q = G.gen_single_pair(0, 9, 120, 90)
    st, comps, _ = O.clip_cpu(P, P.copy(), 1)
    assert st == 0 and len(comps) == 1
    assert abs(O.comps_area(comps) - O.comps_area([P])) <= 1e-14 * O.comps_area([P])
    st, comps, _ = O.clip_cpu(P, Q + 10.0, 1)
    assert st == 2 and comps == []
    inner = 0.2 * P
    st, comps, _ = O.clip_cpu(P, inner, 1)
    assert st == 0 and len(comps) == 1
    np.testing.assert_array_equal(comps[0], inner)


def test_phase1_invariants():
    P, Q = G.gen_single_pair(0, 0, 2000, 1500)
    t = O.phase1(P, Q, 1)
    assert t["invalid"] == 0
    # both views count every intersection once (count1 totals agree)
    assert t["psP1"][-1] == t["psQ1"][-1] > 0
    # slots = vertices + vertex-adding intersections on each side
    assert t["cP"] == len(P) + int(np.sum(t["alphaP"] != -100)) - int(
        np.sum(t["alphaP"][t["psP2"][:-1]] != -100))
    # neighbour links are mutual (1-based)
    nbP, nbQ = t["neighborP"], t["neighborQ"]
    for i in np.nonzero(t["alphaP"] != -100)[0]:
        assert nbQ[nbP[i] - 1] == i + 1
    # alpha keys ascend along each edge
    for e in range(len(P)):
        a = t["alphaP"][t["psP2"][e] + 1:t["psP2"][e + 1]]
        assert np.all(np.diff(a) >= 0)


def _prefilter_cases():
    """Pairs that stress the CPU-route edge-pair prefilter of the batched
    clip: general position, grid-snapped (shared vertices, collinear edges),
    near-duplicate vertices (edges down to 1e-12), shared and identical
    rings, perpendicular tiny edges far apart along one line (where the
    collinear branch of intersect() reaches far), a huge ring around a tiny
    one (large normalised coordinates), and the large-pair shape."""
    rs = np.random.RandomState(0)
    out = []
    for s in range(6):
        out.append(("rand%d" % s, G.gen_single_pair(0, s, int(rs.randint(8, 300)),
                                                   int(rs.randint(8, 300))), 1))
    for grid in (1e-2, 1e-3, 1e-4):
        out.append(("snap%g" % grid, G.gen_single_pair(1, 50, 1200, 300, snap=grid), 1))

    def dups(P, scale):
        keep = []
        for v in P:
            keep.append(v)
            if rs.rand() < 0.3:
                keep.append(v + rs.randn(2) * scale)
        return np.array(keep)

    for scale in (1e-12, 1e-9, 1e-6):
        P, Q = G.gen_single_pair(0, 7, 300, 150)
        for norm in (1, 0):
            out.append(("dups%g_%d" % (scale, norm), (dups(P, scale), dups(Q, scale)), norm))
    P, _ = G.gen_single_pair(0, 11, 240, 8)
    out.append(("shared_sub", (P, P[::3].copy()), 1))
    mid = 0.5 * (P[:-1] + P[1:])
    out.append(("shared_mid", (P, np.concatenate([P[:120], mid[120:]])), 1))
    out.append(("identical", (P, P.copy()), 1))
    P = np.array([[0, 0], [0, 1e-3], [-1, 1], [-1, -1]], float)
    Q = np.array([[1000, 5e-4], [1000 + 1e-5, 5e-4], [1000, 2], [-0.5, 2], [-0.5, -2],
                  [1000, -2]], float)
    out.append(("perp_far_raw", (P, Q), 0))
    out.append(("perp_far", (P, Q), 1))
    P, Q = G.gen_single_pair(0, 3, 300, 200)
    out.append(("huge_vs_tiny", (P * 1e4, Q * 1e-2 + 0.3), 1))
    out.append(("wavy", G.gen_single_pair(1, 0, 3000, 800), 1))
    out.append(("wavy_snap", G.gen_single_pair(1, 0, 3000, 800, snap=1e-3), 1))
    return out


@pytest.mark.parametrize("name,pq,norm", _prefilter_cases(), ids=lambda v: v if isinstance(
    v, str) else "")
def test_cpu_route_prefilter_is_sound(name, pq, norm):
    """The batched clip's CPU route skips an edge pair only when intersect()
    returns 0 for it, and drops an edge only when it has no nonzero pair:
    checked against intersect() on every pair of edges."""
    P, Q = pq
    if len(P) < len(Q):
        P, Q = Q, P
    rc, d = O.prefilter_check(P, Q, norm)
    assert rc == 0, name
    assert d["bad_skips"] == 0 and d["bad_drops"] == 0, (name, d)
    if name.startswith(("rand", "snap", "wavy", "identical")):
        assert d["skipped"] > 0.9 * d["pairs"], (name, d)  # the filter does its job


def test_prefilter_check_catches_a_wrong_rule():
    """The check itself is not vacuous: on the perpendicular far-apart edges
    the collinear branch of intersect() classifies a pair whose boxes are
    1000 apart, which a plain closed-box test (the GPU route's lsmf) skips."""
    P = np.array([[0, 0], [0, 1e-3], [-1, 1], [-1, -1]], float)
    Q = np.array([[1000, 5e-4], [1000 + 1e-5, 5e-4], [1000, 2], [-0.5, 2], [-0.5, -2],
                  [1000, -2]], float)
    t, al, be = O.intersect_type(P[0], P[1], Q[0], Q[1])
    assert t == 7 and 0 < al < 1  # Q1's foot lies on P; Q1 itself is 1000 away


def test_oracle_reference_large_fixture():
    """The reference's own large pair (GH-CUDA/data/readPolygon/s.txt, 101 242
    integer-coordinate vertices, x c.txt, 72 997) against its committed result
    GH-CUDA/results/Fig-large-R.poly: in the normalised, edge-scaled frame
    (mode 5, the clip's default) the oracle gives the 11 components with the
    sizes of GH-CUDA/ReadMe.txt:59 and every vertex equal to the file at its
    printed precision.  About 80 s: the oracle classifies all 7.4e9 edge
    pairs."""
    from conftest import large_fixture, printed
    S, Cq, R = large_fixture()
    assert [len(c) for c in R] == [5289, 5792, 1247, 1953, 1579, 4808, 4300, 3441, 6188,
                                   9739, 5976]
    st, comps, stats = O.clip_cpu(S, Cq, 5)
    assert st == 0 and stats["n_inter"] == 47
    assert [len(c) for c in comps] == [len(c) for c in R]
    for a, b in zip(comps, R):
        np.testing.assert_array_equal(printed(a), b)
