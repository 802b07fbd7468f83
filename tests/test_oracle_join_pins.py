"""Pins of the join oracle's refine stages that do not restate their formulas:

* segment join (orc_segment_join, after GPU_Join.h:376-445): every flag
  equals exact closed-segment intersection, decided by integer orientation
  tests (the textbook method, not the reference's determinant division), on
  segments of a small integer grid whose slopes are 0, +-1/4, +-1/2, +-1,
  +-2, +-4 or vertical: crossings, touching endpoints, T-junctions, parallel
  lines, collinear overlaps and vertical edges all occur;
* point in polygon (orc_pip, after GPU_Join.h:895-912): on axis-parallel
  rectangles every point of a half-integer grid is inside exactly when
  x0 <= x < x1 and y0 <= y < y1 (the half-open crossing convention, boundary
  points included); on convex rings, every point off the boundary is inside
  exactly when it is strictly on the inner side of every edge, including
  points whose ray runs through ring vertices;
* common-MBR edge filter (orc_cmf, after GPU_Join.h:560-577): an edge (the
  degenerate closing edge v[n-1] -> v[0] included) is kept exactly when its
  closed box meets the closed CMBR, on dyadic coordinates (exact), and a
  pair with no kept edge on either side is dropped.

Each pin must fail when a plausible mistake is planted in the oracle
(oracle_py.mutant): touching segments missed, the parallel or vertical case
lost, a ray through a vertex counted twice, the boundary convention flipped,
an open box test, the closing edge dropped.
"""
from fractions import Fraction as F

import numpy as np
import pytest

import gcmf_amd as G
import oracle_py as O

DIRS = [(1, 0), (0, 1), (1, 1), (1, -1), (2, 1), (1, 2), (4, 1), (1, 4), (2, -1), (1, -2),
        (4, -1), (1, -4)]


# --------------------------------------------------------- segment join ----
def orient(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def on_box(a, b, c):
    return min(a[0], b[0]) <= c[0] <= max(a[0], b[0]) and min(a[1], b[1]) <= c[1] <= max(a[1],
                                                                                          b[1])


def closed_segments_meet(p1, p2, q1, q2):
    d1, d2 = orient(q1, q2, p1), orient(q1, q2, p2)
    d3, d4 = orient(p1, p2, q1), orient(p1, p2, q2)
    if d1 * d2 < 0 and d3 * d4 < 0:
        return True
    return ((d1 == 0 and on_box(q1, q2, p1)) or (d2 == 0 and on_box(q1, q2, p2)) or
            (d3 == 0 and on_box(p1, p2, q1)) or (d4 == 0 and on_box(p1, p2, q2)))


def segment_cases(seed, n):
    rs = np.random.RandomState(seed)
    segs = []

    def seg(p, d):
        t = int(rs.randint(1, 4)) * (1 if rs.rand() < 0.5 else -1)
        return (p, (p[0] + t * d[0], p[1] + t * d[1]))

    for _ in range(n):
        p = (int(rs.randint(0, 7)), int(rs.randint(0, 7)))
        d = DIRS[rs.randint(len(DIRS))]
        b = seg(p, d)
        u = rs.rand()
        if u < 0.25:  # on b's line: collinear overlaps, touches and gaps
            s0 = int(rs.randint(-4, 5))
            o = seg((p[0] + s0 * d[0], p[1] + s0 * d[1]), d)
        elif u < 0.35:  # parallel, one grid step off b's line
            o = seg((p[0] + int(rs.randint(-1, 2)), p[1] + 1), d)
        else:
            o = seg((int(rs.randint(0, 7)), int(rs.randint(0, 7))), DIRS[rs.randint(len(DIRS))])
        segs.append([b, o])
    # hand-made: collinear overlap, collinear touching, vertical collinear,
    # vertical parallel, T-junction, crossing, disjoint parallel diagonals
    segs += [(((0, 0), (4, 2)), ((2, 1), (6, 3))), (((0, 0), (2, 1)), ((2, 1), (4, 2))),
             (((3, 0), (3, 4)), ((3, 2), (3, 6))), (((3, 0), (3, 4)), ((4, 0), (4, 4))),
             (((0, 0), (4, 0)), ((2, 0), (2, 3))), (((0, 0), (4, 4)), ((0, 4), (4, 0))),
             (((0, 0), (4, 4)), ((1, 0), (5, 4))), (((0, 0), (2, 8)), ((1, 4), (3, 12)))]
    return segs


def seg_layers(segs):
    base = G.Layer.from_rings([np.array(s[0], float) for s in segs])
    over = G.Layer.from_rings([np.array(s[1], float) for s in segs])
    return base, over


def run_segment_pin(segs):
    base, over = seg_layers(segs)
    K = len(segs)
    pairs = np.stack([np.arange(K), np.arange(K)], axis=1)
    one = np.ones(K, np.int32)
    zero = np.zeros(K, np.int16)  # edge 0 of each two-vertex ring: v0 -> v1
    flags, hits = O.segment_join(pairs, base.as_dict(), over.as_dict(), one, one, zero, zero)
    want = np.array([closed_segments_meet(*s[0], *s[1]) for s in segs], dtype=np.int32)
    return flags, hits, want


def classify(s):
    (p1, p2), (q1, q2) = s
    d = [orient(q1, q2, p1), orient(q1, q2, p2), orient(p1, p2, q1), orient(p1, p2, q2)]
    if d[0] == d[1] == 0:
        return "collinear_meet" if closed_segments_meet(p1, p2, q1, q2) else "collinear_apart"
    if closed_segments_meet(p1, p2, q1, q2):
        return "touch" if 0 in d else "cross"
    return "parallel" if orient((0, 0), (p2[0] - p1[0], p2[1] - p1[1]),
                                (q2[0] - q1[0], q2[1] - q1[1])) == 0 else "apart"


def test_segment_join_exact_pin():
    segs = segment_cases(0, 6000)
    flags, hits, want = run_segment_pin(segs)
    np.testing.assert_array_equal(flags, want)
    assert hits == int(want.sum())
    kinds = {}
    for s in segs:
        k = classify(s)
        kinds[k] = kinds.get(k, 0) + 1
    for k in ("collinear_meet", "collinear_apart", "touch", "cross", "parallel", "apart"):
        assert kinds.get(k, 0) >= 20, kinds
    vertical = sum(1 for s in segs if s[0][0][0] == s[0][1][0] and s[1][0][0] == s[1][1][0]
                   and s[0][0][0] == s[1][0][0] and closed_segments_meet(*s[0], *s[1]))
    assert vertical >= 5


@pytest.mark.parametrize("m", [O.MUT_SEG_OPEN, O.MUT_SEG_NO_PARALLEL, O.MUT_SEG_NO_VERTICAL])
def test_segment_join_pin_catches_mutant(m):
    segs = segment_cases(0, 6000)
    with O.mutant(m):
        flags, _, want = run_segment_pin(segs)
    assert (flags != want).any()


# ------------------------------------------------------- point in polygon ----
def pip_run(rings, points):
    """PiP of points[k] against rings[k] through orc_pip: the point is the
    first vertex of a small triangle in the base layer (pipType 1)."""
    tri = [np.array([p, (p[0] + 0.125, p[1]), (p[0], p[1] + 0.125)], float) for p in points]
    base = G.Layer.from_rings(tri)
    over = G.Layer.from_rings([np.asarray(r, float) for r in rings])
    K = len(points)
    pairs = np.stack([np.arange(K), np.arange(K)], axis=1)
    flags, w = O.pip_flags(pairs, np.ones(K, np.int8), base.as_dict(), over.as_dict())
    # the same through pipType 2 (the point in the overlay layer)
    flags2, _ = O.pip_flags(pairs, np.full(K, 2, np.int8), over.as_dict(), base.as_dict())
    np.testing.assert_array_equal(flags, flags2)
    return flags, w


def rect_cases():
    rings, pts, want = [], [], []
    for (x0, y0, x1, y1) in [(1, 1, 4, 3), (0, 2, 2, 5), (2, 0, 5, 5)]:
        ccw = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
        for ring in (ccw, ccw[::-1], ccw[2:] + ccw[:2]):
            for px in np.arange(-0.5, 6.0, 0.5):
                for py in np.arange(-0.5, 6.0, 0.5):
                    rings.append(ring)
                    pts.append((float(px), float(py)))
                    want.append(int(x0 <= px < x1 and y0 <= py < y1))
    return rings, pts, np.array(want, np.int32)


def convex_cases(seed, n_rings=40):
    rs = np.random.RandomState(seed)
    rings, pts, want = [], [], []
    for _ in range(n_rings):
        P = rs.randint(0, 13, size=(12, 2))
        hull = convex_hull([tuple(map(int, p)) for p in P])
        if len(hull) < 3:
            continue
        if rs.rand() < 0.5:
            hull = hull[::-1]
        sgn = 1 if area2(hull) > 0 else -1
        for px in np.arange(-0.5, 13.5, 0.5):
            for py in range(-1, 14):  # integer y: rays through ring vertices
                q = (F(px), F(py))
                o = [sgn * orient(hull[i], hull[(i + 1) % len(hull)], q) for i in range(len(hull))]
                if 0 in o:
                    continue  # on the boundary line of an edge: left to the rectangle pin
                rings.append(hull)
                pts.append((float(px), float(py)))
                want.append(int(all(v > 0 for v in o)))
    return rings, pts, np.array(want, np.int32)


def area2(r):
    return sum(r[i][0] * r[(i + 1) % len(r)][1] - r[(i + 1) % len(r)][0] * r[i][1]
               for i in range(len(r)))


def convex_hull(points):
    pts = sorted(set(points))
    if len(pts) < 3:
        return pts
    lo, hi = [], []
    for p in pts:
        while len(lo) >= 2 and orient(lo[-2], lo[-1], p) <= 0:
            lo.pop()
        lo.append(p)
    for p in reversed(pts):
        while len(hi) >= 2 and orient(hi[-2], hi[-1], p) <= 0:
            hi.pop()
        hi.append(p)
    return lo[:-1] + hi[:-1]


def test_pip_rectangle_pin():
    rings, pts, want = rect_cases()
    flags, w = pip_run(rings, pts)
    np.testing.assert_array_equal(flags, want)
    assert w == int(want.sum())


def test_pip_convex_pin():
    rings, pts, want = convex_cases(1)
    flags, _ = pip_run(rings, pts)
    np.testing.assert_array_equal(flags, want)
    # rays through ring vertices occur: points level with a vertex, inside and out
    through = sum(1 for r, p in zip(rings, pts) if any(v[1] == p[1] for v in r))
    assert through > 1000 and want.sum() > 500


@pytest.mark.parametrize("m", [O.MUT_PIP_BOTH_ENDS, O.MUT_PIP_CLOSED_Y, O.MUT_PIP_LE])
def test_pip_pins_catch_mutant(m):
    caught = False
    for rings, pts, want in (rect_cases(), convex_cases(1)):
        with O.mutant(m):
            flags, _ = pip_run(rings, pts)
        caught |= bool((flags != want).any())
    assert caught


# ------------------------------------------------------- common-MBR filter ----
def cmf_cases(seed, K=600):
    rs = np.random.RandomState(seed)
    q = 0.25  # dyadic grid: every comparison exact
    brings, orings, boxes = [], [], []
    for _ in range(K):
        for out in (brings, orings):
            n = int(rs.randint(3, 9))
            out.append(rs.randint(0, 17, size=(n, 2)) * q)
        x0, x1 = sorted(rs.randint(0, 17, size=2) * q)
        y0, y1 = sorted(rs.randint(0, 17, size=2) * q)
        boxes.append((x0, y0, x1, y1))
    return brings, orings, np.array(boxes)


def kept_edges(ring, box):
    """Edges i = v[i] -> v[i+1] of the ring with its closing vertex, the last
    one v[n] -> v[0] (a point), whose closed box meets the closed CMBR."""
    v = [tuple(map(F, p)) for p in ring] + [tuple(map(F, ring[0]))]
    x0, y0, x1, y1 = map(F, box)
    out = []
    for i in range(len(v)):
        a, b = v[i], v[(i + 1) % len(v)]
        if max(a[0], b[0]) >= x0 and min(a[0], b[0]) <= x1 and max(a[1], b[1]) >= y0 and \
                min(a[1], b[1]) <= y1:
            out.append(i)
    return out


def run_cmf_pin(seed):
    br, orr, boxes = cmf_cases(seed)
    base, over = G.Layer.from_rings(br), G.Layer.from_rings(orr)
    K = len(br)
    pairs = np.stack([np.arange(K), np.arange(K)], axis=1)
    jf, bc, oc, be, oe, E = O.cmf(pairs, boxes, base.as_dict(), over.as_dict())
    ok = True
    bw = ow = 0
    stats = {"touch": 0, "closing": 0, "dropped": 0}
    for k in range(K):
        kb, ko = kept_edges(br[k], boxes[k]), kept_edges(orr[k], boxes[k])
        if not kb or not ko:
            stats["dropped"] += 1
            ok &= jf[k] == 0 and bc[k] == 0 and oc[k] == 0
            continue
        stats["closing"] += (len(br[k]) in kb) + (len(orr[k]) in ko)
        ok &= jf[k] == 1 and bc[k] == len(kb) and oc[k] == len(ko)
        ok &= list(be[bw:bw + bc[k]]) == kb and list(oe[ow:ow + oc[k]]) == ko
        bw += bc[k]
        ow += oc[k]
        x0, y0, x1, y1 = boxes[k]
        for ring in (br[k], orr[k]):
            stats["touch"] += int(((ring[:, 0] == x0) | (ring[:, 0] == x1) | (ring[:, 1] == y0) |
                                   (ring[:, 1] == y1)).any())
    ok &= E == K - stats["dropped"]
    return bool(ok), stats


def test_cmf_dyadic_pin():
    ok, stats = run_cmf_pin(3)
    assert ok
    assert stats["touch"] > 50 and stats["closing"] > 50 and stats["dropped"] > 20, stats


@pytest.mark.parametrize("m", [O.MUT_CMF_OPEN, O.MUT_CMF_NO_CLOSING])
def test_cmf_pin_catches_mutant(m):
    with O.mutant(m):
        ok, _ = run_cmf_pin(3)
    assert not ok
