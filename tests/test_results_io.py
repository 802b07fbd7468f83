"""The layer file format as the reference reads it, and the result writers
(host library, no GPU).

Layer files (ReadTextFormatPolygon2WithVector, SEQ_Overlay.h:285-389, in the
format dataConverter.py:94-104 writes): the first line is a header and is
skipped; every other line is a polygon id, the MBR corners (min x, min y,
max x, max y) and the ring's vertices, the first vertex repeated at the end
when the source ring was not closed.  The corners are read as the first two
(x, y) pairs and dropped, so VNum counts the ring with its closing vertex,
and the MBR keys come from the text of the first token that attains each
minimum or maximum, which is the corner's text.  A line whose MBR has no
extent in x or y is skipped ("MBR Error!").  No layer file in that format
ships with the reference, so the files here are written by the rules above,
and the expected values are computed from the same text independently.

Results: .poly text at %.17g (savePolygon, polyclip.h:859-865), one WKT
MULTIPOLYGON line per pair with even-odd nesting for complex results, and
the signed ring area (which the reference does not have), pinned by closed
forms and by numpy's shoelace sum."""
import re

import numpy as np

import gcmf_amd as G
import oracle_py as O
from test_oracle_multi import random_complex_pairs


def _fmt_vertex(v):
    return "%.15f" % v


def _write_reference_layer(path, rings, corner_fmt=repr, closed=None):
    """dataConverter.py's line format; returns the expected layer fields."""
    exp = []
    with open(path, "w") as f:
        f.write("id minx miny maxx maxy coordinates\n")  # the skipped header line
        for pid, ring in enumerate(rings):
            toks = [(_fmt_vertex(x), _fmt_vertex(y)) for x, y in ring]
            if closed is None or not closed[pid]:
                toks.append(toks[0])  # dataConverter closes an open ring
            vx = [float(a) for a, _ in toks]
            vy = [float(b) for _, b in toks]
            cmin = (corner_fmt(min(vx)), corner_fmt(min(vy)))
            cmax = (corner_fmt(max(vx)), corner_fmt(max(vy)))
            line = [str(pid), cmin[0], cmin[1], cmax[0], cmax[1]] + [t for p in toks for t in p]
            f.write(" ".join(line) + " \n")
            degenerate = min(vx) == max(vx) or min(vy) == max(vy)
            if not degenerate:
                exp.append({"xy": np.array([vx, vy]).T.reshape(-1),
                            "keys": [int(O.lib().orc_coord_to_mbr(t.encode()))
                                     for t in (cmin[0], cmin[1], cmax[0], cmax[1])],
                            "mbr2": [float(t) for t in (cmin[0], cmin[1], cmax[0], cmax[1])]})
    return exp


def test_reader_reference_format(tmp_path):
    rs = np.random.RandomState(4)
    rings = []
    for k in range(40):
        n = int(rs.randint(3, 30))
        c = rs.uniform(-50, 50, 2)
        t = np.sort(rs.uniform(0, 2 * np.pi, n))
        rings.append(np.stack([c[0] + rs.uniform(0.1, 3) * np.cos(t),
                               c[1] + rs.uniform(0.1, 3) * np.sin(t)], axis=1))
    rings[7] = np.array([[1.5, 0.0], [1.5, 1.0], [1.5, 2.0]])  # no extent in x: skipped
    path = str(tmp_path / "layer.txt")
    # corners printed shorter than the vertices (repr): the keys are the corners' text
    exp = _write_reference_layer(path, rings)
    L = G.Layer.read_text(path)
    assert L.npoly == len(exp) == len(rings) - 1
    np.testing.assert_array_equal(L.vnum, [len(e["xy"]) // 2 for e in exp])
    np.testing.assert_array_equal(L.vpsnum, np.cumsum(L.vnum))
    np.testing.assert_array_equal(L.coords, np.concatenate([e["xy"] for e in exp]))
    np.testing.assert_array_equal(L.mbr.reshape(-1, 4), [e["keys"] for e in exp])
    np.testing.assert_array_equal(L.mbr2.reshape(-1, 4), [e["mbr2"] for e in exp])
    # the corner text matters: repr and %.15f of the same double give different keys
    assert any(int(O.lib().orc_coord_to_mbr(repr(v).encode())) !=
               int(O.lib().orc_coord_to_mbr(("%.15f" % v).encode()))
               for v in L.mbr2)
    # max_poly counts polygons read
    assert G.Layer.read_text(path, 5).npoly == 5


def test_reader_matches_join_generator_layers(tmp_path):
    """A generated join layer written in the reference format and read back
    gives the same coordinates and keys as the generator's own text path."""
    base, _ = G.gen_join_layers(300, 10, seed=9)
    rings = [base.ring(i) for i in range(base.npoly)]
    path = str(tmp_path / "gen.txt")
    _write_reference_layer(path, rings, corner_fmt=_fmt_vertex)
    L = G.Layer.read_text(path)
    np.testing.assert_array_equal(L.coords, base.coords)
    np.testing.assert_array_equal(L.vnum, base.vnum)
    np.testing.assert_array_equal(L.mbr, base.mbr)


def test_ring_area_closed_forms():
    sq = np.array([[0, 0], [2, 0], [2, 1], [0, 1]], float)
    assert G.ring_area(sq) == 2.0 and G.ring_area(sq[::-1]) == -2.0
    tri = np.array([[1e6, 1e6], [1e6 + 3, 1e6], [1e6, 1e6 + 4]])
    assert G.ring_area(tri) == 6.0  # relative to the first vertex: no cancellation
    assert G.ring_area(sq[:2]) == 0.0 and G.ring_area(sq[:0]) == 0.0
    for n in (3, 5, 64, 1000):
        t = 2 * np.pi * np.arange(n) / n
        reg = np.stack([np.cos(t), np.sin(t)], axis=1)
        want = 0.5 * n * np.sin(2 * np.pi / n)
        assert abs(G.ring_area(reg) - want) <= 1e-13 * want


def test_result_areas_match_shoelace():
    rs = np.random.RandomState(1)
    results = []
    for k in range(30):
        comps = []
        for c in range(int(rs.randint(0, 4))):
            n = int(rs.randint(3, 200))
            t = np.sort(rs.uniform(0, 2 * np.pi, n))
            r = rs.uniform(0.5, 1.5, n)
            comps.append(np.stack([5 + r * np.cos(t), -3 + r * np.sin(t)], axis=1))
        results.append(comps)
    got = G.result_areas(results)
    comps = [c for r in results for c in r]
    assert len(got) == len(comps)
    for a, c in zip(got, comps):
        x, y = c[:, 0] - c[0, 0], c[:, 1] - c[0, 1]
        want = 0.5 * float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))
        assert a > 0 and abs(a - want) <= 1e-12 * abs(want)


def _parse_poly(path):
    comps, cur = [], []
    for line in open(path):
        m = re.fullmatch(r"(\S+) (\S+)([,;])\n", line)
        assert m, line
        cur.append((float(m.group(1)), float(m.group(2))))
        if m.group(3) == ";":
            comps.append(np.array(cur))
            cur = []
    assert not cur
    return comps


def test_write_poly_round_trip(tmp_path):
    st, comps, _ = O.clip_cpu(*G.gen_single_pair(0, 3, 300, 200), 5)
    assert st == 0 and len(comps) > 0
    path = str(tmp_path / "r.poly")
    G.result_write(path, [comps], [st])
    back = _parse_poly(path)
    assert len(back) == len(comps)
    for a, b in zip(back, comps):
        np.testing.assert_array_equal(a, b)  # %.17g round-trips every double


def _parse_wkt(line):
    k, st, body = line.rstrip("\n").split("\t")
    if body == "MULTIPOLYGON EMPTY":
        return int(k), int(st), []
    assert body.startswith("MULTIPOLYGON (") and body.endswith(")")
    polys = []
    for ptxt in re.findall(r"\(\((.*?)\)\)", body[len("MULTIPOLYGON ("):-1]):
        rings = []
        for rtxt in ptxt.split("), ("):
            pts = np.array([[float(v) for v in p.split()] for p in rtxt.split(", ")])
            np.testing.assert_array_equal(pts[0], pts[-1])  # closed
            rings.append(pts[:-1])
        polys.append(rings)
    return int(k), int(st), polys


def test_write_wkt_single_rings(tmp_path):
    results, status = [], []
    for s in range(6):
        st, comps, _ = O.clip_cpu(*G.gen_single_pair(0, 40 + s, 150, 90), 5)
        results.append(comps)
        status.append(st)
    results.append([])
    status.append(2)
    path = str(tmp_path / "r.wkt")
    G.result_write(path, results, status, G.WRITE_WKT)
    lines = open(path).readlines()
    assert len(lines) == len(results)
    for i, line in enumerate(lines):
        k, st, polys = _parse_wkt(line)
        assert k == i and st == status[i]
        # every component a polygon of its own, without holes
        assert [len(p) for p in polys] == [1] * len(results[i])
        for p, c in zip(polys, results[i]):
            np.testing.assert_array_equal(p[0], c)


def test_write_wkt_nests_holes(tmp_path):
    """Union results of holed pairs: every hole lies inside its polygon's
    shell, and the polygons' area (shells minus holes) equals the even-odd
    area of all rings."""
    results, status = [], []
    for P, Q in random_complex_pairs(8, 13):
        st, comps, _ = O.clip_multi(P, Q, 8 | 5)
        results.append(comps)
        status.append(st)
    assert sum(len(c) for c in results) > len(results)
    path = str(tmp_path / "u.wkt")
    G.result_write(path, results, status, G.WRITE_WKT, nest=True)
    n_holes = 0
    for line, comps in zip(open(path), results):
        _, _, polys = _parse_wkt(line)
        assert sum(len(p) for p in polys) == len(comps)  # every ring placed once
        for rings in polys:
            shell = rings[0]
            for h in rings[1:]:
                n_holes += 1
                m = 0.5 * (h[0] + h[1])
                assert _inside(shell, m)
        total = sum(abs(G.ring_area(p[0])) - sum(abs(G.ring_area(h)) for h in p[1:])
                    for p in polys)
        want = _even_odd_area(comps)
        assert abs(total - want) <= 1e-9 * max(1.0, want)
    assert n_holes > 0


def _inside(ring, p):
    x, y = ring[:, 0], ring[:, 1]
    inside = False
    for i in range(len(ring)):
        j = i - 1
        if (y[i] > p[1]) != (y[j] > p[1]) and \
                p[0] < (x[j] - x[i]) * (p[1] - y[i]) / (y[j] - y[i]) + x[i]:
            inside = not inside
    return inside


def _even_odd_area(comps):
    tot = 0.0
    for i, c in enumerate(comps):
        m = 0.5 * (c[0] + c[1])
        depth = sum(_inside(d, m) for j, d in enumerate(comps) if j != i)
        tot += (-1) ** depth * abs(G.ring_area(c))
    return tot
