import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "gcmf-gh_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def gctx():
    """One GPU context for the whole session (tests run in one process)."""
    import gcmf_amd as G
    ctx = G.Context(0)
    yield ctx
    ctx.close()


def load_poly(path):
    """Reads the reference's .poly format ('x y,' lines, ';' ends a ring)."""
    txt = open(path).read()
    polys = []
    for part in txt.split(";"):
        pts = [p.split() for p in part.split(",") if p.split()]
        if pts:
            polys.append(np.array([[float(a), float(b)] for a, b in pts]))
    return polys


GOLDEN = os.path.join(ROOT, "tests", "golden")
REFDATA = os.path.join(GOLDEN, "reference")


def _text(path):
    import gzip
    if path.endswith(".gz"):
        with gzip.open(path, "rt") as f:
            return f.read()
    with open(path) as f:
        return f.read()


def load_poly_text(path):
    """load_poly for plain or gzip-compressed files."""
    polys = []
    for part in _text(path).split(";"):
        pts = [p.split() for p in part.split(",") if p.split()]
        if pts:
            polys.append(np.array([[float(a), float(b)] for a, b in pts]))
    return polys


def load_count_ring(path):
    """Reads the readPolygon format of the reference's large pair: a vertex
    count line, then one "x,y" vertex per line (gpc_read_polygon,
    GH-CUDA/lib/polyclip.h:903-913)."""
    lines = _text(path).split()
    n = int(lines[0])
    return np.array([[float(v) for v in ln.split(",")] for ln in lines[1:1 + n]])


def large_fixture():
    """The reference's large pair s.txt x c.txt and its result
    Fig-large-R.poly (GH-CUDA/ReadMe.txt:59: 11 components)."""
    S = load_count_ring(os.path.join(REFDATA, "s.txt.gz"))
    Cq = load_count_ring(os.path.join(REFDATA, "c.txt.gz"))
    R = load_poly_text(os.path.join(REFDATA, "Fig-large-R.poly.gz"))
    return S, Cq, R


def printed(comp):
    """Vertices as the reference prints them: ostream's default precision,
    6 significant digits (point2D.h operator<<)."""
    return np.array([[float("%.6g" % v) for v in row] for row in comp])
