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
