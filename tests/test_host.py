"""Host-side checks that need no GPU: the C-ABI libraries load and export
every function their headers declare; generators are deterministic and
follow their documented distributions; the text layer format round-trips."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import gcmf_amd as G
from conftest import ROOT


def declared_functions(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"^\s*#.*$", "", src, flags=re.M)  # preprocessor lines
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", src)
    return sorted(set(n for n in names if n not in ("defined",)))


@pytest.mark.parametrize("header,lib", [("gcmf_hip.h", "libgcmf_hip.so"),
                                        ("gcmf_host.h", "libgcmf_host.so")])
def test_library_exports_header(header, lib):
    path = os.path.join(ROOT, "gcmf-gh_amd", lib)
    L = C.CDLL(path)
    names = declared_functions(header)
    assert len(names) > 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_entry_points_exported():
    L = C.CDLL(os.path.join(ROOT, "gcmf-gh_amd", "libgcmf_hip.so"))
    for n in ("SortBaseMBROverlap", "GetCMBR", "PointInPolygonTest", "CountCMF", "ApplyCMF",
              "SegmentIntersectJoin", "calculateIntersections", "GH_CUDA", "ghcuda",
              "ST_Intersect", "spatialJoin", "gcmf_st_intersect_pairs", "gcmf_ref_context"):
        assert hasattr(L, n), n


def test_reference_cxx_caller_links(tmp_path):
    """A C++ caller that declares the reference's entry points as the
    reference's headers do (C++ linkage, tests/ref_link/caller.cpp) links
    against libgcmf_hip.so: every mangled name is exported."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = os.path.join(ROOT, "tests", "ref_link", "caller.cpp")
    out = str(tmp_path / "caller")
    r = subprocess.run(["g++", "-std=c++11", src, "-L" + os.path.join(ROOT, "gcmf-gh_amd"),
                        "-lgcmf_hip", "-Wl,--allow-shlib-undefined", "-o", out],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_generators_deterministic():
    a = G.gen_join_layers(300, 200, seed=4)
    b = G.gen_join_layers(300, 200, seed=4)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.coords, y.coords)
        np.testing.assert_array_equal(x.mbr, y.mbr)
    c = G.gen_join_layers(300, 200, seed=5)
    assert not np.array_equal(a[0].coords, c[0].coords)


def test_join_layer_properties():
    base, over = G.gen_join_layers(2000, 1000, seed=0)
    for lay, n in ((base, 2000), (over, 1000)):
        assert lay.npoly == n
        v = lay.vnum - 1
        assert v.min() >= 16 and v.max() <= 512
        assert lay.vpsnum[-1] == lay.nverts == lay.vnum.sum()
        for i in (0, n // 2, n - 1):
            r = lay.ring(i)
            end = lay.vpsnum[i]
            closing = lay.coords[2 * (end - 1):2 * end]
            np.testing.assert_array_equal(closing, r[0])
            m = lay.mbr2[4 * i:4 * i + 4]
            assert m[0] == r[:, 0].min() and m[2] == r[:, 0].max()
            assert m[1] == r[:, 1].min() and m[3] == r[:, 1].max()
            # coordinates went through %.15f text
            assert all(float("%.15f" % x) == x for x in r[:4, 0])
            assert G.coord_to_mbr("%.15f" % m[0]) == int(lay.mbr[4 * i])
        ext = np.array([lay.mbr2[4 * i + 2] - lay.mbr2[4 * i] for i in range(n)])
        assert ext.min() > 1e-4 and ext.max() < 2 * 1.3 * 5e-3 + 1e-12


def test_pair_layers_properties():
    P, Q = G.gen_pair_layers(5000, seed=0)
    n = np.concatenate([P.vnum, Q.vnum]) - 1
    assert n.min() >= 8 and n.max() <= 50000
    assert 40 < np.median(n) < 100
    # 70 % overlapping pairs have their centre within 1.4, the rest at 3
    cq = np.array([Q.ring(i).mean(axis=0) for i in range(0, 5000, 50)])
    d = np.hypot(cq[:, 0], cq[:, 1])
    assert np.all((d < 1.6) | (d > 2.6))


def test_text_format_round_trip(tmp_path):
    base, _ = G.gen_join_layers(50, 10, seed=2)
    path = str(tmp_path / "layer.txt")
    base.write_text(path)
    back = G.Layer.read_text(path)
    np.testing.assert_array_equal(back.coords, base.coords)
    np.testing.assert_array_equal(back.vnum, base.vnum)
    np.testing.assert_array_equal(back.vpsnum, base.vpsnum)
    np.testing.assert_array_equal(back.mbr, base.mbr)
    np.testing.assert_array_equal(back.mbr2, base.mbr2)


def test_gpu_path_fails_loudly_without_device():
    import gcmf_amd
    if gcmf_amd.hip_lib().gcmf_device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(gcmf_amd.GcmfError):
        gcmf_amd.Context(0)
