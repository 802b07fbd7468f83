"""The result writers and areas on the device's results: gcmf_clip_areas
(computed on the GPU, one thread per component) equals the host's
gcmf_result_areas of the same results bit for bit, and gcmf_clip_write
writes the same .poly and WKT files as the host writers, for a batch of
single-ring pairs (every component its own polygon) and for complex pairs
clipped by the union (components nested by the even-odd rule)."""
import numpy as np
import pytest

import gcmf_amd as G
from test_gpu_complex import run_complex
from test_gpu_clip import fig_cases, random_cases, run_batch
from test_oracle_multi import random_complex_pairs

pytestmark = pytest.mark.gpu


def _same_files(a, b):
    assert open(a).read() == open(b).read()


def test_clip_areas_and_files(gctx, tmp_path):
    cases = fig_cases() + random_cases(40, seed0=31, nmax=400)
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE)
    status, res = gctx.clip_results(st)
    areas = gctx.clip_areas(st)
    np.testing.assert_array_equal(areas, G.result_areas(res))
    assert len(areas) == st["total_comps"] > 0
    for fmt, name in ((G.WRITE_POLY, "r.poly"), (G.WRITE_WKT, "r.wkt")):
        dev, host = str(tmp_path / ("dev_" + name)), str(tmp_path / ("host_" + name))
        gctx.clip_write(dev, fmt)
        G.result_write(host, res, status, fmt)
        _same_files(dev, host)


def test_complex_union_areas_and_wkt(gctx, tmp_path):
    cases = random_complex_pairs(12, 13)
    st, (status, res) = run_complex(gctx, cases, G.CLIP_NORMALIZE | G.CLIP_UNION)
    np.testing.assert_array_equal(gctx.clip_areas(st), G.result_areas(res))
    dev, host = str(tmp_path / "dev.wkt"), str(tmp_path / "host.wkt")
    gctx.clip_write(dev, G.WRITE_WKT)
    G.result_write(host, res, status, G.WRITE_WKT, nest=True)
    _same_files(dev, host)
    assert "), (" in open(dev).read()  # some polygon has a hole


def test_clip_write_rejects_bad_format(gctx, tmp_path):
    cases = random_cases(2, seed0=5)
    run_batch(gctx, cases, G.CLIP_NORMALIZE)
    with pytest.raises(G.GcmfError):
        gctx.clip_write(str(tmp_path / "x"), 7)
