"""The reference's call surface, driven through ctypes with the reference's
own prototypes (Join.h, ghcuda.h, main.h, ST_Intersect.cu), in the order
spatialJoin3.cu:240-412 and main.cpp:14-26 call it.  Every output is checked
against the oracle: candidate pairs and counts (mbr_sweep), CMBRs and PiP
flags, CMF counts, scans and edge lists, segment-intersection flags
(join_refine), phase-1 arrays (phase1, raw frame as calculateIntersections
uses) and clipped polygons (clip_cpu in the default frame, mode 5).
"""
import ctypes as C
import os

import numpy as np
import pytest

import gcmf_amd as G
import oracle_py as O
from conftest import large_fixture, printed

pytestmark = pytest.mark.gpu

NORM = 5  # the oracle's frame mask for the default (normalised, edge-scaled) frame

_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]
_libc.free.restype = None


def _addr(p):
    return p.value if isinstance(p, C.c_void_p) else (C.cast(p, C.c_void_p).value or 0)


def dget(gctx, ptr, dtype, n):
    out = np.zeros(int(n), dtype=dtype)
    if n:
        gctx._check(G.hip_lib().gcmf_memcpy_d2h(gctx.h, out.ctypes.data, _addr(ptr), out.nbytes),
                    "d2h")
    return out


def dfree(gctx, *ptrs):
    for p in ptrs:
        if _addr(p):
            gctx._check(G.hip_lib().gcmf_free(gctx.h, _addr(p)), "free")


def oriented(P, Q):
    return (Q, P) if len(P) < len(Q) else (P, Q)


def joined_from_ref(pairs, ref):
    """within U intersecting, in candidate order (the fixed hand-off)."""
    keep = (ref["pip_flag"] == 1) | (ref["seg_flag"] == 1)
    return pairs[keep]


def check_clips(status, results, pairs, base, over, mode=NORM):
    for k, (b, o) in enumerate(pairs):
        P, Q = oriented(base.ring(int(b)), over.ring(int(o)))
        s_ref, comps, _ = O.clip_cpu(P, Q, mode)
        assert status[k] == s_ref, (k, status[k], s_ref)
        assert [len(c) for c in results[k]] == [len(c) for c in comps], k
        for a, c in zip(results[k], comps):
            np.testing.assert_array_equal(a, c)


@pytest.fixture(scope="module")
def layers(gctx):
    base, over = G.gen_join_layers(3000, 3000, seed=11, rmin=2e-3, rmax=2e-2, nmin=16, nmax=200)
    db, do = G.DeviceLayer(gctx, base), G.DeviceLayer(gctx, over)
    yield base, over, db, do
    db.free()
    do.free()


def _adversarial_mbrs(rs, n):
    """MBR keys with inverted x or y corners, zero heights, tall boxes (so
    that some polygons touch many y strips) and repeated keys."""
    x1, y1 = rs.randint(0, 10 ** 6, n), rs.randint(0, 10 ** 6, n)
    m = np.stack([x1, y1, x1 + rs.randint(0, 3000, n), y1 + rs.randint(0, 3000, n)], 1)
    m = m.astype(np.uint64)
    flip = rs.rand(n) < 0.1
    m[flip, 1], m[flip, 3] = m[flip, 3].copy(), m[flip, 1].copy()
    flip = rs.rand(n) < 0.05
    m[flip, 0], m[flip, 2] = m[flip, 2].copy(), m[flip, 0].copy()
    m[rs.rand(n) < 0.05, 3] = m[0, 1]  # one shared y key
    z = rs.rand(n) < 0.05
    m[z, 3] = m[z, 1]
    t = rs.rand(n) < 0.03
    m[t, 3] = m[t, 1] + np.uint64(400000)
    return m


def _mbr_cases():
    rs = np.random.RandomState(21)
    base, over = G.gen_join_layers(6000, 5000, seed=23)
    cl = [(rs.randint(0, 50, (2000, 4)) * 1000).astype(np.uint64) for _ in range(2)]
    for m, w, h in ((cl[0], 1000, 1000), (cl[1], 500, 2000)):
        m[:, 2] = m[:, 0] + np.uint64(w)
        m[:, 3] = m[:, 1] + np.uint64(h)
    return {"layers": (base.mbr.reshape(-1, 4), over.mbr.reshape(-1, 4)),
            "adversarial": (_adversarial_mbrs(rs, 3000), _adversarial_mbrs(rs, 3500)),
            "clustered": tuple(cl)}


@pytest.mark.parametrize("case", ["layers", "adversarial", "clustered"])
def test_mbr_strip_sweep_equals_sweep(gctx, case):
    """SortBaseMBROverlap with the y-strip sweep forced (strip_min 0) and
    with the rank-range sweep alone: both give the oracle sweep's counts and
    ordered candidate list (orc_mbr_sweep), on generated layers, on keys
    with inverted, zero-height and tall boxes (polygons the strips hand back
    to the sweep), and on clustered boxes with repeated keys."""
    bmbr, ombr = _mbr_cases()[case]
    counts_ref, pairs_ref = O.mbr_sweep(bmbr, ombr)
    assert len(pairs_ref) > 0
    L = G.hip_lib()
    ref = G.RefContext()
    default = ref.tuning()
    db, do = gctx.upload(np.ascontiguousarray(bmbr)), gctx.upload(np.ascontiguousarray(ombr))
    try:
        for strip_min in (0, 1 << 40):
            ref.set_tuning(**dict(default, strip_min=strip_min))
            dcnt, dvec = C.c_void_p(), C.c_void_p()
            K = L.SortBaseMBROverlap(len(bmbr), len(ombr), db.ptr, do.ptr, C.byref(dcnt),
                                     C.byref(dvec), b"\x01", b"\x00")
            assert K == len(pairs_ref), strip_min
            np.testing.assert_array_equal(dget(gctx, dcnt, np.int32, len(bmbr) + len(ombr)),
                                          counts_ref)
            np.testing.assert_array_equal(dget(gctx, dvec, np.int32, 2 * K),
                                          pairs_ref.reshape(-1))
            dfree(gctx, dcnt, dvec)
    finally:
        ref.set_tuning(**default)
        db.free()
        do.free()


def test_join_operators_in_reference_order(gctx, layers):
    base, over, db, do = layers
    L = G.hip_lib()
    B, Onum = base.npoly, over.npoly
    bc, bv, bp, bm, bm2 = [b.ptr for b in db.bufs]
    oc, ov, op, om, om2 = [b.ptr for b in do.bufs]
    counts_ref, pairs_ref = O.mbr_sweep(base.mbr, over.mbr)
    ref = O.join_refine(pairs_ref, base.as_dict(), over.as_dict())

    # SortBaseMBROverlap (Join.h:5), x dimension
    dcnt, dvec = C.c_void_p(), C.c_void_p()
    K = L.SortBaseMBROverlap(B, Onum, bm, om, C.byref(dcnt), C.byref(dvec), b"\x01", b"\x00")
    assert K == len(pairs_ref) > 0
    np.testing.assert_array_equal(dget(gctx, dcnt, np.int32, B + Onum), counts_ref)
    np.testing.assert_array_equal(dget(gctx, dvec, np.int32, 2 * K), pairs_ref.reshape(-1))
    # ... and the y dimension (dimSort 1, dimSelect 1): keys with x and y exchanged
    ycnt, yvec = C.c_void_p(), C.c_void_p()
    Ky = L.SortBaseMBROverlap(B, Onum, bm, om, C.byref(ycnt), C.byref(yvec), b"\x01", b"\x01")
    swap = [1, 0, 3, 2]
    cy, py = O.mbr_sweep(base.mbr.reshape(-1, 4)[:, swap], over.mbr.reshape(-1, 4)[:, swap])
    assert Ky == len(py)
    np.testing.assert_array_equal(dget(gctx, yvec, np.int32, 2 * Ky), py.reshape(-1))
    dfree(gctx, ycnt, yvec, dcnt)

    # GetCMBR (Join.h:397)
    cmbr, pipIdx, pipFlag, pipType, joinFlag = (C.c_void_p() for _ in range(5))
    pipNum = C.c_long()
    L.GetCMBR(K, dvec, bm2, om2, C.byref(cmbr), C.byref(pipIdx), C.byref(pipFlag),
              C.byref(pipType), C.byref(joinFlag), C.byref(pipNum))
    np.testing.assert_array_equal(dget(gctx, cmbr, np.float64, 4 * K).reshape(-1, 4), ref["cmbr"])
    np.testing.assert_array_equal(dget(gctx, pipFlag, np.int32, K), ref["pip_candidate"])
    np.testing.assert_array_equal(dget(gctx, pipType, np.int8, K), ref["pip_type"])
    np.testing.assert_array_equal(dget(gctx, joinFlag, np.int32, K), np.ones(K, np.int32))
    cand = np.flatnonzero(ref["pip_candidate"])
    assert pipNum.value == len(cand)
    np.testing.assert_array_equal(dget(gctx, pipIdx, np.int32, pipNum.value), cand)

    # PointInPolygonTest (Join.h:437)
    wNum = L.PointInPolygonTest(bc, oc, K, pipNum.value, dvec, pipIdx, pipType, bp, op, pipFlag,
                                joinFlag)
    assert wNum == ref["within"]
    np.testing.assert_array_equal(dget(gctx, pipFlag, np.int32, K), ref["pip_flag"])
    np.testing.assert_array_equal(dget(gctx, joinFlag, np.int32, K), ref["join_flag_after_pip"])

    # CountCMF (Join.h:157)
    bps, ops, wps, jxy2, bEdge, oEdge = (C.c_void_p() for _ in range(6))
    eiNum, workLoad = C.c_long(), C.c_long()
    L.CountCMF(bc, oc, K, dvec, joinFlag, bv, ov, bp, op, cmbr, C.byref(bps), C.byref(ops),
               C.byref(wps), C.byref(jxy2), C.byref(bEdge), C.byref(oEdge), C.byref(eiNum),
               C.byref(workLoad))
    bcnt, ocnt = ref["b_count"].astype(np.int64), ref["o_count"].astype(np.int64)
    np.testing.assert_array_equal(dget(gctx, bps, np.int64, K), np.cumsum(bcnt))
    np.testing.assert_array_equal(dget(gctx, ops, np.int64, K), np.cumsum(ocnt))
    np.testing.assert_array_equal(dget(gctx, wps, np.int64, K), np.cumsum(bcnt * ocnt))
    assert workLoad.value == int((bcnt * ocnt).sum())
    np.testing.assert_array_equal(dget(gctx, joinFlag, np.int32, K), ref["join_flag"])
    surv = np.flatnonzero(ref["join_flag"] == 1)
    assert eiNum.value == ref["E"] == len(surv)
    np.testing.assert_array_equal(dget(gctx, jxy2, np.int32, eiNum.value), surv)

    # ApplyCMF (Join.h:257): short edge lists at the scanned offsets
    L.ApplyCMF(bc, oc, K, dvec, eiNum.value, jxy2, bv, ov, bp, op, cmbr, bps, ops, bEdge, oEdge)
    np.testing.assert_array_equal(dget(gctx, bEdge, np.int16, int(bcnt.sum())), ref["b_edges"])
    np.testing.assert_array_equal(dget(gctx, oEdge, np.int16, int(ocnt.sum())), ref["o_edges"])

    # SegmentIntersectJoin (Join.h:356): flags by survivor position
    seg = C.c_void_p()
    hits = L.SegmentIntersectJoin(bc, oc, eiNum.value, dvec, jxy2, bp, op, bps, ops, bEdge, oEdge,
                                  C.byref(seg))
    assert hits == ref["hits"]
    np.testing.assert_array_equal(dget(gctx, seg, np.int32, eiNum.value), ref["seg_flag"][surv])
    dfree(gctx, cmbr, pipIdx, pipFlag, pipType, joinFlag, bps, ops, wps, jxy2, bEdge, oEdge, seg,
          dvec)

    # ST_Intersect (ST_Intersect.cu:18): the reference's prototype returns
    # #within + #intersecting and leaves the caller's jPairs untouched
    joined = joined_from_ref(pairs_ref, ref)
    assert len(joined) == ref["within"] + ref["hits"]
    host = [base.coords, over.coords, base.vnum, over.vnum, base.vpsnum, over.vpsnum, base.mbr,
            over.mbr, base.mbr2, over.mbr2]
    args = [a.ctypes.data for a in host]
    sentinel = np.full(4, 12345, np.int32)
    n = L.ST_Intersect(B, Onum, *args, sentinel.ctypes.data)
    assert n == len(joined)
    assert (sentinel == 12345).all()
    assert L.ST_Intersect(B, Onum, *args, None) == len(joined)
    jp = C.POINTER(C.c_int)()
    n2 = L.gcmf_st_intersect_pairs(B, Onum, *args, C.byref(jp))
    assert n2 == len(joined)
    got = np.ctypeslib.as_array(jp, shape=(2 * n2,)).copy().reshape(-1, 2)
    _libc.free(C.cast(jp, C.c_void_p))
    np.testing.assert_array_equal(got, joined)

    # ghcuda (main.h:8) over the joined pairs on the resident layers
    pid = np.ascontiguousarray(joined[:, 0], np.int32)
    qid = np.ascontiguousarray(joined[:, 1], np.int32)
    ip = C.POINTER(C.c_int)
    assert L.ghcuda(pid.ctypes.data_as(ip), qid.ctypes.data_as(ip), len(joined), bc, oc, bv, bp,
                    ov, op) == 0
    rc = G.RefContext()
    st = rc.last_clip_stats()
    assert st["pairs"] == len(joined) and st["failed"] == 0
    status, res = rc.clip_results(st)
    sample = list(range(0, len(joined), max(1, len(joined) // 150)))
    check_clips(status[sample], [res[k] for k in sample], joined[sample], base, over)


def test_gh_cuda_swapped_argument(gctx, layers):
    """GH_CUDA(..., swapped) clips P = overlay ring when swapped = 1 and
    P = base ring otherwise (ghcuda.cpp:196-219), whatever the sizes."""
    base, over, db, do = layers
    L = G.hip_lib()
    bc, bv, bp = db.bufs[0].ptr, db.bufs[1].ptr, db.bufs[2].ptr
    oc, ov, op = do.bufs[0].ptr, do.bufs[1].ptr, do.bufs[2].ptr
    _, pairs = O.mbr_sweep(base.mbr, over.mbr)
    rc = G.RefContext()
    for b, o in pairs[:: max(1, len(pairs) // 12)][:12]:
        for swapped in (0, 1):
            inv = L.GH_CUDA(bc, oc, bv, bp, ov, op, int(b), int(o), swapped)
            st = rc.last_clip_stats()
            status, res = rc.clip_results(st)
            P, Q = (over.ring(o), base.ring(b)) if swapped else (base.ring(b), over.ring(o))
            s_ref, comps, _ = O.clip_cpu(P, Q, NORM)
            assert inv == (1 if s_ref == 1 else 0)
            assert status[0] == s_ref, (b, o, swapped)
            assert [len(c) for c in res[0]] == [len(c) for c in comps]
            for a, c in zip(res[0], comps):
                np.testing.assert_array_equal(a, c)


def test_calculate_intersections_arrays(gctx):
    """calculateIntersections (ghcuda.h:6) on resident layers: host arrays
    bit-exact against the restated GPU phase 1 in the raw frame, as the
    reference computes it; *invalid = 0; arrays freed with free()."""
    from conftest import GOLDEN, load_poly
    cases = []
    for pn, qn in [("Fig1-P", "Fig1-Q"), ("Fig14-P", "Fig14-Q"), ("Fig17-P-Clockwise", "Fig17-Q"),
                   ("Fig8-P", "Fig8-Q")]:
        cases.append((load_poly("%s/fig/%s.poly" % (GOLDEN, pn))[0],
                      load_poly("%s/fig/%s.poly" % (GOLDEN, qn))[0]))
    base = G.Layer.from_rings([c[0] for c in cases])
    over = G.Layer.from_rings([c[1] for c in cases])
    db, do = G.DeviceLayer(gctx, base), G.DeviceLayer(gctx, over)
    L = G.hip_lib()
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    try:
        for k, (P, Q) in enumerate(cases):
            ref = O.phase1(P, Q, 0)
            cP, cQ, inv = C.c_int(), C.c_int(), C.c_int(-7)
            iP, iQ = dp(), dp()
            aP, aQ, lP, lQ, nP, nQ = (ip() for _ in range(6))
            cm = np.zeros(4)
            L.calculateIntersections(db.bufs[0].ptr, do.bufs[0].ptr, len(P), len(Q),
                                     cm.ctypes.data_as(dp), db.bufs[1].ptr, db.bufs[2].ptr,
                                     do.bufs[1].ptr, do.bufs[2].ptr, k, k, C.byref(cP),
                                     C.byref(cQ), C.byref(iP), C.byref(iQ), C.byref(aP),
                                     C.byref(aQ), C.byref(lP), C.byref(lQ), C.byref(nP),
                                     C.byref(nQ), C.byref(inv))
            assert inv.value == ref["invalid"], k
            if ref["invalid"]:
                continue
            assert (cP.value, cQ.value) == (ref["cP"], ref["cQ"])
            g = np.ctypeslib.as_array
            np.testing.assert_array_equal(g(iP, (2 * cP.value,)), ref["intersectionsP"])
            np.testing.assert_array_equal(g(iQ, (2 * cQ.value,)), ref["intersectionsQ"])
            np.testing.assert_array_equal(g(aP, (cP.value,)), ref["alphaP"])
            np.testing.assert_array_equal(g(aQ, (cQ.value,)), ref["alphaQ"])
            np.testing.assert_array_equal(g(lP, (cP.value,)), ref["initLabelsP"])
            np.testing.assert_array_equal(g(lQ, (cQ.value,)), ref["initLabelsQ"])
            np.testing.assert_array_equal(g(nP, (cP.value,)), ref["neighborP"])
            np.testing.assert_array_equal(g(nQ, (cQ.value,)), ref["neighborQ"])
            for a in (iP, iQ, aP, aQ, lP, lQ, nP, nQ):
                _libc.free(C.cast(a, C.c_void_p))
    finally:
        db.free()
        do.free()


def test_spatial_join_then_ghcuda(gctx, tmp_path, monkeypatch):
    """main.cpp's flow: spatialJoin reads two layer files in the reference's
    text format and returns the joined pairs plus the resident layers; ghcuda
    clips them.  Pairs equal the oracle's join of the layers read back from
    the files; every resident array equals the file's; results equal the
    oracle's clips and the file GCMF_RESULT_FILE names holds them."""
    base0, over0 = G.gen_join_layers(1500, 1500, seed=21, rmin=3e-3, rmax=3e-2, nmin=12,
                                     nmax=120)
    bpath, opath = str(tmp_path / "base.txt"), str(tmp_path / "over.txt")
    base0.write_text(bpath)
    over0.write_text(opath)
    base, over = G.Layer.read_text(bpath), G.Layer.read_text(opath)
    monkeypatch.setenv("GCMF_BASE_FILE", bpath)
    monkeypatch.setenv("GCMF_OVERLAY_FILE", opath)
    L = G.hip_lib()
    ip, lp, dp = C.POINTER(C.c_int), C.POINTER(C.c_long), C.POINTER(C.c_double)
    pl, ql, n = ip(), ip(), C.c_long()
    bcs, ocs = dp(), dp()
    bvn, ovn = ip(), ip()
    bps, ops = lp(), lp()
    argv = (C.c_char_p * 1)(b"program")
    assert L.spatialJoin(1, argv, C.byref(pl), C.byref(ql), C.byref(n), C.byref(bcs),
                         C.byref(ocs), C.byref(bvn), C.byref(bps), C.byref(ovn),
                         C.byref(ops)) == 0
    _, pairs_ref = O.mbr_sweep(base.mbr, over.mbr)
    joined = joined_from_ref(pairs_ref, O.join_refine(pairs_ref, base.as_dict(), over.as_dict()))
    assert n.value == len(joined) > 0
    got = np.stack([np.ctypeslib.as_array(pl, (n.value,)), np.ctypeslib.as_array(ql, (n.value,))],
                   axis=1)
    np.testing.assert_array_equal(got, joined)
    np.testing.assert_array_equal(dget(gctx, bcs, np.float64, len(base.coords)), base.coords)
    np.testing.assert_array_equal(dget(gctx, ocs, np.float64, len(over.coords)), over.coords)
    np.testing.assert_array_equal(dget(gctx, bvn, np.int32, base.npoly), base.vnum)
    np.testing.assert_array_equal(dget(gctx, ovn, np.int32, over.npoly), over.vnum)
    np.testing.assert_array_equal(dget(gctx, bps, np.int64, base.npoly), base.vpsnum)
    np.testing.assert_array_equal(dget(gctx, ops, np.int64, over.npoly), over.vpsnum)
    rpath = str(tmp_path / "result.poly")
    monkeypatch.setenv("GCMF_RESULT_FILE", rpath)
    assert L.ghcuda(pl, ql, n.value, C.cast(bcs, C.c_void_p), C.cast(ocs, C.c_void_p),
                    C.cast(bvn, C.c_void_p), C.cast(bps, C.c_void_p), C.cast(ovn, C.c_void_p),
                    C.cast(ops, C.c_void_p)) == 0
    rc = G.RefContext()
    st = rc.last_clip_stats()
    status, res = rc.clip_results(st)
    check_clips(status, res, joined, base, over)
    # the result file: every component, %.17g (exact round trip)
    from conftest import load_poly
    comps = [c for r in res for c in r]
    written = load_poly(rpath)
    assert [len(c) for c in written] == [len(c) for c in comps]
    for a, b in zip(written, comps):
        np.testing.assert_array_equal(a, b)
    _libc.free(C.cast(pl, C.c_void_p))
    _libc.free(C.cast(ql, C.c_void_p))
    dfree(gctx, bcs, ocs, bvn, ovn, bps, ops)


@pytest.mark.parametrize("raw", [0, 1])
def test_ghcuda_reference_large_fixture(gctx, monkeypatch, raw):
    """The reference's own large pair through ghcuda, in the default frame and
    with GCMF_NORMALIZE=0 (raw coordinates, the reference's frame): status 0
    and the 11 components of GH-CUDA/results/Fig-large-R.poly, vertices equal
    at the file's printed precision."""
    S, Cq, R = large_fixture()
    base, over = G.Layer.from_rings([S]), G.Layer.from_rings([Cq])
    db, do = G.DeviceLayer(gctx, base), G.DeviceLayer(gctx, over)
    if raw:
        monkeypatch.setenv("GCMF_NORMALIZE", "0")
    L = G.hip_lib()
    z = np.zeros(1, np.int32)
    ip = C.POINTER(C.c_int)
    try:
        assert L.ghcuda(z.ctypes.data_as(ip), z.ctypes.data_as(ip), 1, db.bufs[0].ptr,
                        do.bufs[0].ptr, db.bufs[1].ptr, db.bufs[2].ptr, do.bufs[1].ptr,
                        do.bufs[2].ptr) == 0
        rc = G.RefContext()
        st = rc.last_clip_stats()
        status, res = rc.clip_results(st)
        assert status[0] == 0 and st["n_intersections"] == 47
        assert [len(c) for c in res[0]] == [len(c) for c in R]
        for a, b in zip(res[0], R):
            np.testing.assert_array_equal(printed(a), b)
    finally:
        db.free()
        do.free()


def test_ref_context_release(gctx):
    """gcmf_ref_release frees the default context; the next reference-named
    call creates a new one and works."""
    L = G.hip_lib()
    L.gcmf_ref_release()
    P, Q = G.gen_single_pair(0, 3, 60, 50)
    base, over = G.Layer.from_rings([P]), G.Layer.from_rings([Q])
    db, do = G.DeviceLayer(gctx, base), G.DeviceLayer(gctx, over)
    try:
        inv = L.GH_CUDA(db.bufs[0].ptr, do.bufs[0].ptr, db.bufs[1].ptr, db.bufs[2].ptr,
                        do.bufs[1].ptr, do.bufs[2].ptr, 0, 0, 0)
        assert inv == 0
        rc = G.RefContext()
        status, res = rc.clip_results(rc.last_clip_stats())
        s_ref, comps, _ = O.clip_cpu(P, Q, NORM)
        assert status[0] == s_ref == 0
        assert [len(c) for c in res[0]] == [len(c) for c in comps]
    finally:
        db.free()
        do.free()


REF_MAIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                        "_ref", "ref_main")


@pytest.mark.skipif(not os.path.exists(REF_MAIN),
                    reason="oracle/_ref/ref_main is built only where the reference tree is")
def test_reference_main_links_unchanged(tmp_path):
    """The reference's own caller, main.cpp (spatialJoin then ghcuda,
    main.cpp:5-33), compiled from its source and linked unchanged against
    libgcmf_hip.so (oracle/Makefile), run as a child process on two layer
    files: the result file holds exactly the oracle's clips of the joined
    pairs, in pair order."""
    import subprocess
    base0, over0 = G.gen_join_layers(800, 800, seed=31, rmin=3e-3, rmax=3e-2, nmin=12, nmax=100)
    bpath, opath = str(tmp_path / "base.txt"), str(tmp_path / "over.txt")
    rpath = str(tmp_path / "result.poly")
    base0.write_text(bpath)
    over0.write_text(opath)
    base, over = G.Layer.read_text(bpath), G.Layer.read_text(opath)
    env = dict(os.environ, GCMF_BASE_FILE=bpath, GCMF_OVERLAY_FILE=opath, GCMF_RESULT_FILE=rpath)
    r = subprocess.run([REF_MAIN], env=env, timeout=120, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    _, pairs = O.mbr_sweep(base.mbr, over.mbr)
    joined = joined_from_ref(pairs, O.join_refine(pairs, base.as_dict(), over.as_dict()))
    want = []
    for b, o in joined:
        P, Q = oriented(base.ring(int(b)), over.ring(int(o)))
        s_ref, comps, _ = O.clip_cpu(P, Q, NORM)
        want += comps
    from conftest import load_poly
    got = load_poly(rpath)
    assert len(want) > 100
    assert [len(c) for c in got] == [len(c) for c in want]
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
