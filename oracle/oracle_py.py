"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and for the
reference's own CPU GH behind oracle/_ref/libghref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("ncomp", C.c_int), ("nverts", C.c_int),
                ("comp_sizes", C.POINTER(C.c_int)), ("xy", C.POINTER(C.c_double)),
                ("n_eval", C.c_longlong), ("n_inter", C.c_longlong)]


class Phase1(C.Structure):
    _fields_ = [("nP", C.c_int), ("nQ", C.c_int), ("invalid", C.c_int),
                ("cP", C.c_int), ("cQ", C.c_int),
                ("psP1", C.POINTER(C.c_int)), ("psP2", C.POINTER(C.c_int)),
                ("psQ1", C.POINTER(C.c_int)), ("psQ2", C.POINTER(C.c_int)),
                ("boolP", C.POINTER(C.c_int)), ("boolQ", C.POINTER(C.c_int)),
                ("neighborMapQ", C.POINTER(C.c_longlong)),
                ("intersectionsP", C.POINTER(C.c_double)),
                ("intersectionsQ", C.POINTER(C.c_double)),
                ("origP", C.POINTER(C.c_double)), ("origQ", C.POINTER(C.c_double)),
                ("alphaP", C.POINTER(C.c_int)), ("alphaQ", C.POINTER(C.c_int)),
                ("initLabelsP", C.POINTER(C.c_int)), ("initLabelsQ", C.POINTER(C.c_int)),
                ("neighborP", C.POINTER(C.c_int)), ("neighborQ", C.POINTER(C.c_int)),
                ("cmbr", C.c_double * 4), ("norm", C.c_double * 3)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        L.orc_clip_cpu.argtypes = [f64p, C.c_int, f64p, C.c_int, C.c_int, C.POINTER(Result)]
        L.orc_clip_gpu_route.argtypes = L.orc_clip_cpu.argtypes
        L.orc_clip_multi.argtypes = [f64p, i32p, C.c_int, f64p, i32p, C.c_int, C.c_int,
                                     C.POINTER(Result)]
        L.orc_result_free.argtypes = [C.POINTER(Result)]
        L.orc_phase1.argtypes = [f64p, C.c_int, f64p, C.c_int, C.c_int, C.POINTER(Phase1)]
        L.orc_phase1_free.argtypes = [C.POINTER(Phase1)]
        L.orc_area_boundary.argtypes = [f64p, C.c_int, f64p, C.c_int]
        L.orc_area_boundary.restype = C.c_double
        L.orc_prefilter_check.argtypes = [f64p, C.c_int, f64p, C.c_int, C.c_int, i64p]
        L.orc_result_area.argtypes = [C.POINTER(Result), C.c_double, C.c_double]
        L.orc_result_area.restype = C.c_double
        L.orc_intersect_type.argtypes = [f64p, f64p, f64p, f64p, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
        L.orc_initial_label.argtypes = [C.c_int, C.c_int]
        L.orc_coord_to_mbr.argtypes = [C.c_char_p]
        L.orc_coord_to_mbr.restype = C.c_uint64
        L.orc_mbr_sweep.argtypes = [C.c_longlong, C.c_longlong, u64p, u64p, i32p, C.c_void_p]
        L.orc_mbr_sweep.restype = C.c_longlong
        L.orc_endpoint_sort.argtypes = [C.c_longlong, C.c_longlong, u64p, u64p, i32p, i32p]
        L.orc_get_cmbr.argtypes = [C.c_longlong, i32p, f64p, f64p, f64p, i32p, i8p, i32p]
        L.orc_pip.argtypes = [C.c_longlong, i32p, i8p, f64p, i64p, f64p, i64p, i32p, i32p]
        L.orc_pip.restype = C.c_longlong
        L.orc_cmf.argtypes = [C.c_longlong, i32p, f64p, f64p, i32p, i64p, f64p, i32p, i64p,
                              i32p, i32p, i32p, C.c_void_p, C.c_void_p]
        L.orc_cmf.restype = C.c_longlong
        L.orc_segment_join.argtypes = [C.c_longlong, i32p, i32p, f64p, i64p, f64p, i64p,
                                       i32p, i32p, i16p, i16p, i32p]
        L.orc_segment_join.restype = C.c_longlong
        _lib = L
    return _lib


def ref_lib():
    """The reference's CPU GH (oracle/_ref), or None when it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libghref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_clip.argtypes = [f64p, C.c_int, f64p, C.c_int, f64p, C.c_int, i32p, C.c_int,
                               C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ref_clip_multi.argtypes = [f64p, i32p, C.c_int, f64p, i32p, C.c_int, C.c_int, f64p,
                                     C.c_int, i32p, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
        _ref = L
    return _ref


def _ring(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


def _unpack(r):
    comps = []
    off = 0
    for c in range(r.ncomp):
        n = r.comp_sizes[c]
        xy = np.ctypeslib.as_array(r.xy, shape=(2 * r.nverts,))[2 * off:2 * (off + n)]
        comps.append(xy.reshape(-1, 2).copy())
        off += n
    return comps


def clip_cpu(P, Q, normalize=True):
    """CPU-route GH intersection.  Returns (status, [component arrays], stats)."""
    L = lib()
    p, q = _ring(P), _ring(Q)
    r = Result()
    L.orc_clip_cpu(p, len(p) // 2, q, len(q) // 2, int(normalize), C.byref(r))
    out = (r.status, _unpack(r), {"n_eval": r.n_eval, "n_inter": r.n_inter})
    L.orc_result_free(C.byref(r))
    return out


def _rings(rings):
    rings = [np.asarray(r, dtype=np.float64).reshape(-1, 2) for r in rings]
    xy = np.ascontiguousarray(np.concatenate(rings).reshape(-1)) if rings else np.zeros(0)
    return xy, np.array([len(r) for r in rings], dtype=np.int32)


UNION = 8  # mode bit of clip_multi


def clip_multi(Ps, Qs, mode=5):
    """CPU route over complex polygons (lists of rings).  mode bits: 1
    normalise, 2 no clean-up, 4 edge scale, 8 union.  Returns (status,
    [component arrays], stats)."""
    L = lib()
    p, ps = _rings(Ps)
    q, qs = _rings(Qs)
    r = Result()
    L.orc_clip_multi(p, ps, len(ps), q, qs, len(qs), int(mode), C.byref(r))
    out = (r.status, _unpack(r), {"n_eval": r.n_eval, "n_inter": r.n_inter})
    L.orc_result_free(C.byref(r))
    return out


def ref_clip_multi(Ps, Qs, union=False):
    """The reference's CPU GH over complex polygons (raw coordinates)."""
    L = ref_lib()
    if L is None:
        return None
    p, ps = _rings(Ps)
    q, qs = _rings(Qs)
    cap = 8 * (len(p) + len(q)) + 64
    xy = np.zeros(2 * cap, dtype=np.float64)
    sizes = np.zeros(cap, dtype=np.int32)
    nc, nv = C.c_int(), C.c_int()
    rc = L.ref_clip_multi(p, ps, len(ps), q, qs, len(qs), int(union), xy, cap, sizes, cap,
                          C.byref(nc), C.byref(nv))
    if rc != 0:
        raise RuntimeError("reference clip output overflow")
    comps, off = [], 0
    for c in range(nc.value):
        n = int(sizes[c])
        comps.append(xy[2 * off:2 * (off + n)].reshape(-1, 2).copy())
        off += n
    return comps


def clip_gpu_route(P, Q, normalize=True):
    L = lib()
    p, q = _ring(P), _ring(Q)
    r = Result()
    L.orc_clip_gpu_route(p, len(p) // 2, q, len(q) // 2, int(normalize), C.byref(r))
    out = (r.status, _unpack(r), {"n_inter": r.n_inter})
    L.orc_result_free(C.byref(r))
    return out


def phase1(P, Q, normalize=True):
    """Reference GPU phase-1 arrays for one pair, as numpy arrays."""
    L = lib()
    p, q = _ring(P), _ring(Q)
    t = Phase1()
    rc = L.orc_phase1(p, len(p) // 2, q, len(q) // 2, int(normalize), C.byref(t))
    out = {"rc": rc, "invalid": t.invalid, "nP": t.nP, "nQ": t.nQ, "cP": t.cP, "cQ": t.cQ,
           "cmbr": np.array(t.cmbr[:]), "norm": np.array(t.norm[:])}
    if rc != 2:
        g = np.ctypeslib.as_array
        out.update(psP1=g(t.psP1, (t.nP + 1,)).copy(), psP2=g(t.psP2, (t.nP + 1,)).copy(),
                   psQ1=g(t.psQ1, (t.nQ + 1,)).copy(), psQ2=g(t.psQ2, (t.nQ + 1,)).copy(),
                   boolP=g(t.boolP, (t.nP,)).copy(), boolQ=g(t.boolQ, (t.nQ,)).copy())
    if rc == 0:
        g = np.ctypeslib.as_array
        cP, cQ = max(t.cP, 1), max(t.cQ, 1)
        out.update(
            neighborMapQ=g(t.neighborMapQ, (cQ,))[:t.cQ].copy(),
            intersectionsP=g(t.intersectionsP, (2 * cP,))[:2 * t.cP].copy(),
            intersectionsQ=g(t.intersectionsQ, (2 * cQ,))[:2 * t.cQ].copy(),
            origP=g(t.origP, (2 * cP,))[:2 * t.cP].copy(),
            origQ=g(t.origQ, (2 * cQ,))[:2 * t.cQ].copy(),
            alphaP=g(t.alphaP, (cP,))[:t.cP].copy(), alphaQ=g(t.alphaQ, (cQ,))[:t.cQ].copy(),
            initLabelsP=g(t.initLabelsP, (cP,))[:t.cP].copy(),
            initLabelsQ=g(t.initLabelsQ, (cQ,))[:t.cQ].copy(),
            neighborP=g(t.neighborP, (cP,))[:t.cP].copy(),
            neighborQ=g(t.neighborQ, (cQ,))[:t.cQ].copy())
    L.orc_phase1_free(C.byref(t))
    return out


def prefilter_check(P, Q, normalize=True):
    """Counts of the CPU-route prefilter check: pairs, skipped, unsound
    skips, dropped edges, unsound edge drops (see orc_prefilter_check)."""
    P, Q = _ring(P), _ring(Q)
    out = np.zeros(5, np.int64)
    rc = lib().orc_prefilter_check(P, len(P) // 2, Q, len(Q) // 2, int(normalize), out)
    return rc, dict(zip(["pairs", "skipped", "bad_skips", "dropped", "bad_drops"], out.tolist()))


def area_boundary(P, Q):
    p, q = _ring(P), _ring(Q)
    return lib().orc_area_boundary(p, len(p) // 2, q, len(q) // 2)


def comps_area(comps, ox=0.0, oy=0.0):
    """Sum of |shoelace| areas of result components about (ox, oy)."""
    tot = 0.0
    for c in comps:
        x = c[:, 0] - ox
        y = c[:, 1] - oy
        tot += abs(0.5 * float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)))
    return tot


def intersect_type(P1, P2, Q1, Q2):
    a, b = C.c_double(), C.c_double()
    t = lib().orc_intersect_type(_ring(P1), _ring(P2), _ring(Q1), _ring(Q2), C.byref(a),
                                 C.byref(b))
    return t, a.value, b.value


def ref_clip(P, Q):
    """Run the reference's CPU GH (raw coordinates, EPSILON 1e-7 absolute)."""
    L = ref_lib()
    if L is None:
        return None
    p, q = _ring(P), _ring(Q)
    cap = 8 * (len(p) + len(q)) + 64
    xy = np.zeros(2 * cap, dtype=np.float64)
    sizes = np.zeros(cap, dtype=np.int32)
    nc, nv = C.c_int(), C.c_int()
    rc = L.ref_clip(p, len(p) // 2, q, len(q) // 2, xy, cap, sizes, cap, C.byref(nc), C.byref(nv))
    if rc != 0:
        raise RuntimeError("reference clip output overflow")
    comps, off = [], 0
    for c in range(nc.value):
        n = int(sizes[c])
        comps.append(xy[2 * off:2 * (off + n)].reshape(-1, 2).copy())
        off += n
    return comps


# ---------------------------------------------------------------- join ----
def mbr_sweep(bMBR, oMBR):
    L = lib()
    b = np.ascontiguousarray(bMBR, dtype=np.uint64).reshape(-1)
    o = np.ascontiguousarray(oMBR, dtype=np.uint64).reshape(-1)
    B, O = len(b) // 4, len(o) // 4
    counts = np.zeros(B + O, dtype=np.int32)
    K = L.orc_mbr_sweep(B, O, b, o, counts, None)
    pairs = np.zeros(2 * max(K, 1), dtype=np.int32)
    L.orc_mbr_sweep(B, O, b, o, counts, pairs.ctypes.data_as(C.c_void_p))
    return counts, pairs[:2 * K].reshape(-1, 2)


def endpoint_sort(bMBR, oMBR):
    L = lib()
    b = np.ascontiguousarray(bMBR, dtype=np.uint64).reshape(-1)
    o = np.ascontiguousarray(oMBR, dtype=np.uint64).reshape(-1)
    n = 2 * (len(b) // 4 + len(o) // 4)
    perm = np.zeros(max(n, 1), dtype=np.int32)
    rank = np.zeros(max(n, 1), dtype=np.int32)
    L.orc_endpoint_sort(len(b) // 4, len(o) // 4, b, o, perm, rank)
    return perm[:n], rank[:n]


def join_refine(pairs, base, over):
    """GetCMBR -> PiP -> CMF -> segment join on layer dicts (coords, vnum,
    vpsnum, mbr2).  Returns a dict of the per-pair flags and edge lists."""
    L = lib()
    pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
    K = len(pr) // 2
    cm = np.zeros(4 * max(K, 1))
    pip = np.zeros(max(K, 1), dtype=np.int32)
    ptype = np.zeros(max(K, 1), dtype=np.int8)
    jf = np.zeros(max(K, 1), dtype=np.int32)
    L.orc_get_cmbr(K, pr, base["mbr2"], over["mbr2"], cm, pip, ptype, jf)
    pip_cand = pip.copy()
    within = L.orc_pip(K, pr, ptype, base["coords"], base["vpsnum"], over["coords"],
                       over["vpsnum"], pip, jf)
    bc = np.zeros(max(K, 1), dtype=np.int32)
    oc = np.zeros(max(K, 1), dtype=np.int32)
    jf_cmf = jf.copy()
    E = L.orc_cmf(K, pr, cm, base["coords"], base["vnum"], base["vpsnum"], over["coords"],
                  over["vnum"], over["vpsnum"], jf_cmf, bc, oc, None, None)
    be = np.zeros(max(int(bc.sum()), 1), dtype=np.int16)
    oe = np.zeros(max(int(oc.sum()), 1), dtype=np.int16)
    jf2 = jf.copy()
    L.orc_cmf(K, pr, cm, base["coords"], base["vnum"], base["vpsnum"], over["coords"],
              over["vnum"], over["vpsnum"], jf2, bc, oc, be.ctypes.data_as(C.c_void_p),
              oe.ctypes.data_as(C.c_void_p))
    seg = np.zeros(max(K, 1), dtype=np.int32)
    hits = L.orc_segment_join(K, pr, jf2, base["coords"], base["vpsnum"], over["coords"],
                              over["vpsnum"], bc, oc, be, oe, seg)
    return {"cmbr": cm[:4 * K].reshape(-1, 4), "pip_candidate": pip_cand[:K],
            "pip_type": ptype[:K], "pip_flag": pip[:K], "within": within,
            "join_flag_after_pip": jf[:K], "join_flag": jf2[:K], "E": E,
            "b_count": bc[:K], "o_count": oc[:K], "b_edges": be[:int(bc.sum())],
            "o_edges": oe[:int(oc.sum())], "seg_flag": seg[:K], "hits": hits}


# ------------------------------------------------------ planted mistakes ----
MUT_SEG_OPEN, MUT_SEG_NO_PARALLEL, MUT_SEG_NO_VERTICAL = 1, 2, 3
MUT_PIP_BOTH_ENDS, MUT_PIP_CLOSED_Y, MUT_PIP_LE = 4, 5, 6
MUT_CMF_OPEN, MUT_CMF_NO_CLOSING = 7, 8


class mutant:
    """Plants mistake m in the join oracle for the duration of a with-block
    (orc_mutant), so a test can show that a pin fails on a wrong oracle."""

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        C.c_int.in_dll(lib(), "orc_mutant").value = self.m
        return self

    def __exit__(self, *exc):
        C.c_int.in_dll(lib(), "orc_mutant").value = 0
        return False


def pip_flags(pairs, pip_type, base, over):
    """orc_pip on explicit pip types: (flags, within)."""
    L = lib()
    pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
    K = len(pr) // 2
    pt = np.ascontiguousarray(pip_type, dtype=np.int8)
    flag = np.zeros(max(K, 1), dtype=np.int32)
    jf = np.ones(max(K, 1), dtype=np.int32)
    w = L.orc_pip(K, pr, pt, base["coords"], base["vpsnum"], over["coords"], over["vpsnum"],
                  flag, jf)
    return flag[:K], w


def cmf(pairs, cmbr, base, over):
    """orc_cmf on explicit CMBRs with every join flag set: (flags, b_count,
    o_count, b_edges, o_edges, E)."""
    L = lib()
    pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
    K = len(pr) // 2
    cm = np.ascontiguousarray(cmbr, dtype=np.float64).reshape(-1)
    jf = np.ones(max(K, 1), dtype=np.int32)
    bc = np.zeros(max(K, 1), dtype=np.int32)
    oc = np.zeros(max(K, 1), dtype=np.int32)
    L.orc_cmf(K, pr, cm, base["coords"], base["vnum"], base["vpsnum"], over["coords"],
              over["vnum"], over["vpsnum"], jf.copy(), bc, oc, None, None)
    be = np.zeros(max(int(bc.sum()), 1), dtype=np.int16)
    oe = np.zeros(max(int(oc.sum()), 1), dtype=np.int16)
    E = L.orc_cmf(K, pr, cm, base["coords"], base["vnum"], base["vpsnum"], over["coords"],
                  over["vnum"], over["vpsnum"], jf, bc, oc, be.ctypes.data_as(C.c_void_p),
                  oe.ctypes.data_as(C.c_void_p))
    return jf[:K], bc[:K], oc[:K], be[:int(bc.sum())], oe[:int(oc.sum())], E


def segment_join(pairs, base, over, b_count, o_count, b_edges, o_edges):
    """orc_segment_join with every join flag set: (seg_flag, hits)."""
    L = lib()
    pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
    K = len(pr) // 2
    jf = np.ones(max(K, 1), dtype=np.int32)
    seg = np.zeros(max(K, 1), dtype=np.int32)
    hits = L.orc_segment_join(K, pr, jf, base["coords"], base["vpsnum"], over["coords"],
                              over["vpsnum"], np.ascontiguousarray(b_count, np.int32),
                              np.ascontiguousarray(o_count, np.int32),
                              np.ascontiguousarray(b_edges, np.int16),
                              np.ascontiguousarray(o_edges, np.int16), seg)
    return seg[:K], hits

