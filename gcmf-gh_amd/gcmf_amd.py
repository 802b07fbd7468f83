"""Python host mirror of the GCMF-GH join-and-clip path on MI355X.

Loads two in-tree shared libraries:
  libgcmf_host.so -- layers, CoordToMBR keys, text format, generators (CPU)
  libgcmf_hip.so  -- the HIP kernels behind the C ABI of include/gcmf_hip.h

Operation names follow the reference's operator API (Join.h, ghcuda.cpp,
gppolyclip_mDIV512.cu).  Every GPU operation calls the HIP library; there is
no CPU fallback: a missing library or device raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_host = None
_hip = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class GcmfError(RuntimeError):
    pass


def host_lib():
    global _host
    if _host is None:
        path = os.path.join(HERE, "libgcmf_host.so")
        if not os.path.exists(path):
            raise GcmfError("libgcmf_host.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.gcmf_coord_to_mbr.argtypes = [C.c_char_p]
        L.gcmf_coord_to_mbr.restype = C.c_uint64
        for name, rt in [("gcmf_layer_npoly", C.c_long), ("gcmf_layer_nverts", C.c_long),
                         ("gcmf_layer_coords", vp), ("gcmf_layer_vnum", vp),
                         ("gcmf_layer_vpsnum", vp), ("gcmf_layer_mbr", vp),
                         ("gcmf_layer_mbr2", vp)]:
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = rt
        L.gcmf_layer_free.argtypes = [vp]
        L.gcmf_layer_from_rings.argtypes = [f64p, i32p, C.c_long, C.c_int]
        L.gcmf_layer_from_rings.restype = vp
        L.gcmf_layer_read_text.argtypes = [C.c_char_p, C.c_long]
        L.gcmf_layer_read_text.restype = vp
        L.gcmf_layer_write_text.argtypes = [vp, C.c_char_p]
        L.gcmf_gen_join_layers.argtypes = [C.c_uint64, C.c_long, C.c_long, C.c_double,
                                           C.c_double, C.c_int, C.c_int, C.POINTER(vp),
                                           C.POINTER(vp)]
        L.gcmf_gen_pair_layers.argtypes = [C.c_uint64, C.c_long, C.c_double, C.c_double,
                                           C.c_int, C.c_int, C.c_double, C.POINTER(vp),
                                           C.POINTER(vp)]
        L.gcmf_gen_single_pair.argtypes = [C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_double,
                                           C.POINTER(vp), C.POINTER(vp)]
        L.gcmf_ring_area.argtypes = [f64p, C.c_long]
        L.gcmf_ring_area.restype = C.c_double
        L.gcmf_result_areas.argtypes = [f64p, i32p, C.c_long, f64p]
        L.gcmf_result_write_poly.argtypes = [C.c_char_p, f64p, i32p, C.c_long]
        L.gcmf_result_write_wkt.argtypes = [C.c_char_p, f64p, i32p, i32p, i32p, C.c_long,
                                            C.c_int]
        _host = L
    return _host


class Layer:
    """A polygon layer in the reference's layout, as numpy arrays."""

    def __init__(self, coords, vnum, vpsnum, mbr, mbr2):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.vnum = np.ascontiguousarray(vnum, dtype=np.int32)
        self.vpsnum = np.ascontiguousarray(vpsnum, dtype=np.int64)
        self.mbr = np.ascontiguousarray(mbr, dtype=np.uint64)
        self.mbr2 = np.ascontiguousarray(mbr2, dtype=np.float64)

    @property
    def npoly(self):
        return len(self.vnum)

    @property
    def nverts(self):
        return len(self.coords) // 2

    def ring(self, i):
        """Ring i without its closing vertex, shape (VNum-1, 2)."""
        end = int(self.vpsnum[i])
        start = end - int(self.vnum[i])
        return self.coords[2 * start:2 * (end - 1)].reshape(-1, 2)

    def as_dict(self):
        return {"coords": self.coords, "vnum": self.vnum, "vpsnum": self.vpsnum,
                "mbr": self.mbr, "mbr2": self.mbr2}

    @staticmethod
    def _from_handle(h):
        L = host_lib()
        n = L.gcmf_layer_npoly(h)
        nv = L.gcmf_layer_nverts(h)

        def arr(ptr, ctype, count, dtype):
            if count == 0:
                return np.zeros(0, dtype=dtype)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,)).astype(
                dtype, copy=True)

        lay = Layer(arr(L.gcmf_layer_coords(h), C.c_double, 2 * nv, np.float64),
                    arr(L.gcmf_layer_vnum(h), C.c_int, n, np.int32),
                    arr(L.gcmf_layer_vpsnum(h), C.c_long, n, np.int64),
                    arr(L.gcmf_layer_mbr(h), C.c_uint64, 4 * n, np.uint64),
                    arr(L.gcmf_layer_mbr2(h), C.c_double, 4 * n, np.float64))
        L.gcmf_layer_free(h)
        return lay

    @staticmethod
    def from_rings(rings, text_round_trip=False):
        n = np.array([len(r) for r in rings], dtype=np.int32)
        xy = np.ascontiguousarray(np.concatenate([np.asarray(r, dtype=np.float64).reshape(-1)
                                                  for r in rings]))
        h = host_lib().gcmf_layer_from_rings(xy, n, len(n), int(text_round_trip))
        return Layer._from_handle(h)

    @staticmethod
    def read_text(path, max_poly=-1):
        h = host_lib().gcmf_layer_read_text(path.encode(), max_poly)
        if not h:
            raise GcmfError("cannot read layer file %s" % path)
        return Layer._from_handle(h)

    def write_text(self, path):
        L = host_lib()
        rings = [self.ring(i) for i in range(self.npoly)]
        n = np.array([len(r) for r in rings], dtype=np.int32)
        xy = np.ascontiguousarray(np.concatenate([r.reshape(-1) for r in rings]))
        h = L.gcmf_layer_from_rings(xy, n, len(n), 0)
        rc = L.gcmf_layer_write_text(h, path.encode())
        L.gcmf_layer_free(h)
        if rc != 0:
            raise GcmfError("cannot write %s" % path)


def coord_to_mbr(token):
    return int(host_lib().gcmf_coord_to_mbr(token.encode() if isinstance(token, str) else token))


def gen_join_layers(n_base=20000, n_over=20000, seed=0, rmin=1e-4, rmax=5e-3, nmin=16, nmax=512):
    """Two join layers of star-shaped polygons in the unit square."""
    b, o = C.c_void_p(), C.c_void_p()
    host_lib().gcmf_gen_join_layers(seed, n_base, n_over, rmin, rmax, nmin, nmax, C.byref(b),
                                    C.byref(o))
    return Layer._from_handle(b), Layer._from_handle(o)


def gen_pair_layers(n_pairs=200000, seed=0, median=64.0, sigma=1.5, nmin=8, nmax=50000,
                    overlap=0.7):
    """Independent uneven pairs (P[i], Q[i]) as two layers."""
    p, q = C.c_void_p(), C.c_void_p()
    host_lib().gcmf_gen_pair_layers(seed, n_pairs, median, sigma, nmin, nmax, overlap,
                                    C.byref(p), C.byref(q))
    return Layer._from_handle(p), Layer._from_handle(q)


def gen_single_pair(kind=0, seed=0, nP=2000, nQ=1500, snap=0.0):
    """One pair as rings (P, Q) without closing vertices."""
    p, q = C.c_void_p(), C.c_void_p()
    host_lib().gcmf_gen_single_pair(kind, seed, nP, nQ, snap, C.byref(p), C.byref(q))
    return Layer._from_handle(p).ring(0), Layer._from_handle(q).ring(0)


# ---------------------------------------------------------------- GPU ----
class JoinStats(C.Structure):
    _fields_ = [(n, C.c_long) for n in ("candidates", "pip_candidates", "within",
                                        "cmf_candidates", "survivors", "b_edges", "o_edges",
                                        "intersecting", "joined")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ClipStats(C.Structure):
    _fields_ = [(n, C.c_long) for n in ("pairs", "ok", "invalid", "disjoint", "failed",
                                        "total_comps", "total_verts", "edges_p", "edges_q",
                                        "slots_p", "slots_q")] + \
               [("n_eval", C.c_longlong), ("n_intersections", C.c_longlong),
                ("n_tested", C.c_longlong)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("cull_min", C.c_long), ("big_nodes", C.c_longlong), ("count_block", C.c_int),
                ("lds_max_q", C.c_int), ("label_lds", C.c_int), ("grid_nodes", C.c_longlong),
                ("grid_clean", C.c_int), ("strip_min", C.c_longlong), ("cmf_group", C.c_int),
                ("label_steal", C.c_int), ("block_cap", C.c_int), ("count_stage", C.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class DLayer(C.Structure):
    _fields_ = [("npoly", C.c_long), ("d_coords", C.c_void_p), ("d_vnum", C.c_void_p),
                ("d_vpsnum", C.c_void_p), ("d_mbr", C.c_void_p), ("d_mbr2", C.c_void_p)]


# field ids of gcmf_join_get / gcmf_clip_get (include/gcmf_hip.h)
J_FIELDS = {"pairs": (0, np.int32), "counts": (1, np.int32), "perm": (2, np.int32),
            "rank": (3, np.int32), "cmbr": (4, np.float64), "pip_type": (5, np.int8),
            "pip_flag": (6, np.int32), "join_flag": (7, np.int32), "b_count": (8, np.int32),
            "o_count": (9, np.int32), "b_edges": (10, np.int32), "o_edges": (11, np.int32),
            "seg_flag": (12, np.int32), "joined": (13, np.int32)}
C_FIELDS = {"status": (0, np.int32), "swapped": (1, np.int32), "ncomp": (2, np.int32),
            "nverts": (3, np.int32), "comp_off": (4, np.int64), "vert_off": (5, np.int64),
            "comp_sizes": (6, np.int32), "xy": (7, np.float64), "label_path": (8, np.int32),
            "label_ticks": (9, np.int64), "offP": (20, np.int64),
            "offQ": (21, np.int64), "psP1": (22, np.int64), "psP2": (23, np.int64),
            "psQ1": (24, np.int64), "psQ2": (25, np.int64), "boolP": (26, np.int32),
            "boolQ": (27, np.int32), "nmapQ": (28, np.int32), "iP": (29, np.float64),
            "iQ": (30, np.float64), "oP": (31, np.float64), "oQ": (32, np.float64),
            "aP": (33, np.int32), "aQ": (34, np.int32), "lP": (35, np.int32),
            "lQ": (36, np.int32), "nbP": (37, np.int32), "nbQ": (38, np.int32),
            "norm": (39, np.float64)}

# labelling path of a pair (clip field "label_path")
LP_SKIPPED, LP_SEGMENTS, LP_SEG_CLEANUP, LP_NO_CROSSING, LP_SEQUENTIAL, LP_FAILED = range(6)

CLIP_NORMALIZE = 1
CLIP_NO_CLEANUP = 2
CLIP_SEQ_LABEL = 4
CLIP_GPU_ROUTE = 8
CLIP_SCALE_EDGES = 16  # accepted, no effect: the edge scale is part of CLIP_NORMALIZE
CLIP_UNIT_FRAME = 32
CLIP_UNION = 64
WRITE_POLY, WRITE_WKT = 0, 1
NOALLOC_SORT_MAX = 1 << 21


def hip_lib():
    """The HIP library.  Raises when it is missing: there is no CPU path."""
    global _hip
    if _hip is None:
        path = os.path.join(HERE, "libgcmf_hip.so")
        if not os.path.exists(path):
            raise GcmfError("libgcmf_hip.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.gcmf_device_count.restype = C.c_int
        L.gcmf_init.argtypes = [C.c_int, C.POINTER(vp)]
        L.gcmf_shutdown.argtypes = [vp]
        L.gcmf_last_error.argtypes = [vp]
        L.gcmf_last_error.restype = C.c_char_p
        L.gcmf_malloc.argtypes = [vp, C.POINTER(vp), C.c_size_t]
        L.gcmf_free.argtypes = [vp, vp]
        L.gcmf_memcpy_h2d.argtypes = [vp, vp, vp, C.c_size_t]
        L.gcmf_memcpy_d2h.argtypes = [vp, vp, vp, C.c_size_t]
        L.gcmf_synchronize.argtypes = [vp]
        L.gcmf_mem_info.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                    C.POINTER(C.c_size_t)]
        L.gcmf_set_timing.argtypes = [vp, C.c_int]
        L.gcmf_set_timing_only.argtypes = [vp, C.c_char_p]
        L.gcmf_timer_reset.argtypes = [vp]
        L.gcmf_timer_get.argtypes = [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_long)]
        L.gcmf_join.argtypes = [vp, C.POINTER(DLayer), C.POINTER(DLayer), C.POINTER(JoinStats)]
        L.gcmf_join_get.argtypes = [vp, C.c_int, vp, C.c_long]
        L.gcmf_clip.argtypes = [vp, C.POINTER(DLayer), C.POINTER(DLayer), vp, C.c_long, C.c_int,
                                C.POINTER(ClipStats)]
        L.gcmf_join_clip.argtypes = [vp, C.POINTER(DLayer), C.POINTER(DLayer), C.c_int,
                                     C.POINTER(JoinStats), C.POINTER(ClipStats)]
        L.gcmf_clip_capture.argtypes = [vp, C.POINTER(DLayer), C.POINTER(DLayer), vp, C.c_long,
                                        C.c_int, C.POINTER(ClipStats)]
        L.gcmf_join_clip_capture.argtypes = [vp, C.POINTER(DLayer), C.POINTER(DLayer), C.c_int,
                                             C.POINTER(JoinStats), C.POINTER(ClipStats)]
        L.gcmf_graph_launch.argtypes = [vp, C.POINTER(C.c_int)]
        L.gcmf_graph_destroy.argtypes = [vp]
        L.gcmf_clip_get.argtypes = [vp, C.c_int, vp, C.c_long]
        L.gcmf_clip_areas.argtypes = [vp, vp]
        L.gcmf_clip_write.argtypes = [vp, C.c_char_p, C.c_int]
        L.gcmf_radix_sort_u64.argtypes = [vp, vp, vp, vp, vp, C.c_long, C.c_int]
        L.gcmf_radix_sort_u64_noalloc.argtypes = [vp, vp, vp, vp, vp, C.c_long, C.c_int]
        L.gcmf_scan_i32_i64.argtypes = [vp, vp, vp, C.c_long]
        L.gcmf_compact_flags.argtypes = [vp, vp, C.c_long, vp, C.POINTER(C.c_long)]
        L.gcmf_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.gcmf_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.gcmf_tune_clip.argtypes = [vp, C.POINTER(DLayer), C.POINTER(DLayer), vp, C.c_long,
                                     C.c_int, C.POINTER(Tuning)]
        L.gcmf_clip_last_stats.argtypes = [vp, C.POINTER(ClipStats)]
        L.gcmf_clip_complex.argtypes = [vp, C.POINTER(DLayer), C.POINTER(DLayer), vp, C.c_long,
                                        C.c_int, C.POINTER(ClipStats)]
        _ref_prototypes(L)
        _hip = L
    return _hip


def _ref_prototypes(L):
    """ctypes prototypes of the reference-named entry points (include/gcmf_hip.h),
    as a caller of Join.h / ghcuda.h / main.h / ST_Intersect.cu binds them."""
    vp, lp, ip, dp = C.c_void_p, C.c_long, C.POINTER(C.c_int), C.POINTER(C.c_double)
    P = C.POINTER
    L.SortBaseMBROverlap.argtypes = [lp, lp, vp, vp, P(vp), P(vp), C.c_char, C.c_char]
    L.SortBaseMBROverlap.restype = lp
    L.GetCMBR.argtypes = [lp, vp, vp, vp, P(vp), P(vp), P(vp), P(vp), P(vp), P(lp)]
    L.GetCMBR.restype = None
    L.PointInPolygonTest.argtypes = [vp, vp, lp, lp, vp, vp, vp, vp, vp, vp, vp]
    L.PointInPolygonTest.restype = lp
    L.CountCMF.argtypes = [vp, vp, lp, vp, vp, vp, vp, vp, vp, vp, P(vp), P(vp), P(vp), P(vp),
                           P(vp), P(vp), P(lp), P(lp)]
    L.CountCMF.restype = None
    L.ApplyCMF.argtypes = [vp, vp, lp, vp, lp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.ApplyCMF.restype = None
    L.SegmentIntersectJoin.argtypes = [vp, vp, lp, vp, vp, vp, vp, vp, vp, vp, vp, P(vp)]
    L.SegmentIntersectJoin.restype = lp
    L.calculateIntersections.argtypes = [vp, vp, C.c_int, C.c_int, dp, vp, vp, vp, vp, C.c_int,
                                         C.c_int, ip, ip, P(dp), P(dp), P(ip), P(ip), P(ip),
                                         P(ip), P(ip), P(ip), ip]
    L.calculateIntersections.restype = None
    L.GH_CUDA.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int]
    L.GH_CUDA.restype = C.c_int
    L.ghcuda.argtypes = [ip, ip, C.c_int, vp, vp, vp, vp, vp, vp]
    L.ghcuda.restype = C.c_int
    L.ST_Intersect.argtypes = [lp, lp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.ST_Intersect.restype = C.c_int
    L.gcmf_st_intersect_pairs.argtypes = [lp, lp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, P(ip)]
    L.gcmf_st_intersect_pairs.restype = C.c_int
    L.spatialJoin.argtypes = [C.c_int, P(C.c_char_p), P(ip), P(ip), P(lp), P(dp), P(dp), P(ip),
                              P(P(C.c_long)), P(ip), P(P(C.c_long))]
    L.spatialJoin.restype = C.c_int
    L.gcmf_ref_context.argtypes = []
    L.gcmf_ref_context.restype = vp
    L.gcmf_ref_release.argtypes = []
    L.gcmf_ref_release.restype = None


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.ptr = C.c_void_p()
        ctx._check(hip_lib().gcmf_malloc(ctx.h, C.byref(self.ptr), max(int(nbytes), 1)), "malloc")
        self.nbytes = nbytes

    def free(self):
        if self.ptr:
            hip_lib().gcmf_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()


class DeviceLayer:
    """A layer resident in HBM (the reference's device layout)."""

    def __init__(self, ctx, layer):
        self.layer = layer
        self.bufs = [ctx.upload(a) for a in (layer.coords, layer.vnum, layer.vpsnum, layer.mbr,
                                             layer.mbr2)]
        self.desc = DLayer(layer.npoly, *[b.ptr for b in self.bufs])

    def free(self):
        for b in self.bufs:
            b.free()


class Context:
    """A GPU context of libgcmf_hip.so: one HIP stream and its workspace."""

    def __init__(self, device=0):
        L = hip_lib()
        if L.gcmf_device_count() <= 0:
            raise GcmfError("no HIP device visible")
        h = C.c_void_p()
        rc = L.gcmf_init(device, C.byref(h))
        if rc != 0:
            raise GcmfError("gcmf_init failed (%d)" % rc)
        self.h = h

    def _check(self, rc, what):
        if rc != 0:
            msg = hip_lib().gcmf_last_error(self.h)
            raise GcmfError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def close(self):
        if self.h:
            hip_lib().gcmf_shutdown(self.h)
            self.h = None

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        buf = DeviceBuffer(self, arr.nbytes)
        if arr.nbytes:
            self._check(hip_lib().gcmf_memcpy_h2d(self.h, buf.ptr, arr.ctypes.data, arr.nbytes),
                        "h2d")
        return buf

    def download(self, buf, dtype, count):
        out = np.zeros(count, dtype=dtype)
        if out.nbytes:
            self._check(hip_lib().gcmf_memcpy_d2h(self.h, out.ctypes.data, buf.ptr, out.nbytes),
                        "d2h")
        return out

    def synchronize(self):
        self._check(hip_lib().gcmf_synchronize(self.h), "synchronize")

    def mem_info(self):
        """(free, total) device bytes of the context's GPU and the bytes its
        grow-only workspace holds."""
        f, t, w = C.c_size_t(), C.c_size_t(), C.c_size_t()
        self._check(hip_lib().gcmf_mem_info(self.h, C.byref(f), C.byref(t), C.byref(w)),
                    "mem_info")
        return f.value, t.value, w.value

    def set_timing(self, on, only=None):
        """Kernel timing on/off; `only` restricts it to one kernel name."""
        L = hip_lib()
        self._check(L.gcmf_set_timing_only(self.h, only.encode() if only else None),
                    "set_timing_only")
        self._check(L.gcmf_set_timing(self.h, int(on)), "set_timing")

    def timer_reset(self):
        self._check(hip_lib().gcmf_timer_reset(self.h), "timer_reset")

    def timer(self, kernel):
        ms, n = C.c_double(), C.c_long()
        self._check(hip_lib().gcmf_timer_get(self.h, kernel.encode(), C.byref(ms), C.byref(n)),
                    "timer_get")
        return ms.value, n.value

    # ------------------------------------------------------------ tuning
    def tuning(self):
        t = Tuning()
        self._check(hip_lib().gcmf_get_tuning(self.h, C.byref(t)), "get_tuning")
        return t.as_dict()

    def set_tuning(self, cull_min, big_nodes, count_block, lds_max_q=None, label_lds=None,
                   grid_nodes=None, grid_clean=None, strip_min=None, cmf_group=None,
                   label_steal=None, block_cap=None, count_stage=None):
        cur = self.tuning()
        if lds_max_q is None:
            lds_max_q = cur["lds_max_q"]
        if label_lds is None:
            label_lds = cur["label_lds"]
        if grid_nodes is None:
            grid_nodes = cur["grid_nodes"]
        if grid_clean is None:
            grid_clean = cur["grid_clean"]
        if strip_min is None:
            strip_min = cur["strip_min"]
        if cmf_group is None:
            cmf_group = cur["cmf_group"]
        if label_steal is None:
            label_steal = cur["label_steal"]
        if block_cap is None:
            block_cap = cur["block_cap"]
        if count_stage is None:
            count_stage = cur["count_stage"]
        t = Tuning(cull_min, big_nodes, count_block, lds_max_q, label_lds, grid_nodes,
                   grid_clean, strip_min, cmf_group, label_steal, block_cap, count_stage)
        self._check(hip_lib().gcmf_set_tuning(self.h, C.byref(t)), "set_tuning")

    def tune_clip(self, base, over, d_pairs, M, flags=CLIP_NORMALIZE):
        """Autotunes the batched clip's shapes on one batch (device pairs
        buffer `d_pairs` of M pairs); returns the chosen values."""
        t = Tuning()
        self._check(hip_lib().gcmf_tune_clip(self.h, C.byref(base.desc), C.byref(over.desc),
                                             d_pairs.ptr, int(M), int(flags), C.byref(t)),
                    "tune_clip")
        return t.as_dict()

    # ------------------------------------------------------- primitives
    def radix_sort(self, d_keys_in, d_vals_in, d_keys_out, d_vals_out, n, bits):
        """Stable sort of n u64 keys (device buffers) by their low `bits` bits."""
        self._check(hip_lib().gcmf_radix_sort_u64(self.h, d_keys_in.ptr, d_vals_in.ptr,
                                                  d_keys_out.ptr, d_vals_out.ptr, int(n),
                                                  int(bits)), "radix_sort")

    def radix_sort_noalloc(self, d_keys_in, d_vals_in, d_keys_out, d_vals_out, n, bits):
        """The same sort with its scratch in a static device array: allocates
        nothing, n <= NOALLOC_SORT_MAX, stream-ordered (not synchronised)."""
        self._check(hip_lib().gcmf_radix_sort_u64_noalloc(self.h, d_keys_in.ptr, d_vals_in.ptr,
                                                          d_keys_out.ptr, d_vals_out.ptr, int(n),
                                                          int(bits)), "radix_sort_noalloc")

    def scan(self, d_in, d_out, n):
        """Exclusive scan of n i32 into n + 1 i64 (device buffers)."""
        self._check(hip_lib().gcmf_scan_i32_i64(self.h, d_in.ptr, d_out.ptr, int(n)), "scan")

    def compact(self, d_flags, d_idx, n):
        """Indices of the nonzero flags; returns their number."""
        cnt = C.c_long()
        self._check(hip_lib().gcmf_compact_flags(self.h, d_flags.ptr, int(n), d_idx.ptr,
                                                 C.byref(cnt)), "compact")
        return cnt.value

    # ------------------------------------------------------------- join
    def join(self, base, over):
        st = JoinStats()
        self._check(hip_lib().gcmf_join(self.h, C.byref(base.desc), C.byref(over.desc),
                                        C.byref(st)), "gcmf_join")
        return st.as_dict()

    def join_get(self, field, count):
        fid, dt = J_FIELDS[field]
        out = np.zeros(int(count), dtype=dt)
        if count:
            self._check(hip_lib().gcmf_join_get(self.h, fid, out.ctypes.data, int(count)),
                        "join_get")
        return out

    # ------------------------------------------------------------- clip
    def clip(self, base, over, pairs, flags=CLIP_NORMALIZE):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        dp = self.upload(pairs)
        st = ClipStats()
        try:
            self._check(hip_lib().gcmf_clip(self.h, C.byref(base.desc), C.byref(over.desc),
                                            dp.ptr, len(pairs) // 2, flags, C.byref(st)),
                        "gcmf_clip")
        finally:
            dp.free()
        return st.as_dict()

    def clip_complex(self, base, over, parts, flags=CLIP_NORMALIZE):
        """Complex polygons: parts = (M, 4) of (first base ring, base ring
        count, first overlay ring, overlay ring count); CLIP_UNION for the
        union.  Read the results with clip_results."""
        parts = np.ascontiguousarray(parts, dtype=np.int32).reshape(-1)
        dp = self.upload(parts)
        st = ClipStats()
        try:
            self._check(hip_lib().gcmf_clip_complex(self.h, C.byref(base.desc),
                                                    C.byref(over.desc), dp.ptr,
                                                    len(parts) // 4, flags, C.byref(st)),
                        "gcmf_clip_complex")
        finally:
            dp.free()
        return st.as_dict()

    def join_clip(self, base, over, flags=CLIP_NORMALIZE):
        js, cs = JoinStats(), ClipStats()
        self._check(hip_lib().gcmf_join_clip(self.h, C.byref(base.desc), C.byref(over.desc),
                                             flags, C.byref(js), C.byref(cs)), "gcmf_join_clip")
        return js.as_dict(), cs.as_dict()

    # ------------------------------------------------------ graph capture
    def clip_capture(self, base, over, d_pairs, M, flags=CLIP_NORMALIZE):
        """Clips the batch eagerly (device pairs buffer d_pairs of M pairs)
        and captures the launches as a graph replayed by graph_launch;
        returns the eager run's stats."""
        st = ClipStats()
        self._check(hip_lib().gcmf_clip_capture(self.h, C.byref(base.desc), C.byref(over.desc),
                                                d_pairs.ptr, int(M), flags, C.byref(st)),
                    "gcmf_clip_capture")
        return st.as_dict()

    def join_clip_capture(self, base, over, flags=CLIP_NORMALIZE):
        js, cs = JoinStats(), ClipStats()
        self._check(hip_lib().gcmf_join_clip_capture(self.h, C.byref(base.desc),
                                                     C.byref(over.desc), flags, C.byref(js),
                                                     C.byref(cs)), "gcmf_join_clip_capture")
        return js.as_dict(), cs.as_dict()

    def graph_launch(self, wait=True):
        """Replays the captured graph; with wait, returns whether every size
        matched the captured one (the results are valid)."""
        if not wait:
            self._check(hip_lib().gcmf_graph_launch(self.h, None), "gcmf_graph_launch")
            return None
        v = C.c_int(0)
        self._check(hip_lib().gcmf_graph_launch(self.h, C.byref(v)), "gcmf_graph_launch")
        return bool(v.value)

    def graph_destroy(self):
        self._check(hip_lib().gcmf_graph_destroy(self.h), "gcmf_graph_destroy")

    def clip_get(self, field, count):
        fid, dt = C_FIELDS[field]
        out = np.zeros(int(count), dtype=dt)
        if count:
            self._check(hip_lib().gcmf_clip_get(self.h, fid, out.ctypes.data, int(count)),
                        "clip_get")
        return out

    def clip_results(self, stats):
        """Per-pair result components (list of lists of (n,2) arrays) + status."""
        M = stats["pairs"]
        status = self.clip_get("status", M)
        comp_off = self.clip_get("comp_off", M + 1)
        sizes = self.clip_get("comp_sizes", stats["total_comps"])
        xy = self.clip_get("xy", 2 * stats["total_verts"]).reshape(-1, 2)
        out = []
        v = 0
        for k in range(M):
            comps = []
            for c in range(int(comp_off[k]), int(comp_off[k + 1])):
                n = int(sizes[c])
                comps.append(xy[v:v + n])
                v += n
            out.append(comps)
        return status, out


    def clip_areas(self, stats):
        """Signed area of every component of the last clip, computed on the
        device (counter-clockwise positive)."""
        n = int(stats["total_comps"])
        d = DeviceBuffer(self, 8 * max(n, 1))
        try:
            self._check(hip_lib().gcmf_clip_areas(self.h, d.ptr), "clip_areas")
            return self.download(d, np.float64, n)
        finally:
            d.free()

    def clip_write(self, path, fmt=WRITE_POLY):
        """Writes the last clip's results (.poly or one WKT line per pair)."""
        self._check(hip_lib().gcmf_clip_write(self.h, path.encode(), int(fmt)), "clip_write")


def _results_arrays(results):
    """(xy, comp_sizes, ncomp per pair) of per-pair lists of (n, 2) arrays."""
    comps = [c for r in results for c in r]
    xy = (np.concatenate([np.asarray(c, np.float64).reshape(-1) for c in comps])
          if comps else np.zeros(0))
    sizes = np.array([len(c) for c in comps], np.int32)
    ncomp = np.array([len(r) for r in results], np.int32)
    return np.ascontiguousarray(xy), sizes, ncomp


def ring_area(ring):
    """Signed area of one ring (gcmf_ring_area)."""
    r = np.ascontiguousarray(ring, np.float64)
    return host_lib().gcmf_ring_area(r, len(r))


def result_areas(results):
    """gcmf_result_areas over per-pair lists of components."""
    xy, sizes, _ = _results_arrays(results)
    out = np.zeros(len(sizes))
    rc = host_lib().gcmf_result_areas(xy, sizes, len(sizes), out)
    if rc != 0:
        raise GcmfError("gcmf_result_areas failed")
    return out


def result_write(path, results, status, fmt=WRITE_POLY, nest=False):
    """Writes per-pair lists of components (gcmf_result_write_poly / _wkt)."""
    xy, sizes, ncomp = _results_arrays(results)
    st = np.ascontiguousarray(status, np.int32)
    L = host_lib()
    if fmt == WRITE_POLY:
        rc = L.gcmf_result_write_poly(path.encode(), xy, sizes, len(sizes))
    else:
        rc = L.gcmf_result_write_wkt(path.encode(), xy, sizes, ncomp, st, len(ncomp),
                                     1 if nest else 0)
    if rc != 0:
        raise GcmfError("cannot write " + path)


class RefContext(Context):
    """The process-default context of the reference-named entry points
    (gcmf_ref_context): reads the device results of their last calls."""

    def __init__(self):
        h = hip_lib().gcmf_ref_context()
        if not h:
            raise GcmfError("no default context (no HIP device)")
        self.h = C.c_void_p(h)

    def close(self):
        self.h = None

    def last_clip_stats(self):
        st = ClipStats()
        self._check(hip_lib().gcmf_clip_last_stats(self.h, C.byref(st)), "clip_last_stats")
        return st.as_dict()
