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
