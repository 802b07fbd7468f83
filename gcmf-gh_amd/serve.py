"""Request loop over the device-resident join-and-clip path.

Clip requests (one polygon pair each, taken from a pair pool resident in
HBM) and join requests (two layers resident in HBM) arrive over time.  Clip
requests are continuously batched: whenever the previous batch has finished,
every request that has arrived since is admitted, in arrival order, while the
free-HBM budget allows, and the admitted pairs are clipped by one batched
gcmf_clip call.  Join requests run one at a time on a second context (its own
HIP stream) in a thread of their own, so that joins and clip batches overlap
on the device.  Each request's latency runs from its arrival to the moment
its result is ready on the host side of the loop.

Admission is by measured device memory.  The budget is the free HBM that
hipMemGetInfo reports once the pools are resident (gcmf_mem_info; at most
`hbm_budget`), less a reserve.  A request is charged the device memory it
needs while it is admitted and released when it completes: bytes per vertex
of its pair (clips) or per polygon of its layers (joins), both measured
from the workspace growth of calibration runs on the server's own contexts
(the largest pool pairs clipped as one batch; every layer set joined once).
A request that does not fit waits for the next batch.

The reference runs pairs one by one (ghcuda.cpp:385-425) and has no serving
loop; this module has no reference counterpart.
"""
import ctypes as C
import hashlib
import threading
import time

import numpy as np

import gcmf_amd as G

# floors of the per-vertex / per-polygon charges (the measured values are
# used when larger), the share of free HBM left unbudgeted, and the number
# of largest pool pairs in the clip calibration batch
CLIP_BYTES_PER_VERTEX = 256
JOIN_BYTES_PER_POLYGON = 256
HBM_RESERVE = 0.1
CALIBRATION_PAIRS = 256


class Request:
    __slots__ = ("rid", "kind", "arrival", "key", "cost", "t_start", "t_done", "batch",
                 "status", "summary", "digest", "result")

    def __init__(self, rid, kind, arrival, key):
        self.rid, self.kind, self.arrival, self.key = rid, kind, arrival, key
        self.cost = 0
        self.t_start = self.t_done = None
        self.batch = -1
        self.status = None
        self.summary = None
        self.digest = None
        self.result = None

    @property
    def latency(self):
        return self.t_done - self.arrival


def poisson_requests(rate, duration, clip_fraction, n_pairs, n_join_sets, seed=0):
    """Arrivals of a Poisson process of `rate` requests/s over `duration` s;
    each request is a clip of a uniformly drawn pool pair with probability
    `clip_fraction`, else a join of a uniformly drawn layer set."""
    rs = np.random.RandomState(seed)
    t, out = 0.0, []
    while True:
        t += rs.exponential(1.0 / rate)
        if t >= duration:
            break
        if rs.rand() < clip_fraction or n_join_sets == 0:
            out.append(Request(len(out), "clip", t, int(rs.randint(n_pairs))))
        else:
            out.append(Request(len(out), "join", t, int(rs.randint(n_join_sets))))
    return out


def _digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


class Server:
    """pair_P / pair_Q: the clip pool (pair k = (pair_P ring k, pair_Q ring
    k)); join_sets: list of (base layer, overlay layer)."""

    def __init__(self, pair_P, pair_Q, join_sets=(), device=0, hbm_budget=200e9,
                 max_batch=65536, flags=G.CLIP_NORMALIZE, keep_results=False):
        self.cctx = G.Context(device)
        self.jctx = G.Context(device)
        self.dP = G.DeviceLayer(self.cctx, pair_P)
        self.dQ = G.DeviceLayer(self.cctx, pair_Q)
        self.joins = [(G.DeviceLayer(self.jctx, b), G.DeviceLayer(self.jctx, o))
                      for b, o in join_sets]
        self.max_batch = int(max_batch)
        self.flags = flags
        self.keep = keep_results
        self.d_pairs = self.cctx.upload(np.zeros(2 * self.max_batch, np.int32))
        self.h_pairs = np.zeros(2 * self.max_batch, np.int32)
        self.pair_verts = (pair_P.vnum.astype(np.int64) - 1) + (pair_Q.vnum.astype(np.int64) - 1)
        self.join_polys = [b.npoly + o.npoly for b, o in join_sets]
        self._calibrate()
        self.pair_cost = self.clip_bytes_per_vertex * self.pair_verts
        self.join_cost = [self.join_bytes_per_polygon * n for n in self.join_polys]
        free, total, _ = self.cctx.mem_info()
        self.hbm_free = free
        self.budget = min(float(hbm_budget), (1.0 - HBM_RESERVE) * free)
        self._lock = threading.Lock()
        self._inflight = 0.0
        self.batches = []     # (t_start, t_done, size)
        self.deferred = 0     # admissions put off by the budget
        self.join_error = None

    def close(self):
        self.d_pairs.free()
        self.dP.free()
        self.dQ.free()
        for b, o in self.joins:
            b.free()
            o.free()
        self.cctx.close()
        self.jctx.close()

    # ------------------------------------------------------- calibration
    def _calibrate(self):
        """Measured charges: the clip context's workspace growth over one
        batch of the largest pool pairs, per vertex, and the join context's
        over one join of every layer set, per polygon of the largest set."""
        n = len(self.pair_verts)
        k = min(CALIBRATION_PAIRS, n, self.max_batch)
        keys = np.argsort(-self.pair_verts, kind="stable")[:k].astype(np.int32)
        w0 = self.cctx.mem_info()[2]
        self.cctx.clip(self.dP, self.dQ, np.stack([keys, keys], axis=1), self.flags)
        w1 = self.cctx.mem_info()[2]
        self.clip_bytes_per_vertex = max(CLIP_BYTES_PER_VERTEX,
                                         int(np.ceil(w1 / max(1, int(self.pair_verts[keys].sum())))))
        self.calibration_clip_bytes = w1 - w0
        w0 = self.jctx.mem_info()[2]
        for b, o in self.joins:
            self.jctx.join(b, o)
        w1 = self.jctx.mem_info()[2]
        big = max(self.join_polys) if self.join_polys else 1
        self.join_bytes_per_polygon = max(JOIN_BYTES_PER_POLYGON, int(np.ceil(w1 / big)))
        self.calibration_join_bytes = w1 - w0

    # ------------------------------------------------------------ budget
    def _admit(self, cost):
        with self._lock:
            if self._inflight + cost > self.budget and self._inflight > 0:
                return False
            self._inflight += cost
            return True

    def _release(self, cost):
        with self._lock:
            self._inflight -= cost

    # --------------------------------------------------------------- run
    def run(self, requests):
        """Serves `requests` (arrival times in seconds from now); returns them
        with t_start / t_done / status / summary filled."""
        clips = sorted((r for r in requests if r.kind == "clip"), key=lambda r: r.arrival)
        joins = sorted((r for r in requests if r.kind == "join"), key=lambda r: r.arrival)
        for r in clips:
            r.cost = float(self.pair_cost[r.key])
        for r in joins:
            r.cost = float(self.join_cost[r.key])
        t0 = time.perf_counter()
        jt = threading.Thread(target=self._join_loop, args=(joins, t0))
        jt.start()
        try:
            self._clip_loop(clips, t0)
        finally:
            jt.join()
        if self.join_error is not None:
            raise self.join_error
        return requests

    def _clip_loop(self, reqs, t0):
        L = G.hip_lib()
        i, n = 0, len(reqs)
        while i < n:
            now = time.perf_counter() - t0
            if reqs[i].arrival > now:
                time.sleep(min(reqs[i].arrival - now, 0.0005))
                continue
            batch = []
            while i < n and reqs[i].arrival <= now and len(batch) < self.max_batch:
                if not self._admit(reqs[i].cost):
                    self.deferred += 1
                    break
                batch.append(reqs[i])
                i += 1
            if not batch:
                time.sleep(0.0002)
                continue
            ts = time.perf_counter() - t0
            M = len(batch)
            keys = np.fromiter((r.key for r in batch), np.int32, M)
            self.h_pairs[0:2 * M:2] = keys
            self.h_pairs[1:2 * M:2] = keys
            self.cctx._check(L.gcmf_memcpy_h2d(self.cctx.h, self.d_pairs.ptr,
                                               self.h_pairs.ctypes.data, 8 * M), "h2d")
            st = G.ClipStats()
            self.cctx._check(L.gcmf_clip(self.cctx.h, C.byref(self.dP.desc),
                                         C.byref(self.dQ.desc), self.d_pairs.ptr, M, self.flags,
                                         C.byref(st)), "gcmf_clip")
            stats = st.as_dict()
            status = self.cctx.clip_get("status", M)
            nverts = self.cctx.clip_get("nverts", M)
            ncomp = self.cctx.clip_get("ncomp", M)
            results = self.cctx.clip_results(stats)[1] if self.keep else None
            td = time.perf_counter() - t0
            self.batches.append((ts, td, M))
            for j, r in enumerate(batch):
                r.t_start, r.t_done, r.batch = ts, td, len(self.batches) - 1
                r.status = int(status[j])
                r.summary = (int(ncomp[j]), int(nverts[j]))
                if results is not None:
                    comps = results[j]
                    r.digest = _digest(np.array([len(c) for c in comps], np.int64),
                                       *[np.asarray(c, np.float64) for c in comps])
                    r.result = comps
                self._release(r.cost)

    def _join_loop(self, reqs, t0):
        try:
            for r in reqs:
                while True:
                    now = time.perf_counter() - t0
                    if r.arrival > now:
                        time.sleep(min(r.arrival - now, 0.0005))
                        continue
                    if self._admit(r.cost):
                        break
                    self.deferred += 1
                    time.sleep(0.0002)
                r.t_start = time.perf_counter() - t0
                b, o = self.joins[r.key]
                js = self.jctx.join(b, o)
                r.status = 0
                r.summary = (js["candidates"], js["joined"])
                if self.keep:
                    cand = self.jctx.join_get("pairs", 2 * js["candidates"])
                    joined = self.jctx.join_get("joined", 2 * js["joined"])
                    r.digest = _digest(cand, joined)
                    r.result = (cand.reshape(-1, 2), joined.reshape(-1, 2))
                r.t_done = time.perf_counter() - t0
                self._release(r.cost)
        except Exception as e:  # surfaced by run()
            self.join_error = e


def clip_alone(ctx, dP, dQ, key, flags=G.CLIP_NORMALIZE):
    """One pool pair clipped by itself: (status, summary, digest) as the
    server records them."""
    st = ctx.clip(dP, dQ, np.array([[key, key]], np.int32), flags)
    status, res = ctx.clip_results(st)
    comps = res[0]
    return (int(status[0]), (len(comps), sum(len(c) for c in comps)),
            _digest(np.array([len(c) for c in comps], np.int64),
                    *[np.asarray(c, np.float64) for c in comps]))


def join_alone(ctx, db, do):
    js = ctx.join(db, do)
    return (0, (js["candidates"], js["joined"]),
            _digest(ctx.join_get("pairs", 2 * js["candidates"]),
                    ctx.join_get("joined", 2 * js["joined"])))


def latency_summary(reqs):
    out = {}
    for kind in ("clip", "join", "all"):
        lat = [r.latency for r in reqs if r.t_done is not None and (kind == "all" or
                                                                  r.kind == kind)]
        if lat:
            out[kind] = {"n": len(lat),
                         "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
                         "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3)}
    return out
