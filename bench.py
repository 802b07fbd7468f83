"""Benchmark of the join-and-clip hot path on one MI355X.

Workload: two layers of 20,000 star-shaped polygons each in the unit square
(centres uniform, radii log-uniform in [1e-4, 5e-3], 16-512 vertices, FP64,
%.15f text round trip for the integer MBR keys, seed 0), resident in HBM.
One step = the whole path over that input: MBR sort-and-sweep filter ->
CMBR/PiP refine -> CMF edge filter -> segment-intersection refine ->
batched Greiner-Hormann clip of every joined pair.

value = clipped polygon pairs per second over the timed steps (inputs
already on the device).  Join candidates per second and per-step p50/p99
latency are reported beside it.  The two largest kernels' roofline
fractions are measured live with HIP events on the library's stream; the
CPU baseline is the repository's C restatement (the oracle) over the whole
workload, on one core and on a pool of the host's cores, and its outputs
check the GPU's bit for bit.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gcmf-gh_amd"))

import numpy as np  # noqa: E402

METRIC = "clipped polygon pairs/s + join candidates/s at ≤1e-9 err; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
# FP64 work of the count kernel per edge pair it classifies: four
# orientation determinants (7 flops each), two subtractions and two
# divisions; per intersection point it emits, 7 flops for (1-a)P1 + aP2
FLOPS_PER_EVAL = 32
FLOPS_PER_POINT = 7
# the two largest kernels of the clip, timed live with HIP events inside the
# timed region
COUNT_KERNEL = "gh_count"
LABEL_KERNEL = "gh_label_trace"
LIVE_KERNELS = COUNT_KERNEL + "," + LABEL_KERNEL
BREAKDOWN_STEPS = 3
PROFILES = os.path.join(ROOT, "profiles")


def traffic_bytes(kernel, workload, size):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (each scaled by a plain copy's reading) for this workload at
    this size, from profiles/<kernel>_traffic.json; None otherwise."""
    fn = os.path.join(PROFILES, "%s_traffic.json" % kernel)
    if not os.path.exists(fn):
        return None
    with open(fn) as f:
        t = json.load(f)
    rec = t.get(workload)
    if not rec or rec.get("size") != size:
        return None
    return rec["bytes_per_launch"]


def count_roofline(total_ms, launches, cs, traffic):
    """The count kernel (k_count_cpu; its rings are all within the culling
    threshold here, so the LDS-binned pass does not launch) priced as FP64 vector
    work: 32 flops per edge pair it classifies plus 7 per intersection
    point, against MI355X's 78.6 TFLOP/s FP64 peak.  Box rejects are not
    counted: they are the kernel's overhead, not the algorithm's work."""
    avg_ms = total_ms / max(launches, 1)
    flops = FLOPS_PER_EVAL * cs["n_eval"] + FLOPS_PER_POINT * cs["n_intersections"]
    achieved = flops / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    return {"bound": "fp64-valu", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFS,
            "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFS, 6), "traffic": traffic,
            "kernel": "k_count_cpu", "avg_launch_ms": round(avg_ms, 5), "launches": launches,
            "work": "%d classified edge pairs x %d + %d intersection points x %d FP64 flops; "
                    "%d edge pairs visited" % (cs["n_eval"], FLOPS_PER_EVAL,
                                                cs["n_intersections"], FLOPS_PER_POINT,
                                                cs["n_tested"])}


def label_bytes(cs):
    """Algorithmic HBM bytes of labelling and tracing one batch: every slot
    (ring vertex or inserted intersection) read once -- its normalised point
    (16 B), original point (16 B) and neighbour link (4 B) --, every edge's
    first slot (8 B) to flag the source vertices, and the result written
    once (16 B per vertex, 4 B per component) with 8 B of counts per pair."""
    return (36 * (cs["slots_p"] + cs["slots_q"]) + 8 * (cs["edges_p"] + cs["edges_q"]) +
            16 * cs["total_verts"] + 4 * cs["total_comps"] + 8 * cs["pairs"])


def label_roofline(total_ms, launches, cs, traffic, kernel="k_label_wave"):
    avg_ms = total_ms / max(launches, 1)
    b = label_bytes(cs)
    achieved = b / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
            "kernel": kernel, "avg_launch_ms": round(avg_ms, 5), "launches": launches,
            "work": "%d algorithmic bytes per launch: 36 B x %d slots + 8 B x %d edges + "
                    "16 B x %d result vertices + 4 B x %d components + 8 B x %d pairs"
                    % (b, cs["slots_p"] + cs["slots_q"], cs["edges_p"] + cs["edges_q"],
                       cs["total_verts"], cs["total_comps"], cs["pairs"])}


# ------------------------------------------------ CPU path (the oracle) ----
_W = {}


def _oracle_worker_init(n, seed):
    """Pool worker: the same layers, regenerated from the seed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import gcmf_amd as G
    _W["layers"] = G.gen_join_layers(n, n, seed=seed)


def _clip_pairs(base, over, pairs):
    import oracle_py as O
    out = []
    for b, o in pairs:
        P, Q = base.ring(int(b)), over.ring(int(o))
        if len(P) < len(Q):
            P, Q = Q, P
        st, comps, _ = O.clip_cpu(P, Q, 5)  # the GCMF_CLIP_NORMALIZE frame
        out.append((st, comps))
    return out


def _digest(st, comps):
    import hashlib
    h = hashlib.sha1()
    for c in comps:
        h.update(np.ascontiguousarray(c, np.float64).tobytes())
    return st, tuple(len(c) for c in comps), h.hexdigest()


def _oracle_worker_clip(pairs):
    """Clips a chunk; returns digests, so that the timed region is the
    clipping and not the transfer of the results."""
    base, over = _W["layers"]
    return [_digest(st, comps) for st, comps in _clip_pairs(base, over, pairs)]


def cpu_reference(base, over, n, seed):
    """The CPU path (the repository's C restatement of the reference's CPU
    route) over the whole workload: the join (MBR sweep, CMBR, PiP, CMF,
    segment refine) and the clip of every joined pair, on one core and then
    with the clips spread over a pool of worker processes (the join stays on
    one core).  Returns the baseline record and the oracle's outputs."""
    import multiprocessing as mp
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    t0 = time.perf_counter()
    _, cand = O.mbr_sweep(base.mbr, over.mbr)
    ref = O.join_refine(cand, base.as_dict(), over.as_dict())
    joined = cand[(ref["pip_flag"] == 1) | (ref["seg_flag"] == 1)]
    t_join = time.perf_counter() - t0
    print("cpu: join %.2f s, %d joined pairs" % (t_join, len(joined)), file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    results = _clip_pairs(base, over, joined)
    t_clip1 = time.perf_counter() - t0
    print("cpu: clip on one core %.2f s" % t_clip1, file=sys.stderr, flush=True)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    W = max(1, min(16, cores))
    chunks = [joined[i::4 * W] for i in range(4 * W)]
    with mp.get_context("spawn").Pool(W, initializer=_oracle_worker_init,
                                       initargs=(n, seed)) as pool:
        pool.map(len, [[0]] * W)  # workers up
        t0 = time.perf_counter()
        parts = pool.map(_oracle_worker_clip, chunks)
        t_clipW = time.perf_counter() - t0
    # the pool's results are the single-core run's (interleaved back)
    same = True
    for i, part in enumerate(parts):
        for j, d in enumerate(part):
            same &= d == _digest(*results[i + 4 * W * j])
    print("cpu: clip on %d processes %.2f s" % (W, t_clipW), file=sys.stderr, flush=True)
    rec = {"value": round(len(joined) / (t_join + t_clipW), 2), "unit": "pairs/s", "cores": W,
           "kind": "port",
           "sample": "the whole workload, no sampling: the join of both layers (%.2f s, one "
                     "core) and the clip of all %d joined pairs (%.2f s over %d processes, "
                     "%.2f s on one core)" % (t_join, len(joined), t_clipW, W, t_clip1),
           "one_core": {"value": round(len(joined) / (t_join + t_clip1), 2), "unit": "pairs/s",
                        "cores": 1, "seconds": round(t_join + t_clip1, 3)},
           "pool_matches_one_core": bool(same)}
    return rec, cand, joined, results


def parity(ctx, js, cs, cand, joined, results):
    """The GPU outputs of the last timed step against the oracle's: the
    candidate pair list (order included), the joined list, and every
    clipped pair's status, component sizes and vertices, bit for bit."""
    g_cand = ctx.join_get("pairs", 2 * js["candidates"]).reshape(-1, 2)
    g_join = ctx.join_get("joined", 2 * js["joined"]).reshape(-1, 2)
    status, res = ctx.clip_results(cs)
    mism = 0
    if len(res) == len(results):
        for k, (st, comps) in enumerate(results):
            ok = int(status[k]) == st and len(res[k]) == len(comps) and all(
                np.array_equal(a, b) for a, b in zip(res[k], comps))
            mism += 0 if ok else 1
    else:
        mism = max(len(res), len(results))
    return {"candidates_equal": bool(np.array_equal(g_cand, cand)),
            "joined_equal": bool(np.array_equal(g_join, joined)),
            "pairs_compared": len(results), "pairs_mismatched": mism,
            "checked": "candidate list, joined list and every clipped pair (status, component "
                       "sizes, vertex coordinates) bit-exact against the oracle"}


def graph_replay(ctx, steps, capture, total_verts):
    """The same step captured as one HIP graph (gcmf_*_capture) and replayed
    `steps` times back to back; the last replay is waited for and its size
    checks read.  Every kernel runs in every replay; the result vertices of
    the last replay are compared with the eager run's, bit for bit."""
    xy0 = ctx.clip_get("xy", 2 * total_verts)
    capture()
    ctx.graph_launch()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps - 1):
        ctx.graph_launch(wait=False)
    valid = ctx.graph_launch()
    el = time.perf_counter() - t0
    same = bool(np.array_equal(ctx.clip_get("xy", 2 * total_verts), xy0))
    ctx.graph_destroy()
    return {"ms_per_step": round(1e3 * el / steps, 4), "replays": steps,
            "sizes_match_capture": valid, "results_equal_eager": same}


def copy_calibration():
    import torch
    a = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * a.element_size()}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=20000, help="polygons per layer")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the captured-graph replay after the timed steps (for profiles)")
    ap.add_argument("--workload", choices=["join-clip", "batched", "large", "serve", "prims"],
                    default="join-clip",
                    help="join-clip: the two-layer join + clip (default); batched: "
                         "independent uneven pairs clipped in one pass; large: one pair of "
                         "1e6 x 1e5 vertices; serve: Poisson stream of clip and join requests")
    ap.add_argument("--rate", type=float, default=2000.0, help="requests/s for --workload serve")
    ap.add_argument("--duration", type=float, default=60.0,
                    help="seconds of arrivals for --workload serve")
    ap.add_argument("--saturate", type=int, default=20000,
                    help="clip requests arriving at once after the stream (--workload serve): "
                         "the saturated served rate")
    ap.add_argument("--hbm-budget", type=float, default=200e9,
                    help="free-HBM admission budget in bytes for --workload serve")
    ap.add_argument("--large-n", type=int, nargs=2, default=[1000000, 100000],
                    help="ring sizes for --workload large")
    ap.add_argument("--snap", type=float, default=0.0,
                    help="grid for --workload large (1e-3: the degenerate variant)")
    ap.add_argument("--reference-eps", action="store_true",
                    help="--workload large in the unit CMBR frame (GCMF_CLIP_UNIT_FRAME: the "
                         "reference's absolute epsilon without the edge scale)")
    ap.add_argument("--pairs", type=int, default=200000, help="pairs for --workload batched")
    ap.add_argument("--no-tune", action="store_true",
                    help="batched: keep the default shapes instead of autotuning them")
    ap.add_argument("--per-pair-sample", type=int, default=2000,
                    help="pairs clipped one call each for the per-pair comparison (batched)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one plain 256 MiB device-to-device copy and exit (the reference "
                         "reading for the HBM byte counters FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return batched(args)
    if args.workload == "large":
        return large(args)
    if args.workload == "serve":
        return serve(args)
    if args.workload == "prims":
        return prims(args)

    import gcmf_amd as G
    print("generating layers ...", file=sys.stderr, flush=True)
    base, over = G.gen_join_layers(args.n, args.n, seed=0)
    cpu = None
    if not args.no_cpu:
        # before the GPU is touched: the pool's processes start clean
        cpu = cpu_reference(base, over, args.n, 0)
    ctx = G.Context(0)
    db, do = G.DeviceLayer(ctx, base), G.DeviceLayer(ctx, over)
    flags = G.CLIP_NORMALIZE
    for _ in range(args.warmup):
        js, cs = ctx.join_clip(db, do, flags)
    ctx.synchronize()
    # timed region: only the count and label kernels carry events
    ctx.timer_reset()
    ctx.set_timing(True, only=LIVE_KERNELS)
    lat = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        js, cs = ctx.join_clip(db, do, flags)
        lat.append(time.perf_counter() - s0)
        print("step %d: %.2f ms" % (i, 1e3 * lat[-1]), file=sys.stderr, flush=True)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_timing(False)
    cnt_ms, cnt_n = ctx.timer(COUNT_KERNEL)
    lab_ms, lab_n = ctx.timer(LABEL_KERNEL)
    check = parity(ctx, js, cs, *cpu[1:]) if cpu else None
    # per-kernel breakdown from separate, fully timed steps
    ctx.timer_reset()
    ctx.set_timing(True)
    for _ in range(BREAKDOWN_STEPS):
        ctx.join_clip(db, do, flags)
    ctx.synchronize()
    ctx.set_timing(False)

    kernels = ["mbr_endpoints", "sort_hist", "sort_scatter", "scan_tile", "scan_reduce",
               "scan_downsweep", "scan_lookback", "compact_lookback", "mbr_sweep_prep",
               "mbr_sweep_count", "mbr_sweep_write", "get_cmbr", "pip", "cmf_count", "cmf_apply",
               "segment_join", "gh_pair_setup", "gh_pair_fill", "gh_count", "gh_transpose",
               "gh_nmap", "gh_calc_own", "gh_calc", "gh_sort_p", "gh_sort_runs", "gh_sort_q",
               "gh_init_label", "gh_cpu_fill", "gh_cpu_sort", "gh_cpu_link", "gh_label_trace",
               "gh_gather_out", "fill_i32"]
    times = {k: ctx.timer(k) for k in kernels}
    # the step as one captured graph, replayed (reported beside the eager figure)
    graph = None if args.no_graph else graph_replay(
        ctx, args.steps, lambda: ctx.join_clip_capture(db, do, flags), cs["total_verts"])
    ms_step = 1e3 * elapsed / args.steps
    pairs = cs["pairs"]
    value = pairs / (elapsed / args.steps)
    size = args.n
    roofs = [count_roofline(cnt_ms, cnt_n, cs, traffic_bytes("k_count_cpu", "join-clip", size)),
             label_roofline(lab_ms, lab_n, cs, traffic_bytes("k_label_wave", "join-clip", size))]
    roofs.sort(key=lambda r: -r["avg_launch_ms"])  # the dominant kernel first
    line = {"metric": METRIC, "value": round(value, 2), "unit": "pairs/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded generator, no dataset)",
            "config": {"workload": "spatial join + GH clip of two layers of %d star polygons "
                                   "(unit square, radii logU[1e-4,5e-3], 16-512 vertices, "
                                   "FP64, seed 0)" % args.n,
                       "candidates": js["candidates"], "joined_pairs": pairs,
                       "within": js["within"], "intersecting": js["intersecting"],
                       "result_components": cs["total_comps"],
                       "result_vertices": cs["total_verts"], "failed_pairs": cs["failed"],
                       "invalid_pairs": cs["invalid"]},
            "join_candidates_per_s": round(js["candidates"] / (elapsed / args.steps), 2),
            "latency_ms_p50": round(1e3 * float(np.percentile(lat, 50)), 4),
            "latency_ms_p99": round(1e3 * float(np.percentile(lat, 99)), 4),
            "kernel_ms_per_step": {k: round(v[0] / BREAKDOWN_STEPS, 4)
                                   for k, v in times.items() if v[1]},
            "roofline": roofs[0], "roofline_next": roofs[1], "graph_replay": graph}
    if cpu:
        line["cpu_baseline"] = cpu[0]
        line["parity"] = check
    db.free()
    do.free()
    ctx.close()
    print(json.dumps(line))


def batched(args):
    """Independent pairs with log-normal vertex counts (median 64, sigma 1.5,
    clamped 8..50000), 70 % overlapping, all clipped in one batched pass."""
    import gcmf_amd as G
    print("generating %d pairs ..." % args.pairs, file=sys.stderr, flush=True)
    P, Q = G.gen_pair_layers(args.pairs, seed=0)
    ctx = G.Context(0)
    dP, dQ = G.DeviceLayer(ctx, P), G.DeviceLayer(ctx, Q)
    pairs = np.stack([np.arange(args.pairs), np.arange(args.pairs)], axis=1).astype(np.int32)
    d_pairs = ctx.upload(pairs)
    import ctypes as C
    L = G.hip_lib()

    def step():
        st = G.ClipStats()
        ctx._check(L.gcmf_clip(ctx.h, C.byref(dP.desc), C.byref(dQ.desc), d_pairs.ptr,
                               args.pairs, G.CLIP_NORMALIZE, C.byref(st)), "gcmf_clip")
        return st.as_dict()

    # launch shapes chosen by the autotuner on this batch (untimed)
    t0 = time.perf_counter()
    tuned = ctx.tune_clip(dP, dQ, d_pairs, args.pairs) if not args.no_tune else ctx.tuning()
    tune_s = time.perf_counter() - t0
    print("tuning: %s (%.1f s)" % (tuned, tune_s), file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        cs = step()
    ctx.timer_reset()
    ctx.set_timing(True, only=LIVE_KERNELS)
    lat = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        cs = step()
        lat.append(time.perf_counter() - s0)
        print("step %d: %.2f ms" % (i, 1e3 * lat[-1]), file=sys.stderr, flush=True)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_timing(False)
    cnt_ms, cnt_n = ctx.timer(COUNT_KERNEL)
    lab_ms, lab_n = ctx.timer(LABEL_KERNEL)
    ctx.timer_reset()
    ctx.set_timing(True)
    step()
    ctx.synchronize()
    ctx.set_timing(False)
    kernel_ms = {k: round(ctx.timer(k)[0], 4) for k in (
        "gh_pair_setup", "gh_pair_fill", "gh_count", "gh_transpose", "gh_nmap", "gh_calc_own",
        "gh_calc",
        "gh_sort_p", "gh_sort_runs", "gh_sort_q", "gh_init_label", "gh_cpu_fill",
        "gh_cpu_sort", "gh_cpu_link", "gh_label_trace", "gh_gather_out", "sort_hist",
        "sort_scatter")}
    # the plain per-pair path (one clip call per pair, as the reference's
    # ghcuda loop runs GH_CUDA per pair) on an evenly strided sample of the
    # same pairs; each pair's result must match its result in the batch
    nverts_batch = ctx.clip_get("nverts", args.pairs)
    status_batch = ctx.clip_get("status", args.pairs)
    # the batch as one captured graph, replayed
    graph = None if args.no_graph else graph_replay(
        ctx, args.steps, lambda: ctx.clip_capture(dP, dQ, d_pairs, args.pairs), cs["total_verts"])
    stride = max(1, args.pairs // max(1, args.per_pair_sample))
    sample = list(range(0, args.pairs, stride))[:args.per_pair_sample]
    mism = 0
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in sample:
        st = G.ClipStats()
        ctx._check(L.gcmf_clip(ctx.h, C.byref(dP.desc), C.byref(dQ.desc),
                               C.c_void_p(d_pairs.ptr.value + 8 * i), 1, G.CLIP_NORMALIZE,
                               C.byref(st)), "gcmf_clip")
        s1 = int(ctx.clip_get("status", 1)[0])
        if s1 != int(status_batch[i]) or (s1 == 0 and st.total_verts != int(nverts_batch[i])):
            mism += 1
    ctx.synchronize()
    t_pp = time.perf_counter() - t0
    per_pair_rate = len(sample) / t_pp
    print("per-pair path: %d pairs in %.2f s, %d mismatches" % (len(sample), t_pp, mism),
          file=sys.stderr, flush=True)
    kernels = ["gh_pair_setup", "gh_pair_fill", "gh_count", "gh_transpose", "gh_nmap",
               "gh_calc_own", "gh_calc", "gh_sort_p", "gh_sort_runs", "gh_sort_q", "gh_init_label",
               "gh_cpu_fill", "gh_cpu_sort", "gh_cpu_link",
               "gh_label_trace", "gh_gather_out", "sort_hist", "sort_scatter"]
    line = {"metric": METRIC, "value": round(args.pairs / (elapsed / args.steps), 2),
            "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded generator, no dataset)",
            "config": {"workload": "batched clip of %d independent pairs, log-normal vertex "
                                   "counts (median 64, sigma 1.5, 8..50000), 70%% "
                                   "overlapping, FP64, seed 0" % args.pairs,
                       "ok": cs["ok"], "disjoint": cs["disjoint"], "failed": cs["failed"],
                       "result_vertices": cs["total_verts"],
                       "autotuned": None if args.no_tune else tuned,
                       "tuning_s": round(tune_s, 2)},
            "latency_ms_p50": round(1e3 * float(np.percentile(lat, 50)), 4),
            "latency_ms_p99": round(1e3 * float(np.percentile(lat, 99)), 4),
            "kernel_ms_per_step": kernel_ms,
            "per_pair_path": {"pairs_per_s": round(per_pair_rate, 2),
                              "sample": "%d pairs (every %d-th), one gcmf_clip call each"
                                        % (len(sample), stride),
                              "result_mismatches_vs_batch": mism,
                              "batched_speedup": round(args.pairs / (elapsed / args.steps)
                                                       / per_pair_rate, 2)},
            "roofline": label_roofline(lab_ms, lab_n, cs, None),
            "roofline_next": count_roofline(cnt_ms, cnt_n, cs, None), "graph_replay": graph}
    d_pairs.free()
    dP.free()
    dQ.free()
    ctx.close()
    print(json.dumps(line))


def large(args):
    """One large pair: a base ring of 1e6 and a clip ring of 1e5 vertices,
    radii 1 + 0.2 (1/6) sum_k sin(2^k t + phi_k) + 0.02 U(-1, 1) about (0, 0)
    and (0.3, 0.2), optionally snapped to a grid (consecutive duplicates
    dropped).  One step = one clip of the pair."""
    import gcmf_amd as G
    nP, nQ = args.large_n
    P, Q = G.gen_single_pair(1, 0, nP, nQ, snap=args.snap)
    ctx = G.Context(0)
    base, over = G.Layer.from_rings([P]), G.Layer.from_rings([Q])
    db, do = G.DeviceLayer(ctx, base), G.DeviceLayer(ctx, over)
    pairs = np.array([[0, 0]], np.int32)
    flags = G.CLIP_NORMALIZE | (G.CLIP_UNIT_FRAME if args.reference_eps else 0)
    for _ in range(args.warmup):
        cs = ctx.clip(db, do, pairs, flags)
    ctx.timer_reset()
    ctx.set_timing(True, only=LIVE_KERNELS)
    lat = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        cs = ctx.clip(db, do, pairs, flags)
        lat.append(time.perf_counter() - s0)
        print("step %d: %.2f ms" % (i, 1e3 * lat[-1]), file=sys.stderr, flush=True)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_timing(False)
    cnt_ms, cnt_n = ctx.timer(COUNT_KERNEL)
    lab_ms, lab_n = ctx.timer(LABEL_KERNEL)
    ctx.timer_reset()
    ctx.set_timing(True)
    ctx.clip(db, do, pairs, flags)
    ctx.synchronize()
    ctx.set_timing(False)
    kernel_ms = {k: round(ctx.timer(k)[0], 4) for k in (
        "gh_pair_setup", "gh_pair_fill", "gh_count", "gh_transpose", "gh_cpu_fill",
        "gh_cpu_sort",
        "gh_sort_runs", "gh_cpu_link", "gh_label_trace", "gh_gather_out", "sort_hist",
        "sort_scatter", "scan_reduce", "scan_downsweep")}
    path = int(ctx.clip_get("label_path", 1)[0])
    status = int(ctx.clip_get("status", 1)[0])
    # the kernel that labels the pair: by its node count (ring vertices and
    # intersection copies) against the context's thresholds
    tune = ctx.tuning()
    nodes = cs["slots_p"] + cs["slots_q"]
    lab_kernel = ("k_label_grid" if nodes >= tune["grid_nodes"] else
                  "k_label_block" if nodes >= tune["big_nodes"] else "k_label_wave")
    line = {"metric": METRIC, "value": round(args.steps / elapsed, 4), "unit": "pairs/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded generator, no dataset)",
            "config": {"workload": "one pair of %d x %d vertices (six-sinusoid radii, "
                                   "snap %g), FP64, seed 0" % (len(P), len(Q), args.snap),
                       "predicate_frame": "CMBR-normalised" + (
                           ", unit frame (GCMF_CLIP_UNIT_FRAME)" if args.reference_eps
                           else ", edge-scaled (the default)"),
                       "status": status, "label_path": path,
                       "result_components": cs["total_comps"],
                       "result_vertices": cs["total_verts"],
                       "intersection_pairs": cs["n_intersections"],
                       "n_eval_build": cs["n_eval"],
                       "n_eval_reference_algorithm": len(P) * len(Q)},
            "latency_ms_p50": round(1e3 * float(np.percentile(lat, 50)), 3),
            "latency_ms_p99": round(1e3 * float(np.percentile(lat, 99)), 3),
            "kernel_ms_per_step": kernel_ms,
            "roofline": label_roofline(lab_ms, lab_n, cs, None, lab_kernel),
            "roofline_next": count_roofline(cnt_ms, cnt_n, cs, None)}
    db.free()
    do.free()
    ctx.close()
    print(json.dumps(line))


def prims(args):
    """HBM sweep of the device primitives and the MBR filter at 2^20..2^27
    elements: achieved algorithmic GB/s and the fraction of MI355X's 8 TB/s.
    Algorithmic bytes: sort P passes x (8 + 4 read + 8 + 4 written) per key
    (P = bits / 8); scan 4 read + 8 written per element; compaction 4 read
    per flag + 4 written per kept index; MBR filter (B + O) x 32 read + K x 8
    written (the candidate pairs)."""
    import gcmf_amd as G
    ctx = G.Context(0)
    rs = np.random.RandomState(0)
    out = []

    def timed(fn, reps):
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t0) / reps

    for lg in range(20, 28):
        n = 1 << lg
        keys = rs.randint(0, 1 << 62, size=n, dtype=np.int64).astype(np.uint64)
        bufs = [ctx.upload(keys), ctx.upload(np.arange(n, dtype=np.int32)),
                ctx.upload(np.zeros(n, np.uint64)), ctx.upload(np.zeros(n, np.int32))]
        for bits in (32, 64):
            t = timed(lambda: ctx.radix_sort(bufs[0], bufs[1], bufs[2], bufs[3], n, bits), 3)
            gb = (bits // 8) * n * 24 / t / 1e9
            out.append({"prim": "radix_sort", "n": n, "bits": bits, "ms": round(1e3 * t, 4),
                        "GB/s": round(gb, 1), "frac": round(gb / HBM_PEAK_GBS, 4)})
        for b in bufs:
            b.free()
        a = rs.randint(0, 100, size=n).astype(np.int32)
        d_in, d_out = ctx.upload(a), ctx.upload(np.zeros(n + 1, np.int64))
        t = timed(lambda: ctx.scan(d_in, d_out, n), 5)
        gb = 12 * n / t / 1e9
        out.append({"prim": "scan", "n": n, "ms": round(1e3 * t, 4), "GB/s": round(gb, 1),
                    "frac": round(gb / HBM_PEAK_GBS, 4)})
        d_idx = ctx.upload(np.zeros(n, np.int32))
        flags = (a < 50).astype(np.int32)
        d_f = ctx.upload(flags)
        kept = int(flags.sum())
        t = timed(lambda: ctx.compact(d_f, d_idx, n), 5)
        gb = (4 * n + 4 * kept) / t / 1e9
        out.append({"prim": "compact", "n": n, "ms": round(1e3 * t, 4), "GB/s": round(gb, 1),
                    "frac": round(gb / HBM_PEAK_GBS, 4)})
        for b in (d_in, d_out, d_idx, d_f):
            b.free()
        print(json.dumps(out[-4:]), file=sys.stderr, flush=True)
    # MBR filter (endpoint sort + sweep) on layers of 8-vertex polygons with
    # 2 (B + O) = 2^20 .. 2^24 endpoint keys
    for lg in range(20, 25):
        npoly = (1 << lg) // 4
        r = 1.0 / np.sqrt(npoly)  # a few candidates per polygon at every size
        base, over = G.gen_join_layers(npoly, npoly, seed=lg, rmin=0.2 * r, rmax=0.5 * r, nmin=8,
                                       nmax=8)
        db, do = G.DeviceLayer(ctx, base), G.DeviceLayer(ctx, over)
        js = ctx.join(db, do)
        ctx.timer_reset()
        ctx.set_timing(True)
        ctx.join(db, do)
        ctx.synchronize()
        ctx.set_timing(False)
        names = ["mbr_endpoints", "sort_hist", "sort_scatter", "scan_tile", "scan_reduce",
                 "scan_downsweep", "scan_lookback", "mbr_sweep_prep", "mbr_sweep_count",
                 "mbr_sweep_write", "mbr_strip_build", "mbr_strip_count", "mbr_strip_write"]
        ms = sum(ctx.timer(k)[0] for k in names)
        sweep_ms = sum(ctx.timer(k)[0] for k in ("mbr_sweep_prep", "mbr_sweep_count",
                                                  "mbr_sweep_write", "mbr_strip_build",
                                                  "mbr_strip_count", "mbr_strip_write"))
        K = js["candidates"]
        gb = (2 * npoly * 32 + K * 8) / (ms * 1e-3) / 1e9
        out.append({"prim": "mbr_filter", "keys": 4 * npoly, "candidates": K,
                    "ms": round(ms, 4), "sweep_ms": round(sweep_ms, 4), "GB/s": round(gb, 1),
                    "frac": round(gb / HBM_PEAK_GBS, 4)})
        # CMF edge filter (count + apply) over the same join's candidates that
        # the PiP test did not settle: every vertex of both polygons read
        # (16 B) and every surviving edge index written (4 B)
        cmf_ms = ctx.timer("cmf_count")[0] + ctx.timer("cmf_apply")[0]
        pairs = ctx.join_get("pairs", 2 * K).reshape(-1, 2)
        pip = ctx.join_get("pip_flag", K)
        todo = pairs[pip != 1]
        verts = int(base.vnum[todo[:, 0]].sum()) + int(over.vnum[todo[:, 1]].sum())
        cbytes = 16 * verts + 4 * (js["b_edges"] + js["o_edges"])
        cgb = cbytes / (cmf_ms * 1e-3) / 1e9
        out.append({"prim": "cmf_filter", "keys": 4 * npoly, "pairs": int(len(todo)),
                    "vertices": verts, "surviving_edges": int(js["b_edges"] + js["o_edges"]),
                    "ms": round(cmf_ms, 4), "GB/s": round(cgb, 1),
                    "frac": round(cgb / HBM_PEAK_GBS, 4)})
        db.free()
        do.free()
        print(json.dumps(out[-2:]), file=sys.stderr, flush=True)
    ctx.close()
    print(json.dumps({"workload": "primitive HBM sweep", "peak_GBs": HBM_PEAK_GBS,
                      "results": out}))


def serve(args):
    """Served stream: Poisson arrivals at --rate requests/s for --duration s;
    90 % clips of one pair drawn from a resident pool of the batched
    workload's pairs, 10 % joins of two resident layers drawn from four layer
    sets of the join workload's generator at 1000, 5000, 10000 and 20000
    polygons; admission by a free-HBM budget (at most --hbm-budget, of the
    free HBM the device reports).  Then --saturate clip requests arriving at
    once (the saturated served rate), and the per-pair path (one gcmf_clip
    call per pair) on 2000 of the pool's pairs for comparison."""
    import gcmf_amd as G
    import serve as S
    print("generating the pool ...", file=sys.stderr, flush=True)
    P, Q = G.gen_pair_layers(args.pairs, seed=0)
    sets = [G.gen_join_layers(n, n, seed=10 + i) for i, n in enumerate((1000, 5000, 10000,
                                                                        20000))]
    srv = S.Server(P, Q, sets, hbm_budget=args.hbm_budget)
    # warm-up: one batch and one join of each set, untimed
    warm = [S.Request(0, "clip", 0.0, 0)] + [S.Request(1 + i, "join", 0.0, i) for i in range(4)]
    srv.run(warm)
    srv.batches.clear()
    srv.deferred = 0
    reqs = S.poisson_requests(args.rate, args.duration, 0.9, args.pairs, len(sets), seed=0)
    print("serving %d requests over %.0f s ..." % (len(reqs), args.duration), file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    srv.run(reqs)
    wall = time.perf_counter() - t0
    lat = S.latency_summary(reqs)
    clips = [r for r in reqs if r.kind == "clip"]
    joins = [r for r in reqs if r.kind == "join"]
    makespan = max(r.t_done for r in reqs)
    sizes = [b[2] for b in srv.batches]
    stream_batches = list(srv.batches)
    n_stream_batches = len(stream_batches)
    # saturated: every request there at once; the loop batches them as the
    # budget and max_batch allow
    rs = np.random.RandomState(1)
    sat = [S.Request(i, "clip", 0.0, int(k)) for i, k in
           enumerate(rs.randint(args.pairs, size=args.saturate))]
    srv.batches.clear()
    srv.run(sat)
    sat_span = max(r.t_done for r in sat)
    sat_batches = len(srv.batches)
    # the per-pair path on a sample of the pool
    sample = rs.randint(args.pairs, size=2000)
    srv.cctx.synchronize()
    t1 = time.perf_counter()
    for k in sample:
        srv.cctx.clip(srv.dP, srv.dQ, np.array([[k, k]], np.int32))
    srv.cctx.synchronize()
    per_pair = len(sample) / (time.perf_counter() - t1)
    line = {"metric": METRIC, "value": round(len(clips) / makespan, 2), "unit": "pairs/s",
            "n_gpus": 1, "steps": n_stream_batches, "warmup": 1,
            "ms_per_step": round(1e3 * float(np.mean([b[1] - b[0] for b in stream_batches])), 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded generators, no dataset)",
            "config": {"workload": "served stream: Poisson %g req/s for %g s, 90%% clips of "
                                   "one pair from a resident pool of %d log-normal pairs, 10%% "
                                   "joins of resident layer pairs of 1k/5k/10k/20k polygons, "
                                   "free-HBM budget %.0f GB (at most the free HBM measured)"
                                   % (args.rate, args.duration, args.pairs,
                                      srv.budget / 1e9),
                       "requests": len(reqs), "clips": len(clips), "joins": len(joins),
                       "clip_batches": n_stream_batches,
                       "batch_size_mean": round(float(np.mean(sizes)), 2),
                       "batch_size_max": int(max(sizes)),
                       "deferred_by_budget": srv.deferred,
                       "failed_clips": sum(1 for r in clips if r.status not in (0, 2))},
            "latency_ms_p50": lat["all"]["p50_ms"], "latency_ms_p99": lat["all"]["p99_ms"],
            "latency": lat,
            "join_candidates_per_s": round(sum(r.summary[0] for r in joins) / makespan, 2),
            "requests_per_s": round(len(reqs) / makespan, 2),
            "saturated": {"clips": len(sat), "batches": sat_batches,
                          "pairs_per_s": round(len(sat) / sat_span, 2),
                          "vs_per_pair_path": round(len(sat) / sat_span / per_pair, 2)},
            "per_pair_path_pairs_per_s": round(per_pair, 2),
            "admission": {"budget_bytes": int(srv.budget), "free_hbm_bytes": int(srv.hbm_free),
                          "clip_bytes_per_vertex": int(srv.clip_bytes_per_vertex),
                          "join_bytes_per_polygon": int(srv.join_bytes_per_polygon)},
            "wall_s": round(wall, 3)}
    srv.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
