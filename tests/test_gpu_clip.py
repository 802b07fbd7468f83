"""GPU parity of the batched Greiner-Hormann clip (gh.hip) against the oracle.

The clip's default phase 1 is the reference's CPU route (computeIntersections,
polyclip.h:188-281): its results must be identical (same status, components,
vertex counts and coordinates) to the oracle's CPU route on every input,
degenerate ones included.  With GCMF_CLIP_GPU_ROUTE the phase-1 arrays
(counts, scans, CMBR flags, neighbour map, intersection coordinates, alpha
keys, initial labels, neighbour links) must be bit-exact against the
restatement of gppolyclip_mDIV512.cu and the results equal to the GPU-route
oracle.  Traced areas must agree with the independent boundary-integral area
within 1e-12 relative.
"""
import numpy as np
import pytest

import gcmf_amd as G
import oracle_py as O
from conftest import GOLDEN, load_poly

pytestmark = pytest.mark.gpu

AREA_RTOL = 1e-12
# the oracle's frame mask for GCMF_CLIP_NORMALIZE: bit 0 CMBR normalisation,
# bit 2 the edge scale (GCMF_CLIP_UNIT_FRAME drops bit 2)
NORM = 5


def fig_cases():
    out = []
    for fig, pn in [("Fig1", "Fig1-P"), ("Fig14", "Fig14-P"), ("Fig14cw", "Fig14-P-clockWise"),
                    ("Fig15", "Fig15-P"), ("Fig17", "Fig17-P-Clockwise")]:
        qn = fig.replace("cw", "") + "-Q"
        P = load_poly("%s/fig/%s.poly" % (GOLDEN, pn))[0]
        Q = load_poly("%s/fig/%s.poly" % (GOLDEN, qn))[0]
        out.append((fig, P, Q))
    return out


def random_cases(n, seed0=0, nmax=200):
    out = []
    rs = np.random.RandomState(seed0)
    for s in range(n):
        a, b = int(rs.randint(8, nmax)), int(rs.randint(8, nmax))
        P, Q = G.gen_single_pair(0, 1000 + seed0 + s, a, b)
        out.append(("rand%d" % s, P, Q))
    return out


def run_batch(gctx, cases, flags):
    base = G.Layer.from_rings([c[1] for c in cases])
    over = G.Layer.from_rings([c[2] for c in cases])
    db, do = G.DeviceLayer(gctx, base), G.DeviceLayer(gctx, over)
    M = len(cases)
    pairs = np.stack([np.arange(M), np.arange(M)], axis=1).astype(np.int32)
    st = gctx.clip(db, do, pairs, flags)
    db.free()
    do.free()
    return st


def oriented(P, Q):
    """The pair as the clip sees it: P is the ring with more vertices."""
    return (Q, P) if len(P) < len(Q) else (P, Q)


def check_phase1(gctx, st, cases, normalize):
    M = len(cases)
    f = {k: gctx.clip_get(k, n) for k, n in [
        ("status", M), ("offP", M + 1), ("offQ", M + 1), ("psP1", st["edges_p"] + 1),
        ("psP2", st["edges_p"] + 1), ("psQ1", st["edges_q"] + 1), ("psQ2", st["edges_q"] + 1),
        ("boolP", st["edges_p"]), ("boolQ", st["edges_q"]), ("nmapQ", st["slots_q"]),
        ("iP", 2 * st["slots_p"]), ("iQ", 2 * st["slots_q"]), ("oP", 2 * st["slots_p"]),
        ("oQ", 2 * st["slots_q"]), ("aP", st["slots_p"]), ("aQ", st["slots_q"]),
        ("lP", st["slots_p"]), ("lQ", st["slots_q"]), ("nbP", st["slots_p"]),
        ("nbQ", st["slots_q"])]}
    for k, (name, P0, Q0) in enumerate(cases):
        P, Q = oriented(P0, Q0)
        ref = O.phase1(P, Q, normalize)
        if ref["rc"] == 2:
            assert f["status"][k] == 2, name
            continue
        p0, p1 = int(f["offP"][k]), int(f["offP"][k + 1])
        q0, q1 = int(f["offQ"][k]), int(f["offQ"][k + 1])
        assert p1 - p0 == len(P) and q1 - q0 == len(Q), name
        for key, lo, hi in [("psP1", p0, p1), ("psP2", p0, p1), ("psQ1", q0, q1),
                            ("psQ2", q0, q1)]:
            got = f[key][lo:hi + 1] - f[key][lo]
            np.testing.assert_array_equal(got, ref[key], err_msg="%s %s" % (name, key))
        np.testing.assert_array_equal(f["boolP"][p0:p1], ref["boolP"], err_msg=name)
        np.testing.assert_array_equal(f["boolQ"][q0:q1], ref["boolQ"], err_msg=name)
        if ref["invalid"]:
            assert f["status"][k] == 1, name
            continue
        assert f["status"][k] in (0, -1), name
        sbP, sbQ = int(f["psP2"][p0]), int(f["psQ2"][q0])
        cP, cQ = ref["cP"], ref["cQ"]
        assert int(f["psP2"][p1]) - sbP == cP and int(f["psQ2"][q1]) - sbQ == cQ, name
        np.testing.assert_array_equal(f["iP"][2 * sbP:2 * (sbP + cP)], ref["intersectionsP"],
                                      err_msg=name + " iP")
        np.testing.assert_array_equal(f["iQ"][2 * sbQ:2 * (sbQ + cQ)], ref["intersectionsQ"],
                                      err_msg=name + " iQ")
        np.testing.assert_array_equal(f["oP"][2 * sbP:2 * (sbP + cP)], ref["origP"],
                                      err_msg=name + " oP")
        np.testing.assert_array_equal(f["oQ"][2 * sbQ:2 * (sbQ + cQ)], ref["origQ"],
                                      err_msg=name + " oQ")
        np.testing.assert_array_equal(f["aP"][sbP:sbP + cP], ref["alphaP"], err_msg=name)
        np.testing.assert_array_equal(f["aQ"][sbQ:sbQ + cQ], ref["alphaQ"], err_msg=name)
        np.testing.assert_array_equal(f["lP"][sbP:sbP + cP], ref["initLabelsP"], err_msg=name)
        np.testing.assert_array_equal(f["lQ"][sbQ:sbQ + cQ], ref["initLabelsQ"], err_msg=name)
        nbP = f["nbP"][sbP:sbP + cP].astype(np.int64)
        nbQ = f["nbQ"][sbQ:sbQ + cQ].astype(np.int64)
        np.testing.assert_array_equal(np.where(nbP > 0, nbP - sbQ, 0), ref["neighborP"],
                                      err_msg=name + " nbP")
        np.testing.assert_array_equal(np.where(nbQ > 0, nbQ - sbP, 0), ref["neighborQ"],
                                      err_msg=name + " nbQ")
        nm = f["nmapQ"][sbQ:sbQ + cQ]
        np.testing.assert_array_equal(nm, ref["neighborMapQ"], err_msg=name + " nmap")


def check_results(gctx, st, cases, mode, gpu_route=False):
    """CPU route (default): every pair equals the oracle's CPU route of
    polyclip.h.  GPU route: every pair equals the GPU-route oracle (the
    reference's GPU phase 1 spliced and labelled as ghcuda.cpp does), and
    pairs in general position also equal the CPU route.  On degenerate
    (grid-snapped) inputs the reference's two routes can differ: the GPU
    route's phase-1 lists order coincident intersections by truncated alpha,
    the CPU route by exact insertion."""
    status, results = gctx.clip_results(st)
    for k, (name, P0, Q0) in enumerate(cases):
        P, Q = oriented(P0, Q0)
        if gpu_route:
            refs = [O.clip_gpu_route(P, Q, mode)]
            if not name.startswith("snap"):
                refs.append(O.clip_cpu(P, Q, mode))
        else:
            refs = [O.clip_cpu(P, Q, mode)]
        for s_ref, comps, _ in refs:
            if s_ref == 2:
                assert status[k] == 2, name
                continue
            assert status[k] == s_ref, (name, status[k], s_ref)
            got = results[k]
            assert [len(c) for c in got] == [len(c) for c in comps], name
            for a, b in zip(got, comps):
                np.testing.assert_array_equal(a, b, err_msg=name)


@pytest.mark.parametrize("normalize", [1, 0])
def test_phase1_fig_fixtures(gctx, normalize):
    cases = fig_cases()
    st = run_batch(gctx, cases, (G.CLIP_NORMALIZE if normalize else 0) | G.CLIP_GPU_ROUTE)
    check_phase1(gctx, st, cases, NORM if normalize else 0)


def test_phase1_random_batch(gctx):
    cases = random_cases(64)
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE | G.CLIP_GPU_ROUTE)
    assert st["ok"] == 64
    check_phase1(gctx, st, cases, NORM)


def test_phase1_c1_pair(gctx):
    P, Q = G.gen_single_pair(0, 0, 2000, 1500)
    cases = [("c1", P, Q)]
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE | G.CLIP_GPU_ROUTE)
    check_phase1(gctx, st, cases, NORM)
    assert st["n_intersections"] == O.phase1(P, Q, NORM)["psP1"][-1]
    check_results(gctx, st, cases, NORM, gpu_route=True)


@pytest.mark.parametrize("flags,mode", [(G.CLIP_NORMALIZE, NORM), (0, 0),
                                        (G.CLIP_NORMALIZE | G.CLIP_NO_CLEANUP, NORM | 2)])
def test_clip_c1_pair_results(gctx, flags, mode):
    """The single 2000 x 1500 pair: the final clipped polygon (status,
    component sizes, every vertex) equals the CPU route's bit for bit, in
    the default frame, the raw frame and without clean-up."""
    P, Q = G.gen_single_pair(0, 0, 2000, 1500)
    cases = [("c1", P, Q)]
    st = run_batch(gctx, cases, flags)
    assert st["ok"] == 1 and st["total_comps"] > 0
    check_results(gctx, st, cases, mode)


def test_phase1_grid_snapped(gctx):
    # shared vertices and collinear edges (degenerate intersection types)
    cases = []
    for s in range(6):
        P, Q = G.gen_single_pair(1, 50 + s, 3000, 400, snap=1e-2)
        cases.append(("snap%d" % s, P, Q))
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE | G.CLIP_GPU_ROUTE)
    check_phase1(gctx, st, cases, NORM)
    check_results(gctx, st, cases, NORM, gpu_route=True)


def snapped_cases(seed0, n=4, grid=1e-2):
    out = []
    for s in range(n):
        P, Q = G.gen_single_pair(1, seed0 + s, 1500, 300, snap=grid)
        out.append(("snap%d" % s, P, Q))
    return out


@pytest.mark.parametrize("seq", [0, 1])
@pytest.mark.parametrize("mode", [1, 3, 0])
def test_clip_matches_cpu_route(gctx, mode, seq):
    # seq=1 runs labelling/tracing one thread per pair, seq=0 the wave path;
    # grid-snapped pairs (shared vertices, collinear edges) included
    cases = fig_cases() + random_cases(48, seed0=77) + snapped_cases(90)
    flags = (G.CLIP_NORMALIZE if mode & 1 else 0) | (G.CLIP_NO_CLEANUP if mode & 2 else 0)
    flags |= G.CLIP_SEQ_LABEL if seq else 0
    st = run_batch(gctx, cases, flags)
    check_results(gctx, st, cases, mode | (4 if mode & 1 else 0))
    if not seq:
        # pairs in general position take the parallel paths, not the
        # sequential fallback
        path = gctx.clip_get("label_path", len(cases))
        for k, (name, _, _) in enumerate(cases):
            if name.startswith("rand"):
                assert path[k] in (G.LP_SEGMENTS, G.LP_SEG_CLEANUP, G.LP_NO_CROSSING), (name,
                                                                                      path[k])


def test_clip_gpu_route_results(gctx):
    """The GPU route gives the GPU-route oracle's results (and the CPU
    route's in general position), through the wave path."""
    cases = fig_cases() + random_cases(48, seed0=77) + snapped_cases(90)
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE | G.CLIP_GPU_ROUTE)
    check_results(gctx, st, cases, NORM, gpu_route=True)


def test_clip_degenerate_wavy(gctx):
    """A 20000 x 5000 pair of the large-pair shape (radii from six
    random-phase sinusoids plus noise) and its copy snapped to a 1e-3 grid
    (shared vertices, collinear edges): identical to the CPU route; the
    unsnapped pair's traced area equals the boundary integral."""
    P, Q = G.gen_single_pair(1, 0, 20000, 5000)
    Ps, Qs = G.gen_single_pair(1, 0, 20000, 5000, snap=1e-3)
    cases = [("wavy", P, Q), ("wavy_snap", Ps, Qs)]
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE)
    assert st["ok"] == 2
    check_results(gctx, st, cases, NORM)
    st = run_batch(gctx, cases[:1], G.CLIP_NORMALIZE | G.CLIP_NO_CLEANUP)
    _, res = gctx.clip_results(st)
    A = O.area_boundary(P, Q)
    allxy = np.concatenate([P, Q])
    ox, oy = 0.5 * (allxy[:, 0].min() + allxy[:, 0].max()), 0.5 * (
        allxy[:, 1].min() + allxy[:, 1].max())
    assert abs(O.comps_area(res[0], ox, oy) - A) <= AREA_RTOL * A


def test_clip_edge_scaled_fine_pair(gctx):
    """A 50000 x 5000 pair of the large-pair shape in the edge-scaled frame
    (GCMF_CLIP_SCALE_EDGES, mode bit 4 of the oracle): identical to the CPU
    route in the same frame, and its traced area equals the boundary
    integral within 1e-12."""
    P, Q = G.gen_single_pair(1, 0, 50000, 5000)
    cases = [("fine", P, Q)]
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE | G.CLIP_SCALE_EDGES)
    assert st["ok"] == 1
    check_results(gctx, st, cases, 5)
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE | G.CLIP_SCALE_EDGES | G.CLIP_NO_CLEANUP)
    check_results(gctx, st, cases, 7)
    _, res = gctx.clip_results(st)
    A = O.area_boundary(P, Q)
    allxy = np.concatenate([P, Q])
    ox, oy = 0.5 * (allxy[:, 0].min() + allxy[:, 0].max()), 0.5 * (
        allxy[:, 1].min() + allxy[:, 1].max())
    assert abs(O.comps_area(res[0], ox, oy) - A) <= AREA_RTOL * A


def test_clip_large_pair_fixture(gctx):
    """The 200000 x 20000 pair of the large-pair shape in the edge-scaled
    frame, against the oracle's result committed in
    tests/golden/large_pair_200k.json (made by make_large_pair.py, 2.5 min on
    one core): the same status and component sizes, bit-identical vertices
    (SHA-256 of the float64 coordinates).  The committed result is cleaned
    (cleanUpResult drops vertices whose doubled turn area is below the
    absolute 1e-7), which moves its area by 3.5e-10 relative from the
    boundary integral; the traced area without clean-up is checked at 1e-12
    on smaller pairs."""
    import hashlib
    import json
    with open(GOLDEN + "/large_pair_200k.json") as f:
        ref = json.load(f)
    assert abs(ref["area_result"] - ref["area_boundary"]) <= 1e-9 * ref["area_boundary"]
    P, Q = G.gen_single_pair(1, 0, 200000, 20000)
    st = run_batch(gctx, [("large", P, Q)], G.CLIP_NORMALIZE | G.CLIP_SCALE_EDGES)
    status, res = gctx.clip_results(st)
    assert status[0] == ref["status"] == 0
    assert st["n_intersections"] == ref["n_inter"]
    assert [len(c) for c in res[0]] == ref["comp_sizes"]
    xy = np.ascontiguousarray(np.concatenate(res[0]), np.float64)
    assert hashlib.sha256(xy.tobytes()).hexdigest() == ref["xy_sha256"]


def test_clip_area_oracle(gctx):
    cases = random_cases(40, seed0=300, nmax=400)
    P, Q = G.gen_single_pair(0, 0, 2000, 1500)
    cases.append(("c1", P, Q))
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE | G.CLIP_NO_CLEANUP)
    status, results = gctx.clip_results(st)
    for k, (name, P0, Q0) in enumerate(cases):
        A = O.area_boundary(P0, Q0)
        allxy = np.concatenate([P0, Q0])
        ox, oy = 0.5 * (allxy[:, 0].min() + allxy[:, 0].max()), 0.5 * (
            allxy[:, 1].min() + allxy[:, 1].max())
        got = O.comps_area(results[k], ox, oy)
        assert abs(got - A) <= AREA_RTOL * max(A, 1e-300), (name, got, A)


def test_clip_batch_independence(gctx):
    """Each pair's result is the same alone and inside a mixed batch.  Calls
    of at most 32 pairs run the per-pair set-up with 1024 threads per pair,
    split over several blocks (512 for one pair, 17 for 30); 40 pairs run it
    one wave per pair: the results and the number of edge pairs the
    prefilter lets through (which follows the filter constants) agree."""
    cases = random_cases(30, seed0=500, nmax=300)
    extra = random_cases(10, seed0=700, nmax=300)
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE)
    _, mixed = gctx.clip_results(st)
    for k in (0, 7, 29):
        st1 = run_batch(gctx, [cases[k]], G.CLIP_NORMALIZE)
        _, alone = gctx.clip_results(st1)
        assert len(alone[0]) == len(mixed[k])
        for a, b in zip(alone[0], mixed[k]):
            np.testing.assert_array_equal(a, b)
    st_x = run_batch(gctx, extra, G.CLIP_NORMALIZE)
    st40 = run_batch(gctx, cases + extra, G.CLIP_NORMALIZE)
    _, wide = gctx.clip_results(st40)
    assert st40["n_eval"] == st["n_eval"] + st_x["n_eval"]
    assert st40["n_intersections"] == st["n_intersections"] + st_x["n_intersections"]
    for k in range(30):
        assert len(wide[k]) == len(mixed[k])
        for a, b in zip(wide[k], mixed[k]):
            np.testing.assert_array_equal(a, b)


def test_clip_uneven_batch(gctx):
    """Pairs of log-normal sizes (median 64, sigma 1.5, 8..50000), 70 %
    overlapping, clipped in one batch; a sample is compared with the CPU route
    and the whole batch's per-pair areas with the boundary integral."""
    P, Q = G.gen_pair_layers(3000, seed=1, nmax=5000)
    dP, dQ = G.DeviceLayer(gctx, P), G.DeviceLayer(gctx, Q)
    pairs = np.stack([np.arange(3000), np.arange(3000)], axis=1).astype(np.int32)
    st = gctx.clip(dP, dQ, pairs, G.CLIP_NORMALIZE)
    status, res = gctx.clip_results(st)
    dP.free()
    dQ.free()
    assert st["failed"] == 0 and st["invalid"] == 0
    # every classified edge pair was visited first; a batch of 3000 pairs is
    # labelled largest first
    assert st["n_tested"] >= st["n_eval"] > 0
    for k in range(0, 3000, 37):
        a, b = oriented(P.ring(k), Q.ring(k))
        s_ref, comps, _ = O.clip_cpu(a, b, NORM)
        assert status[k] == (2 if s_ref == 2 else 0)
        assert [len(c) for c in res[k]] == [len(c) for c in comps]
        for x, y in zip(res[k], comps):
            np.testing.assert_array_equal(x, y)


def test_clip_disjoint_and_nested(gctx):
    P, Q = G.gen_single_pair(0, 5, 100, 80)
    far = Q + np.array([10.0, 10.0])
    inner = 0.2 * P  # inside P (star about the origin scaled down)
    cases = [("disjoint", P, far), ("nested", P, inner), ("identical", P, P.copy())]
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE)
    status, res = gctx.clip_results(st)
    assert status[0] == 2 and res[0] == []
    s_ref, comps, _ = O.clip_cpu(P, inner, NORM)
    assert status[1] == 0 and len(res[1]) == len(comps) == 1
    np.testing.assert_array_equal(res[1][0], comps[0])
    s_ref, comps, _ = O.clip_cpu(P, P.copy(), NORM)
    assert [len(c) for c in res[2]] == [len(c) for c in comps]


def long_run_cases():
    """Pairs whose edges carry more than 32 intersections, so the per-edge
    alpha sorts leave the per-edge thread: a jagged 4000-vertex star against
    an 8-vertex star (Q runs up to 274, ranked by one block each), a 12-gon
    whose corner 0 is replaced by a 700-point dent well inside the other ring
    against a jagged 690-vertex star (P runs of 33-36 on the 12-gon's
    chords), and a jagged 50000-vertex star against a 6-vertex star (a Q run
    of 3290, beyond one block, through the segmented radix sort)."""
    P1, Q1 = G.gen_single_pair(0, 4242, 4000, 8)
    ang = 2 * np.pi * np.arange(12) / 12
    gon = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    t = np.linspace(-0.5 * np.pi, 0.5 * np.pi, 700)
    dent = np.array([0.5, 0.0]) + 0.05 * np.stack([np.cos(t), np.sin(t)], axis=1)
    P2 = np.concatenate([gon[1:], dent])
    Q2, _ = G.gen_single_pair(0, 4343, 690, 8)
    P3, Q3 = G.gen_single_pair(0, 3, 50000, 6)
    return [("longQ", P1, Q1), ("longP", P2, Q2), ("longQ3290", P3, Q3)]


def test_clip_long_runs(gctx):
    cases = long_run_cases()
    for name, P, Q in cases:
        Po, Qo = oriented(P, Q)
        ref = O.phase1(Po, Qo, NORM)
        side = "psP2" if name == "longP" else "psQ2"
        assert int(np.diff(ref[side]).max()) - 1 > (2048 if name == "longQ3290" else 32), name
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE | G.CLIP_GPU_ROUTE)
    check_phase1(gctx, st, cases, NORM)
    check_results(gctx, st, cases, NORM, gpu_route=True)
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE)
    check_results(gctx, st, cases, NORM)
    for p in gctx.clip_get("label_path", len(cases)):
        assert p in (G.LP_SEGMENTS, G.LP_SEG_CLEANUP)


@pytest.mark.parametrize("gpu_route", [0, 1])
def test_clip_batched_rare_pairs(gctx, gpu_route):
    """Pairs from the batched workload that reach rare paths (a trace that
    does not close, BOUNCING intersections among thousands of crossings)
    equal the oracle of the route, status included, and the BOUNCING pairs
    stay on the parallel labelling path."""
    z = np.load(GOLDEN + "/batched_pairs.npz")
    names = ["p4596", "p156581", "p73779", "p176719"]
    cases = [(n, z[n + "_P"], z[n + "_Q"]) for n in names]
    st = run_batch(gctx, cases, G.CLIP_NORMALIZE | (G.CLIP_GPU_ROUTE if gpu_route else 0))
    status, results = gctx.clip_results(st)
    path = gctx.clip_get("label_path", len(cases))
    for k, (name, P0, Q0) in enumerate(cases):
        P, Q = oriented(P0, Q0)
        s_ref, comps, _ = (O.clip_gpu_route if gpu_route else O.clip_cpu)(P, Q, NORM)
        assert status[k] == s_ref, (name, status[k], s_ref)
        assert [len(c) for c in results[k]] == [len(c) for c in comps], name
        for a, b in zip(results[k], comps):
            np.testing.assert_array_equal(a, b, err_msg=name)
    assert status[0] == -1 and path[0] == G.LP_FAILED
    for k in (1, 2, 3):  # split and unsplit BOUNCING pairs
        assert path[k] in (G.LP_SEGMENTS, G.LP_SEG_CLEANUP), (names[k], path[k])


def test_clip_autotuner_keeps_results(gctx):
    """gcmf_tune_clip picks launch shapes by measurement; results are the
    same under the tuned and the default shapes, and set_tuning rejects
    invalid shapes."""
    P, Q = G.gen_pair_layers(300, seed=9, nmax=6000)
    dP, dQ = G.DeviceLayer(gctx, P), G.DeviceLayer(gctx, Q)
    pairs = np.stack([np.arange(300), np.arange(300)], axis=1).astype(np.int32)
    default = gctx.tuning()
    st0 = gctx.clip(dP, dQ, pairs)
    _, r0 = gctx.clip_results(st0)
    d_pairs = gctx.upload(pairs)
    try:
        tuned = gctx.tune_clip(dP, dQ, d_pairs, 300)
        assert tuned["count_block"] in (64, 128, 256)
        assert tuned == gctx.tuning()
        st1 = gctx.clip(dP, dQ, pairs)
        _, r1 = gctx.clip_results(st1)
        assert len(r0) == len(r1)
        for a, b in zip(r0, r1):
            assert [len(c) for c in a] == [len(c) for c in b]
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)
        with pytest.raises(G.GcmfError):
            gctx.set_tuning(512, 8192, 100)
        for bad in ({"label_lds": 100}, {"grid_nodes": 0}, {"grid_clean": 48},
                    {"strip_min": -1}, {"cmf_group": 32}, {"label_steal": 2},
                    {"block_cap": 0}, {"count_stage": 2}):
            with pytest.raises(G.GcmfError):
                gctx.set_tuning(**dict(default, **bad))
        assert gctx.tuning() == tuned
    finally:
        gctx.set_tuning(**default)
        d_pairs.free()
        dP.free()
        dQ.free()


@pytest.mark.parametrize("big_nodes", [3000, 1 << 40])
def test_clip_block_threshold_keeps_results(gctx, big_nodes):
    """A batch big enough to be labelled largest first, split between the
    wave path and the block path at a node threshold that is not a power of
    two (or with no block path at all): every pair is labelled once, and the
    results equal the default shapes' bit for bit."""
    P, Q = G.gen_pair_layers(1500, seed=4, nmax=6000)
    dP, dQ = G.DeviceLayer(gctx, P), G.DeviceLayer(gctx, Q)
    pairs = np.stack([np.arange(1500), np.arange(1500)], axis=1).astype(np.int32)
    default = gctx.tuning()
    try:
        st0 = gctx.clip(dP, dQ, pairs)
        s0, r0 = gctx.clip_results(st0)
        gctx.set_tuning(default["cull_min"], big_nodes, default["count_block"])
        st1 = gctx.clip(dP, dQ, pairs)
        s1, r1 = gctx.clip_results(st1)
        np.testing.assert_array_equal(s0, s1)
        assert len(r0) == len(r1)
        for a, b in zip(r0, r1):
            assert [len(c) for c in a] == [len(c) for c in b]
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)
    finally:
        gctx.set_tuning(**default)
        dP.free()
        dQ.free()


@pytest.mark.parametrize("flags", [G.CLIP_NORMALIZE, 0])
def test_clip_reference_large_fixture(gctx, flags):
    """The reference's own large pair (s.txt, 101 242 integer-coordinate
    vertices, x c.txt, 72 997) through gcmf_clip, in the default frame and in
    the raw frame: status 0, 47 intersections and the 11 components of
    GH-CUDA/results/Fig-large-R.poly (sizes as GH-CUDA/ReadMe.txt:59), every
    vertex equal to the file at its printed precision."""
    from conftest import large_fixture, printed
    S, Cq, R = large_fixture()
    st = run_batch(gctx, [("large", S, Cq)], flags)
    status, res = gctx.clip_results(st)
    assert status[0] == 0 and st["n_intersections"] == 47
    assert [len(c) for c in res[0]] == [len(c) for c in R]
    for a, b in zip(res[0], R):
        np.testing.assert_array_equal(printed(a), b)


def test_count_lds_pass_matches_ring_walk(gctx):
    """The CPU route's count pass with the Q ring binned in LDS (pairs whose
    Q ring is longer than cull_min and at most lds_max_q vertices) gives the
    ring walk's phase-1 arrays and results bit for bit: the Fig fixtures,
    random pairs of the join workload's sizes, grid-snapped pairs (tiny and
    collinear edges), the long-run cases, and an uneven batch whose Q rings
    above 1024 vertices still walk, all in one launch.  Three runs: LDS
    pass from 64 vertices up, the culled walk from 64 vertices up, and the
    walk over whole rings (which leaves out only the runs of 16 edges that
    no lane of a wave can pair with); the LDS pass classifies the same edge
    pairs as the whole-ring walk and visits fewer."""
    cases = fig_cases() + random_cases(64, seed0=900, nmax=520) + snapped_cases(200) + \
        long_run_cases()
    P, Q = G.gen_pair_layers(600, seed=12, nmax=3000)
    cases += [("u%d" % k, P.ring(k), Q.ring(k)) for k in range(600)]
    default = gctx.tuning()
    out = []
    try:
        for cull, lds in ((64, 1024), (64, 0), (1 << 40, 0)):
            gctx.set_tuning(cull, default["big_nodes"], default["count_block"], lds)
            st = run_batch(gctx, cases, G.CLIP_NORMALIZE)
            n = {"psP1": st["edges_p"] + 1, "psP2": st["edges_p"] + 1,
                 "psQ1": st["edges_q"] + 1, "psQ2": st["edges_q"] + 1,
                 "boolP": st["edges_p"], "boolQ": st["edges_q"], "iP": 2 * st["slots_p"],
                 "iQ": 2 * st["slots_q"], "oP": 2 * st["slots_p"], "oQ": 2 * st["slots_q"],
                 "nbP": st["slots_p"], "nbQ": st["slots_q"], "status": len(cases),
                 "comp_sizes": st["total_comps"], "xy": 2 * st["total_verts"]}
            out.append((st, {f: gctx.clip_get(f, c) for f, c in n.items()}))
    finally:
        gctx.set_tuning(**default)
    (s1, a1), (s2, a2), (s0, a0) = out
    assert s1["n_eval"] == s0["n_eval"] == s2["n_eval"]
    assert s1["n_intersections"] == s0["n_intersections"] == s2["n_intersections"]
    assert s1["n_tested"] < s0["n_tested"], (s1["n_tested"], s0["n_tested"])
    for f in a1:
        np.testing.assert_array_equal(a1[f], a0[f], err_msg=f)
        np.testing.assert_array_equal(a2[f], a0[f], err_msg=f)


def test_count_wave_walk_equals_lane_walk(gctx):
    """The count pass's wave walk (the edges of one pair in a wave walk the
    other ring together by 16-edge runs, leaving out runs none of them can
    pair with; count_stage 0) gives the lane-by-lane walk's (count_stage 1)
    phase-1 arrays and results bit for bit, and visits fewer edge pairs: the
    Fig fixtures, random pairs of the join workload's sizes, grid-snapped
    pairs (tiny and collinear edges), the long-run cases and an uneven batch
    whose longest rings take the culled walk, with the culling threshold at
    its default and at 64."""
    cases = fig_cases() + random_cases(64, seed0=950, nmax=520) + snapped_cases(200) + \
        long_run_cases()
    P, Q = G.gen_pair_layers(600, seed=13, nmax=3000)
    cases += [("u%d" % k, P.ring(k), Q.ring(k)) for k in range(600)]
    default = gctx.tuning()
    try:
        for cull in (default["cull_min"], 64):
            out = []
            for stage in (0, 1):
                gctx.set_tuning(cull, default["big_nodes"], default["count_block"], 0,
                                count_stage=stage)
                st = run_batch(gctx, cases, G.CLIP_NORMALIZE)
                n = {"psP1": st["edges_p"] + 1, "psP2": st["edges_p"] + 1,
                     "psQ1": st["edges_q"] + 1, "psQ2": st["edges_q"] + 1,
                     "boolP": st["edges_p"], "boolQ": st["edges_q"], "iP": 2 * st["slots_p"],
                     "iQ": 2 * st["slots_q"], "nbP": st["slots_p"], "nbQ": st["slots_q"],
                     "status": len(cases), "comp_sizes": st["total_comps"],
                     "xy": 2 * st["total_verts"]}
                out.append((st, {f: gctx.clip_get(f, c) for f, c in n.items()}))
            (s0, a0), (s1, a1) = out
            assert s0["n_eval"] == s1["n_eval"] and s0["n_intersections"] == s1["n_intersections"]
            assert s0["n_tested"] < s1["n_tested"], (cull, s0["n_tested"], s1["n_tested"])
            for f in a0:
                np.testing.assert_array_equal(a0[f], a1[f], err_msg="%s cull %d" % (f, cull))
    finally:
        gctx.set_tuning(**default)


@pytest.mark.parametrize("label_lds", [0, 2048, 65536])
def test_label_lds_arena_keeps_results(gctx, label_lds):
    """The wave labelling kernel keeps a pair's node arrays, segment scratch
    and normalised output in LDS as far as label_lds bytes allow, the rest
    in HBM: with none (0), with room for the first arrays of small pairs
    only (2048) and with room for whole pairs (65536), every pair equals the
    oracle, through the segment path, the clean-up, the no-crossing path
    and the sequential fallback (which moves the nodes out of LDS first)."""
    cases = fig_cases() + random_cases(48, seed0=77) + snapped_cases(90)
    rs = np.random.RandomState(5)
    for s in range(6):  # nested and disjoint pairs: the no-crossing path
        P, _ = G.gen_single_pair(0, 500 + s, int(rs.randint(8, 300)), 8)
        c = P.mean(axis=0)
        cases.append(("nest%d" % s, P, c + 0.3 * (P - c)))
    z = np.load(GOLDEN + "/batched_pairs.npz")
    cases += [(n, z[n + "_P"], z[n + "_Q"]) for n in ["p156581", "p73779", "p176719"]]
    default = gctx.tuning()
    try:
        gctx.set_tuning(**dict(default, label_lds=label_lds))
        for flags, mode in ((G.CLIP_NORMALIZE, NORM), (G.CLIP_NORMALIZE | G.CLIP_NO_CLEANUP,
                                                       NORM | 2)):
            st = run_batch(gctx, cases, flags)
            check_results(gctx, st, cases, mode)
            paths = set(gctx.clip_get("label_path", len(cases)).tolist())
            assert {G.LP_SEGMENTS, G.LP_NO_CROSSING, G.LP_SEQUENTIAL} <= paths, paths
            if flags == G.CLIP_NORMALIZE:
                assert G.LP_SEG_CLEANUP in paths, paths
    finally:
        gctx.set_tuning(**default)


def test_clip_grid_labelling_large_pair(gctx):
    """The 200000 x 20000 pair labelled by the whole GPU (one cooperative
    launch, grid_nodes below its node count) gives the committed oracle
    result bit for bit, as the 1024-thread block does."""
    import hashlib
    import json
    with open(GOLDEN + "/large_pair_200k.json") as f:
        ref = json.load(f)
    P, Q = G.gen_single_pair(1, 0, 200000, 20000)
    default = gctx.tuning()
    try:
        gctx.set_tuning(**dict(default, grid_nodes=1 << 17))
        st = run_batch(gctx, [("large", P, Q)], G.CLIP_NORMALIZE)
        status, res = gctx.clip_results(st)
        assert gctx.clip_get("label_path", 1)[0] in (G.LP_SEGMENTS, G.LP_SEG_CLEANUP)
    finally:
        gctx.set_tuning(**default)
    assert status[0] == ref["status"] == 0
    assert [len(c) for c in res[0]] == ref["comp_sizes"]
    xy = np.ascontiguousarray(np.concatenate(res[0]), np.float64)
    assert hashlib.sha256(xy.tobytes()).hexdigest() == ref["xy_sha256"]


def test_label_steal_keeps_results(gctx):
    """The persistent, work-stealing labelling launch (waves taking pairs from
    a device counter, label_steal 1) gives every pair the oracle's result and
    the one-wave-per-pair launch's, on a batch with the block path's pairs,
    the sequential fallback and the no-crossing path, in largest-first order
    (1024 pairs or more) and in index order."""
    z = np.load(GOLDEN + "/batched_pairs.npz")
    cases = [(n, z[n + "_P"], z[n + "_Q"]) for n in ["p4596", "p156581", "p73779", "p176719"]]
    cases += fig_cases() + random_cases(60, seed0=33, nmax=400) + snapped_cases(810, n=4)
    cases += [("wavy%d" % s,) + tuple(G.gen_single_pair(1, 50 + s, 6000, 1500)) for s in range(2)]
    P, Q = G.gen_pair_layers(1100, seed=7, nmax=2000)
    big = cases + [("u%d" % k, P.ring(k), Q.ring(k)) for k in range(1100)]
    default = gctx.tuning()
    for batch in (cases, big):
        st0 = run_batch(gctx, batch, G.CLIP_NORMALIZE)
        s0, r0 = gctx.clip_results(st0)
        try:
            gctx.set_tuning(**dict(default, label_steal=1))
            st = run_batch(gctx, batch, G.CLIP_NORMALIZE)
            s1, r1 = gctx.clip_results(st)
        finally:
            gctx.set_tuning(**default)
        np.testing.assert_array_equal(s1, s0)
        for x, y in zip(r1, r0):
            assert [len(c) for c in x] == [len(c) for c in y]
            for p, q in zip(x, y):
                np.testing.assert_array_equal(p, q)
    check_results(gctx, st, big, NORM)


def dense_square_cases():
    """A square with 1500 vertices on each side against a diamond and a
    7-point star crossing it: the result components carry runs of hundreds
    of collinear vertices that the clean-up drops (runs far beyond the
    map states of the parallel clean-up, so its overflow path is taken)."""
    t = np.linspace(0.0, 1.0, 1500, endpoint=False)
    z, o = np.zeros_like(t), np.ones_like(t)
    sq = np.concatenate([np.stack([t, z], 1), np.stack([o, t], 1), np.stack([1 - t, o], 1),
                         np.stack([z, 1 - t], 1)])
    a = np.arange(4) * np.pi / 2 + 0.1
    rot = np.array([0.5, 0.5]) + 0.62 * np.stack([np.cos(a), np.sin(a)], 1)
    b = np.pi / 2 + np.arange(14) * np.pi / 7
    r = np.where(np.arange(14) % 2 == 0, 0.8, 0.45)
    star = np.array([0.5, 0.45]) + np.stack([r * np.cos(b), r * np.sin(b)], 1)
    return [("dense_rot", sq, rot), ("dense_star", sq, star)]


def test_clip_grid_labelling_batch(gctx):
    """A batch split between the wave, block and grid paths (grid_nodes low
    enough that several pairs take cooperative launches one after another,
    among them the batched workload's rare pairs, a trace that does not
    close included): every pair equals the oracle of the route taken and the
    default shapes' results, status included.  Run twice: with the grid
    clean-up for components above 16384 vertices only, and for every
    component above 48 (grid_clean 49), where the dense squares' collinear
    runs send it to its overflow path."""
    z = np.load(GOLDEN + "/batched_pairs.npz")
    cases = [(n, z[n + "_P"], z[n + "_Q"]) for n in ["p4596", "p156581", "p73779", "p176719"]]
    cases += random_cases(30, seed0=61, nmax=300) + snapped_cases(300, n=3)
    cases += [("wavy%d" % s,) + tuple(G.gen_single_pair(1, 70 + s, 6000, 1500)) for s in range(3)]
    cases += dense_square_cases()
    st0 = run_batch(gctx, cases, G.CLIP_NORMALIZE)
    s0, r0 = gctx.clip_results(st0)
    assert G.LP_SEG_CLEANUP in set(gctx.clip_get("label_path", len(cases)).tolist())
    default = gctx.tuning()
    for grid_clean in (default["grid_clean"], 49):
        try:
            gctx.set_tuning(**dict(default, big_nodes=2000, grid_nodes=6000,
                                   grid_clean=grid_clean))
            st = run_batch(gctx, cases, G.CLIP_NORMALIZE)
            s1, r1 = gctx.clip_results(st)
        finally:
            gctx.set_tuning(**default)
        np.testing.assert_array_equal(s1, s0)
        for a, b in zip(r1, r0):
            assert [len(c) for c in a] == [len(c) for c in b]
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)
        check_results(gctx, st, cases, NORM)
