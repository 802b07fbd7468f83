"""The request loop (gcmf-gh_amd/serve.py): clip requests are continuously
batched, joins run concurrently on a second context, admission is by the
free HBM the device reports and charges measured from workspace growth.

Every served request is checked against the oracle, bit for bit, on a mix
shaped like the served stream's (90 % clips of log-normal pairs of the
batched workload's generator, 10 % joins of layer pairs of the join
workload's generator at several sizes): a clip's status, component sizes
and vertices against the CPU route of polyclip.h (orc_clip_cpu, in the
clip's frame), a join's candidate list (order included) and joined list
against the MBR sweep and the refine (orc_mbr_sweep, orc_join_refine).
With a budget of three large requests, admissions are deferred and the
results do not change."""
import numpy as np
import pytest

import gcmf_amd as G
import oracle_py as O
import serve as S

pytestmark = pytest.mark.gpu

NORM = 5  # the oracle's frame mask for GCMF_CLIP_NORMALIZE


def _pool():
    P, Q = G.gen_pair_layers(300, seed=3, nmax=4000)
    joins = [G.gen_join_layers(n, n, seed=20 + i) for i, n in enumerate((1000, 3000, 6000))]
    return P, Q, joins


def _oracle_clip(P, Q, key):
    a, b = P.ring(key), Q.ring(key)
    if len(a) < len(b):
        a, b = b, a
    st, comps, _ = O.clip_cpu(a, b, NORM)
    return st, comps


def _oracle_join(base, over):
    _, cand = O.mbr_sweep(base.mbr, over.mbr)
    ref = O.join_refine(cand, base.as_dict(), over.as_dict())
    return cand, cand[(ref["pip_flag"] == 1) | (ref["seg_flag"] == 1)]


def _check_oracle(P, Q, joins, reqs):
    clips, jn = {}, {}
    n_clip = n_join = 0
    for r in reqs:
        assert r.t_done is not None, r.rid
        if r.kind == "clip":
            if r.key not in clips:
                clips[r.key] = _oracle_clip(P, Q, r.key)
            st, comps = clips[r.key]
            assert r.status == st, (r.rid, r.key, r.status, st)
            assert [len(c) for c in r.result] == [len(c) for c in comps], (r.rid, r.key)
            for a, b in zip(r.result, comps):
                np.testing.assert_array_equal(a, b, err_msg="clip %d" % r.key)
            n_clip += 1
        else:
            if r.key not in jn:
                jn[r.key] = _oracle_join(*joins[r.key])
            cand, joined = jn[r.key]
            np.testing.assert_array_equal(r.result[0], cand, err_msg="join %d" % r.key)
            np.testing.assert_array_equal(r.result[1], joined, err_msg="join %d" % r.key)
            assert r.summary == (len(cand), len(joined))
            n_join += 1
    return n_clip, n_join


@pytest.mark.parametrize("tight", [False, True])
def test_serve_mix_matches_oracle(tight):
    P, Q, joins = _pool()
    srv = S.Server(P, Q, joins, keep_results=True)
    try:
        free, total, _ = srv.cctx.mem_info()
        assert 0 < srv.budget <= free and free <= total
        assert srv.clip_bytes_per_vertex >= S.CLIP_BYTES_PER_VERTEX
        assert srv.calibration_clip_bytes > 0
        if tight:
            # room for about three of the largest requests at a time:
            # admissions are deferred while earlier work is in flight
            srv.budget = 3.0 * float(srv.pair_cost.max())
        reqs = S.poisson_requests(rate=20000, duration=0.05, clip_fraction=0.9, n_pairs=300,
                                  n_join_sets=len(joins), seed=5)
        srv.run(reqs)
        # continuous batching: requests that arrived during a batch share the next
        assert max(b[2] for b in srv.batches) > 1
        if tight:
            assert srv.deferred > 0
        n_clip, n_join = _check_oracle(P, Q, joins, reqs)
        assert n_clip > 500 and n_join > 30, (n_clip, n_join)
        lat = S.latency_summary(reqs)
        assert lat["all"]["n"] == len(reqs)
    finally:
        srv.close()
