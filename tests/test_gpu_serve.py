"""The request loop (gcmf-gh_amd/serve.py): clip requests are continuously
batched, joins run concurrently on a second context, and every request's
result is the same as when it is run alone (bit-identical: status, component
sizes and SHA-256 of the vertices for clips; candidate and joined pair lists
for joins), also when the HBM budget defers admissions."""
import numpy as np
import pytest

import gcmf_amd as G
import serve as S

pytestmark = pytest.mark.gpu


def _pool():
    P, Q = G.gen_pair_layers(400, seed=3, nmax=3000)
    joins = [G.gen_join_layers(1000, 1000, seed=s, rmin=2e-3, rmax=2e-2, nmin=8, nmax=64)
             for s in (1, 2)]
    return P, Q, joins


def _check_alone(srv, reqs):
    alone = {}
    for r in reqs:
        assert r.t_done is not None, r.rid
        if r.kind == "clip":
            if r.key not in alone:
                alone[r.key] = S.clip_alone(srv.cctx, srv.dP, srv.dQ, r.key)
            ref = alone[r.key]
        else:
            b, o = srv.joins[r.key]
            ref = S.join_alone(srv.jctx, b, o)
        assert (r.status, r.summary, r.digest) == ref, (r.rid, r.kind, r.key)


@pytest.mark.parametrize("budget", [200e9, None])
def test_serve_same_alone_and_in_mix(budget):
    P, Q, joins = _pool()
    nv = (P.vnum.astype(np.int64) - 1) + (Q.vnum.astype(np.int64) - 1)
    if budget is None:
        # room for about three of the largest requests at a time: admissions
        # are deferred while earlier work is in flight
        budget = 3.0 * S.CLIP_BYTES_PER_VERTEX * int(nv.max())
    srv = S.Server(P, Q, joins, hbm_budget=budget, keep_results=True)
    try:
        reqs = S.poisson_requests(rate=20000, duration=0.02, clip_fraction=0.9, n_pairs=400,
                                  n_join_sets=2, seed=5)
        assert sum(r.kind == "join" for r in reqs) > 0
        srv.run(reqs)
        # continuous batching: requests that arrived during a batch share the next
        assert max(b[2] for b in srv.batches) > 1
        if budget < 1e9:
            assert srv.deferred > 0
        _check_alone(srv, reqs)
        lat = S.latency_summary(reqs)
        assert lat["all"]["n"] == len(reqs)
    finally:
        srv.close()
