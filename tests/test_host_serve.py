"""Host logic of the serving loop and of bench.py that needs no GPU: the
Poisson request generator, the latency summary, the free-HBM admission
budget, the roofline arithmetic and the one-GPU refusal."""
import os
import subprocess
import sys
import threading

import numpy as np
import pytest

import serve as S
from conftest import ROOT


def test_poisson_requests_deterministic_and_ordered():
    a = S.poisson_requests(500.0, 2.0, 0.75, 37, 3, seed=5)
    b = S.poisson_requests(500.0, 2.0, 0.75, 37, 3, seed=5)
    assert [(r.kind, r.arrival, r.key) for r in a] == [(r.kind, r.arrival, r.key) for r in b]
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0) and t[0] > 0 and t[-1] < 2.0
    assert [r.rid for r in a] == list(range(len(a)))
    for r in a:
        assert 0 <= r.key < (37 if r.kind == "clip" else 3)


def test_poisson_requests_rate_and_mix():
    # the count of a Poisson process over T at rate l is Poisson(l T): mean
    # and variance 20000 here, so 5 sigma is about 707; the clip share is
    # Binomial(n, 0.75) with sigma about 0.0031
    reqs = S.poisson_requests(4000.0, 5.0, 0.75, 10, 4, seed=1)
    n = len(reqs)
    assert abs(n - 20000) < 5 * np.sqrt(20000)
    share = sum(r.kind == "clip" for r in reqs) / n
    assert abs(share - 0.75) < 5 * np.sqrt(0.75 * 0.25 / n)
    # inter-arrival gaps are exponential with mean 1/rate
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 2.5e-4) < 5 * 2.5e-4 / np.sqrt(n)
    # every key value is drawn
    assert set(r.key for r in reqs if r.kind == "clip") == set(range(10))
    assert set(r.key for r in reqs if r.kind == "join") == set(range(4))


def test_poisson_requests_without_join_sets_are_clips():
    reqs = S.poisson_requests(1000.0, 1.0, 0.1, 5, 0, seed=2)
    assert reqs and all(r.kind == "clip" for r in reqs)


def test_latency_summary_percentiles():
    reqs = []
    for i in range(100):
        r = S.Request(i, "clip", 10.0, 0)
        r.t_done = 10.0 + (i + 1) * 1e-3          # latencies 1 .. 100 ms
        reqs.append(r)
    late = S.Request(100, "join", 0.0, 0)
    late.t_done = 0.25
    pending = S.Request(101, "join", 0.0, 0)      # not done: left out
    out = S.latency_summary(reqs + [late, pending])
    # linear interpolation over 1..100: p50 = 50.5, p99 = 1 + 0.99 * 99
    assert out["clip"] == {"n": 100, "p50_ms": 50.5, "p99_ms": 99.01}
    assert out["join"] == {"n": 1, "p50_ms": 250.0, "p99_ms": 250.0}
    assert out["all"]["n"] == 101 and out["all"]["p50_ms"] == 51.0
    assert S.latency_summary([pending]) == {}


def budget(limit):
    srv = S.Server.__new__(S.Server)    # the budget only: no device contexts
    srv._lock = threading.Lock()
    srv._inflight = 0.0
    srv.budget = float(limit)
    return srv


def test_admission_budget():
    srv = budget(100)
    assert srv._admit(60)
    assert srv._admit(40)               # exactly full
    assert not srv._admit(1)            # over: waits for a release
    srv._release(40)
    assert not srv._admit(41)
    assert srv._admit(40)
    srv._release(60)
    srv._release(40)
    assert srv._inflight == 0
    # a lone request larger than the whole budget is still admitted when
    # nothing is in flight, so it cannot wait forever
    assert srv._admit(1000)
    assert not srv._admit(1)
    srv._release(1000)
    assert srv._inflight == 0


def test_admission_budget_threads_balance():
    srv = budget(1e9)

    def work():
        for _ in range(2000):
            assert srv._admit(3.0)
            srv._release(3.0)

    ts = [threading.Thread(target=work) for _ in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert srv._inflight == 0


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as B
    return B


def test_count_roofline_arithmetic(bench):
    """The count kernel priced as 32 flops per classified edge pair plus 7
    per intersection point (box rejects are not work), against 78.6 TFLOP/s."""
    cs = {"n_eval": 3_000_000, "n_tested": 50_000_000, "n_intersections": 400_000}
    r = bench.count_roofline(4.0, 8, cs, 123456)
    flops = 32 * 3_000_000 + 7 * 400_000
    achieved = flops / 0.5e-3 / 1e12               # 8 launches in 4 ms
    assert r["avg_launch_ms"] == 0.5 and r["launches"] == 8
    assert r["achieved"] == pytest.approx(achieved, abs=1e-4)
    assert r["bound"] == "fp64-valu" and r["peak"] == 78.6 and r["unit"] == "TFLOP/s"
    assert r["frac"] == pytest.approx(achieved / 78.6, abs=1e-6)
    assert r["traffic"] == 123456
    zero = bench.count_roofline(0.0, 0, cs, None)
    assert zero["achieved"] == 0.0 and zero["frac"] == 0.0 and zero["traffic"] is None


def test_label_roofline_arithmetic(bench):
    """The label kernel priced by its algorithmic bytes against 8 TB/s."""
    cs = {"slots_p": 1000, "slots_q": 500, "edges_p": 800, "edges_q": 400,
          "total_verts": 300, "total_comps": 7, "pairs": 3}
    b = 36 * 1500 + 8 * 1200 + 16 * 300 + 4 * 7 + 8 * 3
    assert bench.label_bytes(cs) == b
    r = bench.label_roofline(2.0, 4, cs, None)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["achieved"] == pytest.approx(b / 0.5e-3 / 1e9, abs=1e-2)


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
