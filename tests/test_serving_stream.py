"""CPU checks of the served query stream (pool configurations, mix, Poisson
arrivals) and of the latency summary."""
from collections import Counter

import numpy as np

from graphalgo_amd import serving as S


def test_pool_configs_cover_scales_and_edge_factors():
    cfg = S.pool_configs(1)
    assert [c["scale"] for c in cfg] == list(range(10, 21))
    assert all(8 <= c["edge_factor"] <= 32 for c in cfg)
    assert cfg == S.pool_configs(1)  # deterministic
    assert cfg != S.pool_configs(2)


def test_stream_mix_and_arrivals(G):
    pool = S.make_pool(S.pool_configs(1, scales=range(8, 11)))
    specs, arr = S.make_stream(pool, 4000, 500.0, seed=3)
    assert arr[0] == 0 and (np.diff(arr) >= 0).all()
    # exponential gaps at 500 QPS: mean 2 ms, coefficient of variation 1
    gaps = np.diff(arr)
    assert abs(gaps.mean() - 2.0) < 0.15 and abs(gaps.std() / gaps.mean() - 1.0) < 0.08
    c = Counter(q["kind"] for q in specs)
    assert abs(c["tc"] / 4000 - 0.60) < 0.03 and abs(c["ktruss"] / 4000 - 0.25) < 0.03
    assert abs(c["bc"] / 4000 - 0.15) < 0.03
    for q in specs:
        g = pool[q["graph"]]
        if q["kind"] == "ktruss":
            assert 3 <= q["k"] <= 16
        if q["kind"] == "bc":
            s = q["sources"]
            assert len(s) == 64 and len(set(s.tolist())) == 64 and (g.degrees()[s] > 0).all()
    specs2, arr2 = S.make_stream(pool, 4000, 500.0, seed=3)
    assert np.array_equal(arr, arr2) and [q["kind"] for q in specs] == [q["kind"] for q in specs2]


def test_latency_summary():
    arr = np.array([0.0, 10.0, 20.0, 30.0])
    res = [{"t_end_ms": t} for t in (5.0, 30.0, 25.0, 40.0)]
    s = S.latency_summary(res, arr)
    assert s["max_ms"] == 20.0 and s["makespan_ms"] == 40.0
    assert s["throughput_qps"] == 100.0
    assert s["p50_ms"] == np.percentile([5, 20, 5, 10], 50)
