"""The served query stream: a seeded pool of resident R-MAT graphs and a
Poisson stream of triangle-count, k-truss and BC queries over it.

Pool: one R-MAT graph per scale in 10..20 (a=0.57, b=0.19, c=0.19), edge
factor drawn uniformly from 8..32 per scale by a SplitMix64 hash of
(seed, scale), graph seed = 1000 * seed + scale; symmetrised, deduplicated,
loop-free (the input contract of the reference's TC / k-truss CLI path).

Stream: exponential inter-arrival times at a fixed QPS; each query picks a
pool graph uniformly and is a triangle count (60%), a k-truss for one k
drawn uniformly from 3..16 (25%) or a 64-source BC (15%, sources = seeded
random non-isolated vertices).  Every draw comes from numpy's PCG64 seeded
with the stream seed, so the same call yields the same stream everywhere.

The query kinds are the reference's own calls (src/main.cpp:65-86, :89-146;
include/betweenness_centrality.h:125-237); the reference has no serving
layer to mirror.
"""
from __future__ import annotations

import numpy as np

from . import graphs as G

MIX = (("tc", 0.60), ("ktruss", 0.25), ("bc", 0.15))
K_RANGE = (3, 16)
BC_SOURCES = 64


def pool_configs(seed: int = 1, scales=range(10, 21), ef_range=(8, 32)) -> list[dict]:
    """R-MAT configurations of the resident pool (scale, edge factor, seed)."""
    out = []
    lo, hi = ef_range
    for s in scales:
        h = int(G.splitmix64(np.uint64(seed) * np.uint64(1 << 32) + np.uint64(s)))
        out.append({"scale": int(s), "edge_factor": lo + h % (hi - lo + 1), "seed": 1000 * seed + int(s)})
    return out


def make_pool(configs: list[dict]) -> list[G.CSR]:
    return [G.rmat(c["scale"], c["edge_factor"], seed=c["seed"]) for c in configs]


def make_stream(pool: list[G.CSR], nq: int, qps: float, seed: int = 1) -> tuple[list[dict], np.ndarray]:
    """(query specs, arrival times in ms, ascending, first at 0)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1000.0 / qps, nq)
    arrival = np.concatenate([[0.0], np.cumsum(gaps[1:])]) if nq else np.zeros(0)
    kinds = rng.choice(len(MIX), nq, p=[p for _, p in MIX])
    graph = rng.integers(0, len(pool), nq)
    ks = rng.integers(K_RANGE[0], K_RANGE[1] + 1, nq)
    src_seed = rng.integers(0, 2**31 - 1, nq)
    specs = []
    for i in range(nq):
        kind = MIX[kinds[i]][0]
        q = {"kind": kind, "graph": int(graph[i])}
        if kind == "ktruss":
            q["k"] = int(ks[i])
        elif kind == "bc":
            q["sources"] = G.pick_sources(pool[graph[i]], BC_SOURCES, int(src_seed[i]))
        specs.append(q)
    return specs, arrival


def latency_summary(results: list[dict], arrival_ms: np.ndarray) -> dict:
    """p50 / p99 / mean latency (completion - arrival, ms) and throughput
    (queries / (last completion - first arrival), per second)."""
    end = np.array([r["t_end_ms"] for r in results])
    lat = end - np.asarray(arrival_ms)
    span = float(end.max() - np.min(arrival_ms)) if len(end) else 0.0
    return {
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "max_ms": float(lat.max()),
        "makespan_ms": span,
        "throughput_qps": len(results) / (span / 1000.0) if span > 0 else float("inf"),
    }
