"""Benchmark: truss decomposition (k-truss for every k until empty) on R-MAT
scale 16, edge factor 16 (a=.57 b=.19 c=.19 d=.05, symmetrised, deduplicated,
loops removed), one GPU.

One step = one full truss decomposition of the HBM-resident graph through the
masked-SpGEMM support loop.  value = undirected edges / seconds per step.
The dominant kernel is the masked SpGEMM (k_mxm_items); its roofline entry
uses algorithmic bytes per launch

  4(m+1) + 4 nnz(A) + sum_{(i,k) in A} (8 + 4 nnz(B_k)) + 4(m+1) + 4 nnz(M) + 4 nnz(M)

(A = B = M = the current graph; values are implicit ones) divided by the GPU
time of its launches, measured with HIP events on the library's stream.
The CPU baseline is the OpenMP oracle's truss decomposition on the host.

Other workloads: --workload tc (triangle count on R-MAT 16), tc_er (one
small triangle count), bc (256-source batched BC on R-MAT 18), batch (4096
kNN graphs, batched TC), serve, ewise (eWiseAdd of two R-MAT 20 graphs),
ktruss_one (one k-truss query on graphs of scale 12..20 under each peeling
setting), copy (calibration).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "masked-SpGEMM edges/s + % HBM roofline; served p50/p99 latency vs plain"
HBM_PEAK_GBS = 8000.0


def _sync():
    from graphalgo_amd import _native as N
    N.init()
    N.check(N.galib().ga_synchronize())


def _cpu_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def _measure(step, steps: int, warmup: int):
    """Runs `warmup` untimed steps, then `steps` timed steps with HIP events
    around every masked-SpGEMM launch, then one extra untimed step with the
    algorithmic byte accounting on (every step issues the same launches).
    Returns (seconds per step, last step result, timing, accounting)."""
    from graphalgo_amd import api
    for _ in range(warmup):
        step()
    _sync()
    api.prof_enable(api.PROF_EVENTS)
    api.prof_read(reset=True)
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = step()
    _sync()
    dt = (time.perf_counter() - t0) / steps
    timing = api.prof_read(reset=True)
    api.prof_enable(api.PROF_BYTES)
    step()
    acct = api.prof_read(reset=True)
    api.prof_enable(0)
    timing["steps"] = steps
    return dt, out, timing, acct


def _traffic(key: str):
    """HBM bytes per launch measured offline with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (scaled by the copy-kernel calibration), as recorded in profiles/traffic.json."""
    try:
        rec = json.loads((ROOT / "profiles" / "traffic.json").read_text())[key]
        return rec["traffic_bytes_per_launch"], rec["source"]
    except (OSError, KeyError, ValueError):
        return None, None


def _roofline(timing: dict, acct: dict, kernel: str, traffic_key: str | None = None) -> dict:
    steps = max(timing.get("steps", 1), 1)
    launches = max(timing["launches"] // steps, 1)
    kernel_s = timing["kernel_ms"] / steps / 1e3
    achieved = acct["bytes_alg"] / kernel_s / 1e9 if kernel_s > 0 else 0.0
    traffic, source = _traffic(traffic_key) if traffic_key else (None, None)
    return {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": source,
        "kernel": kernel, "launches_per_step": launches,
        "avg_launch_us": round(kernel_s * 1e6 / launches, 2),
        "bytes_alg_per_launch": int(acct["bytes_alg"] / max(acct["accounted_launches"], 1)),
        "wedges_per_step": acct["wedges"], "wedges_processed_per_step": acct["wedges_processed"],
    }


def bench_ktruss(args) -> dict:
    from graphalgo_amd import api, graphs as G
    g = G.rmat(16, 16, seed=1)
    m = g.m_undirected
    R = api.Graph(g)
    dt, (_, kmax, stats), timing, acct = _measure(lambda: R.truss(want_values=False), args.steps, args.warmup)
    R.close()
    # per-step averages for the roofline (every step repeats the same launches)
    out = {
        "value": m / dt, "unit": "edges/s", "ms_per_step": dt * 1e3, "dtype": "int32",
        "config": {"workload": "truss decomposition (k-truss for all k) on R-MAT scale 16",
                   "graph": "rmat(scale=16, edge_factor=16, a=.57, b=.19, c=.19, seed=1)",
                   "n": g.n, "m_undirected": m, "nnz": g.nnz, "kmax": kmax,
                   "support_passes": stats.get("passes"), "filter_steps": stats.get("steps")},
        "roofline": _roofline(timing, acct, "k_mxm_items (support passes: symmetric half pass; triangle "
                                            "listing while the graph holds >= 2^20 entries)",
                              "k_mxm_items/truss_rmat16"),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_ktruss(g)
    return out


def cpu_ktruss(g) -> dict:
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    O.lib()
    t0 = time.perf_counter()
    O.truss_decomposition(g, max_passes=1)
    dt = time.perf_counter() - t0
    return {"value": g.m_undirected / dt, "unit": "edges/s", "cores": O.num_threads(), "kind": "port",
            "sample": "one full truss decomposition of the same R-MAT 16 graph by the OpenMP oracle "
                      f"({dt:.1f} s)"}


def bench_tc(args) -> dict:
    from graphalgo_amd import api, graphs as G
    g = G.rmat(16, 16, seed=1)
    R = api.Graph(g)
    dt, t, timing, acct = _measure(R.triangle_count, args.steps, args.warmup)
    R.close()
    out = {"value": g.m_undirected / dt, "unit": "edges/s", "ms_per_step": dt * 1e3, "dtype": "int32",
           "config": {"workload": "triangle count C<L>=L*L on R-MAT scale 16", "n": g.n,
                      "m_undirected": g.m_undirected, "triangles": t},
           "roofline": _roofline(timing, acct, "k_mxm_items (C<L> = L*L, sum mode)", "k_mxm_items/tc_rmat16")}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 5.0:
            O.tc(g)
            reps += 1
        dtc = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {"value": g.m_undirected / dtc, "unit": "edges/s", "cores": O.num_threads(),
                               "kind": "port", "sample": f"{reps} oracle triangle counts of the same graph"}
    return out


def bench_tc_er(args) -> dict:
    """Triangle count C<L> = L*L on the Erdos-Renyi graph n = 1000, m = 9990
    (seed 1): a latency-bound single launch; value = edges/s of one call on the
    resident graph, with the per-call latency beside it."""
    from graphalgo_amd import api, graphs as G
    g = G.erdos_renyi(1000, 9990, seed=1)
    R = api.Graph(g)
    dt, t, timing, acct = _measure(R.triangle_count, max(args.steps, 20), args.warmup + 5)
    R.close()
    out = {"value": g.m_undirected / dt, "unit": "edges/s", "ms_per_step": dt * 1e3, "dtype": "int32",
           "config": {"workload": "triangle count C<L>=L*L on Erdos-Renyi n=1000 m=9990 (seed 1)", "n": g.n,
                      "m_undirected": g.m_undirected, "triangles": t, "latency_us": round(dt * 1e6, 2)},
           "roofline": _roofline(timing, acct, "k_sum_subrows (C<L> = L*L, sum mode)")}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 2.0:
            O.tc(g)
            reps += 1
        dtc = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {"value": g.m_undirected / dtc, "unit": "edges/s", "cores": O.num_threads(),
                               "kind": "port", "sample": f"{reps} oracle triangle counts of the same graph "
                                                         f"({dtc * 1e6:.0f} us each)"}
    return out


def bench_ewise(args) -> dict:
    """eWiseAdd C = A + B (include/matrix_la.h:375-436) of two R-MAT scale 20
    graphs (edge factor 16, seeds 1 and 2) with int32 values.  The entry point
    takes host arrays, so value = input entries per second of device time (the
    count and fill passes, HIP events); the rate with the PCIe copies and the
    host allocations is reported beside it."""
    from graphalgo_amd import api, graphs as G
    A, B = G.rmat(20, 16, seed=1), G.rmat(20, 16, seed=2)
    rng = np.random.default_rng(0)
    va = rng.integers(1, 100, size=A.nnz).astype(np.int32)
    vb = rng.integers(1, 100, size=B.nnz).astype(np.int32)

    def step():
        return api.ewise_add(A, va, B, vb, dtype=np.int32)
    dt, (Cm, cv), timing, acct = _measure(step, args.steps, args.warmup)
    dev_s = timing["kernel_ms"] / timing["steps"] / 1e3
    entries = A.nnz + B.nnz
    out = {"value": entries / dev_s, "unit": "input entries/s", "ms_per_step": dev_s * 1e3, "dtype": "int32",
           "config": {"workload": "eWiseAdd C = A + B of two R-MAT scale 20 graphs (edge factor 16), int32 values",
                      "n": A.n, "nnz_A": A.nnz, "nnz_B": B.nnz, "nnz_C": Cm.nnz,
                      "host_inclusive_entries_per_s": entries / dt, "host_inclusive_ms_per_call": dt * 1e3},
           "roofline": _roofline(timing, acct, "k_ewise (count pass + scan, fill pass)", "k_ewise/ewise_add_rmat20")}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O
        t0 = time.perf_counter()
        O.ewise(O.EW_ADD, A.n, A.rowptr, A.col, va, B.rowptr, B.col, vb, dtype=np.int32)
        dtc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": entries / dtc, "unit": "input entries/s", "cores": 1, "kind": "port",
                               "sample": f"one numpy oracle eWiseAdd of the same operands ({dtc:.1f} s)"}
    return out


def bench_copy(args) -> dict:
    """Calibration: a plain copy kernel of known size (read + write bytes)."""
    import ctypes as C
    from graphalgo_amd import _native as N
    N.init()
    nbytes = 1 << 30
    ms = C.c_double()
    N.check(N.galib().ga_stream_copy(nbytes, max(args.steps, 1), C.byref(ms)))
    gbs = 2 * nbytes / (ms.value / 1e3) / 1e9
    return {"value": gbs, "unit": "GB/s", "ms_per_step": ms.value, "dtype": "u8",
            "config": {"workload": "device copy of 1 GiB (calibration)", "bytes_read": nbytes,
                       "bytes_written": nbytes},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": 2 * nbytes}}


def bench_batch(args) -> dict:
    """4096 kNN graphs (n ~ U[100,2000], 4-8 dim Gaussian mixtures, k=30, symmetrised):
    triangle counts of all graphs in one persistent launch, against the plain path
    that counts one resident graph per call."""
    from graphalgo_amd import api, graphs as G
    count = int(os.environ.get("GA_BATCH_GRAPHS", "4096"))
    t0 = time.perf_counter()
    graphs = G.knn_batch(base_seed=20240, count=count, k=30, workers=min(16, _cpu_cores()))
    gen_s = time.perf_counter() - t0
    m_total = sum(g.m_undirected for g in graphs)
    B = api.Batch(graphs)
    (best_w, best_c), tune_ms = B.autotune()
    dt, counts, timing, acct = _measure(B.triangle_counts, args.steps, args.warmup)
    # plain path: every graph resident, one triangle-count call per graph
    handles = [api.Graph(g) for g in graphs]
    for h in handles[:64]:
        h.triangle_count()
    t0 = time.perf_counter()
    plain = [h.triangle_count() for h in handles]
    dt_plain = time.perf_counter() - t0
    for h in handles:
        h.close()
    B.close()
    assert plain == counts.tolist(), "batched and per-graph counts differ"
    roof = _roofline(timing, acct, "k_sum_subrows + k_mxm_items (batched)", "k_sum_subrows/batched_tc_knn4096")
    out = {"value": m_total / dt, "unit": "edges/s", "ms_per_step": dt * 1e3, "dtype": "int32",
           "config": {"workload": f"batched triangle count over {count} kNN graphs (k=30) in one launch",
                      "graphs": count, "m_total": m_total, "triangles_total": int(counts.sum()),
                      "autotuned": {"item_width": best_w, "rows_per_claim": best_c},
                      "autotune_ms": {f"{w}x{c}": round(v, 3) for (w, c), v in tune_ms.items()},
                      "generation_s": round(gen_s, 1)},
           "roofline": roof,
           "plain_path": {"value": m_total / dt_plain, "unit": "edges/s", "ms": dt_plain * 1e3,
                          "speedup_batched_vs_plain": dt_plain / dt}}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O
        sample = graphs[:512]
        t0 = time.perf_counter()
        for g in sample:
            O.tc(g)
        dtc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": sum(g.m_undirected for g in sample) / dtc, "unit": "edges/s",
                               "cores": O.num_threads(), "kind": "port",
                               "sample": f"oracle triangle counts of the first 512 of the {count} graphs"}
    return out


def bench_bc(args) -> dict:
    """Batched betweenness centrality on R-MAT scale 18 (edge factor 16), 256 seeded
    random non-isolated sources as one 256-column pass, FP64 sigma and delta."""
    from graphalgo_amd import api, graphs as G
    g = G.rmat(18, 16, seed=1)
    src = G.pick_sources(g, 256, seed=7)
    R = api.Graph(g)
    dt, bc, timing, acct = _measure(lambda: R.bc(src), args.steps, args.warmup)
    R.close()
    m = g.m_undirected
    roof = _roofline(timing, acct, "k_bc_bottom_up_items / k_bc_discover_items + k_bc_sigma_items + "
                     "k_bc_backward_items (per level)", "k_bc/bc_rmat18")
    roof["launches_note"] = ("one event pair per device segment between host waits: candidate lists, "
                             "forward steps, the backward sweep")
    out = {"value": m * len(src) / dt, "unit": "edges*sources/s", "ms_per_step": dt * 1e3, "dtype": "f64",
           "config": {"workload": "batched BC, 256 sources, R-MAT scale 18", "n": g.n, "m_undirected": m,
                      "sources": len(src), "source_seed": 7, "bc_sum": float(bc.sum())},
           "roofline": roof}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O
        sample = src[:32]
        t0 = time.perf_counter()
        O.bc(g, sample)
        dtc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": m * len(sample) / dtc, "unit": "edges*sources/s",
                               "cores": O.num_threads(), "kind": "port",
                               "sample": f"oracle FP64 batched BC with the first 32 of the 256 sources ({dtc:.1f} s)"}
    return out


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_serve(args) -> dict:
    """Served query stream: Poisson arrivals at a fixed QPS mixing 60% triangle
    count, 25% k-truss (one k in 3..16 per query) and 15% 64-source BC over 11
    resident R-MAT graphs (scales 10..20, edge factors 8..32), admitted by HBM
    footprint into a fixed device arena.  The same stream is replayed on the
    plain path (one query at a time through the per-graph calls).  The QPS is
    set from the measured capacities (all queries at t = 0) unless --qps."""
    from graphalgo_amd import api, serving as S
    cfg = S.pool_configs(1)
    t0 = time.perf_counter()
    pool = S.make_pool(cfg)
    _log(f"pool generated in {time.perf_counter() - t0:.1f} s")
    srv = api.Server(hbm_budget=args.budget_gb << 30, slots=args.slots, max_batch=args.max_batch)
    for g in pool:
        srv.add_graph(g)
    graphs = [api.Graph(g) for g in pool]
    nq = args.queries
    # warm-up: every graph and kind once on both paths (code objects, first touch)
    warm, _ = S.make_stream(pool, 3 * len(pool), 1.0, seed=99)
    srv.replay(warm, np.zeros(len(warm)))
    api.plain_replay(graphs, warm, np.zeros(len(warm)))
    _log("warm-up done")
    sat, _ = S.make_stream(pool, nq, 1.0, seed=2)
    z = np.zeros(nq)
    r_sat = srv.replay(sat, z)
    cap_served = S.latency_summary(r_sat, z)["throughput_qps"]
    _log(f"served capacity {cap_served:.1f} queries/s")
    r_sat_plain = api.plain_replay(graphs, sat, z)
    cap_plain = S.latency_summary(r_sat_plain, z)["throughput_qps"]
    _log(f"plain capacity {cap_plain:.1f} queries/s")
    # the same saturated stream without batching: concurrency alone
    nb = api.Server(hbm_budget=args.budget_gb << 30, slots=args.slots, max_batch=1)
    for g in pool:
        nb.add_graph(g)
    cap_nobatch = S.latency_summary(nb.replay(sat, z), z)["throughput_qps"]
    nb.close()
    _log(f"served capacity without batching {cap_nobatch:.1f} queries/s")
    busy = {}
    for q, r in zip(sat, r_sat_plain):
        key = f"{q['kind']}@{cfg[q['graph']]['scale']}"
        busy[key] = busy.get(key, 0.0) + r["t_end_ms"] - r["t_start_ms"]
    top = sorted(busy.items(), key=lambda kv: -kv[1])[:8]
    qps = args.qps or round(0.8 * cap_served, 1)
    specs, arr = S.make_stream(pool, nq, qps, seed=3)
    served, lat = [], []
    t_served = 0.0
    for step in range(args.steps):
        t0 = time.perf_counter()
        res = srv.replay(specs, arr)
        t_served += time.perf_counter() - t0
        served.append(S.latency_summary(res, arr))
        _log(f"served replay {step}: p50 {served[-1]['p50_ms']:.2f} ms p99 {served[-1]['p99_ms']:.2f} ms")
    plain_res = api.plain_replay(graphs, specs, arr)
    plain = S.latency_summary(plain_res, arr)
    _log(f"plain replay: p50 {plain['p50_ms']:.2f} ms p99 {plain['p99_ms']:.2f} ms")
    for q, a, b in zip(specs, res, plain_res):
        if q["kind"] == "bc":
            assert np.allclose(a["bc"], b["bc"], rtol=1e-9), "served and plain BC differ"
        else:
            assert a["value"] == b["value"], "served and plain results differ"
    st = srv.stats()
    srv.close()
    for h in graphs:
        h.close()
    last = served[-1]
    kinds = {k: sum(q["kind"] == k for q in specs) for k, _ in S.MIX}
    out = {"value": last["throughput_qps"], "unit": "queries/s", "ms_per_step": 1e3 * t_served / args.steps,
           "dtype": "int32+f64",
           "config": {"workload": "served query stream over 11 resident R-MAT graphs (scale 10..20, edge factor "
                                  "8..32): Poisson arrivals, 60% TC / 25% k-truss / 15% 64-source BC",
                      "pool": [(c["scale"], c["edge_factor"]) for c in cfg], "queries": nq, "kinds": kinds,
                      "qps_offered": qps, "slots": args.slots, "arena_gb": args.budget_gb, "stream_seed": 3},
           "roofline": None,
           "served": {**{k: round(v, 3) for k, v in last.items()},
                      "p50_ms_all_steps": [round(x["p50_ms"], 3) for x in served],
                      "p99_ms_all_steps": [round(x["p99_ms"], 3) for x in served],
                      "capacity_qps": round(cap_served, 2), "capacity_qps_without_batching": round(cap_nobatch, 2),
                      "runs": st["batches"], "joined_runs": st["joined"], "queries_completed": st["completed"],
                      "peak_inflight": st["peak_inflight"],
                      "peak_arena_bytes": st["peak_bytes"], "resident_bytes": st["resident_bytes"]},
           "plain_path": {**{k: round(v, 3) for k, v in plain.items()}, "capacity_qps": round(cap_plain, 2),
                          "busy_ms_top_kind_at_scale": {k: round(v, 1) for k, v in top}},
           "served_vs_plain": {"throughput": last["throughput_qps"] / plain["throughput_qps"],
                               "capacity": cap_served / cap_plain,
                               "p50_latency": plain["p50_ms"] / last["p50_ms"],
                               "p99_latency": plain["p99_ms"] / last["p99_ms"]},
           "roofline_note": "queries of every kind run concurrently on 'slots' streams; per-kernel roofline "
                            "is reported by the ktruss, tc, bc and batch workloads"}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O
        t0 = time.perf_counter()
        done = 0
        for q in specs:
            g = pool[q["graph"]]
            if g.n > (1 << 15):
                continue
            if q["kind"] == "tc":
                O.tc(g)
            elif q["kind"] == "ktruss":
                O.ktruss(g, q["k"])
            else:
                O.bc(g, q["sources"])
            done += 1
            if time.perf_counter() - t0 > 15.0:
                break
        dtc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done / dtc, "unit": "queries/s", "cores": O.num_threads(), "kind": "port",
                               "sample": f"oracle on the stream's first {done} queries over pool graphs of scale "
                                         f"<= 15, back to back ({dtc:.1f} s)"}
    return out


def bench_ktruss_one(args) -> dict:
    """One k-truss query (k_truss for one k, src/main.cpp:89-146) on resident
    R-MAT graphs from scale 12 to the served pool's scale 18 and 20, k = 3 and
    8, with the default peeling settings and with each one forced: the loop
    (0 = host-driven, 1 = device-driven steps, 2 = steps replayed from a HIP
    graph), full support passes (no selective passes, no triangle listing)
    and no triangle listing.  Median wall time of `steps` queries after one
    warm-up query.  value = undirected edges / s of the scale-20 graph at
    k = 8 with the defaults."""
    from graphalgo_amd import api, graphs as G, serving as S
    pool = {c["scale"]: c for c in S.pool_configs()}
    graphs = [("rmat12_ef16_s5", dict(scale=12, edge_factor=16, seed=5)),
              ("rmat14_ef16_s5", dict(scale=14, edge_factor=16, seed=5)),
              ("rmat16_ef16_s1", dict(scale=16, edge_factor=16, seed=1))]
    scales = {int(x) for x in args.ktruss_scales.split(",")} if args.ktruss_scales else None
    for sc in (18, 20):
        c = pool[sc]
        graphs.append((f"rmat{sc}_ef{c['edge_factor']}_s{c['seed']}",
                       dict(scale=sc, edge_factor=c["edge_factor"], seed=c["seed"])))
    variants = {"default": {}, "host_loop": {"truss_loop": 0}, "device_steps": {"truss_loop": 1},
                "graph_replay": {"truss_loop": 2}, "full_passes": {"selective_support": 0, "support_tri": 0},
                "no_tri": {"support_tri": 0}, "tri_full_passes": {"selective_support": 0, "support_tri": 1}}
    if args.ktruss_variants:
        variants = {k: v for k, v in variants.items() if k in args.ktruss_variants.split(",")}
    if scales:
        graphs = [(nm, cfg) for nm, cfg in graphs if cfg["scale"] in scales]
    rows, headline = [], None
    for name, cfg in graphs:
        g = G.rmat(**cfg)
        R = api.Graph(g)
        for k in (3, 8):
            for vname, opts in variants.items():
                with api.options(**opts):
                    R.ktruss(k)
                    ts = []
                    for _ in range(max(1, args.steps)):
                        t0 = time.perf_counter()
                        nnz, passes = R.ktruss(k)
                        ts.append(time.perf_counter() - t0)
                ms = float(np.median(ts)) * 1e3
                rows.append({"graph": name, "nnz": g.nnz, "k": k, "variant": vname, "ms": round(ms, 3),
                             "truss_nnz": nnz, "passes": passes})
                if cfg["scale"] == 20 and k == 8 and vname == "default":
                    headline = (g.m_undirected, ms)
            _log(f"ktruss_one {name} k={k}: " + ", ".join(
                f"{r['variant']} {r['ms']:.2f}" for r in rows if r["graph"] == name and r["k"] == k))
        R.close()
    if headline is None:  # a subset without the headline case: the first case
        headline = (rows[0]["nnz"] // 2, rows[0]["ms"]) if rows else (0, float("nan"))
    m, ms = headline
    return {"value": m / (ms / 1e3), "unit": "edges/s", "ms_per_step": ms, "dtype": "int32",
            "config": {"workload": "k-truss for one k on resident R-MAT graphs (scale 12..20), each peeling "
                                   "setting forced in turn; headline: the served pool's scale-20 graph, k = 8, "
                                   "default settings", "queries_per_case": max(1, args.steps)},
            "roofline": None,
            "roofline_note": "wall time of whole queries (many launches and host steps); the support kernel's "
                             "roofline is the ktruss workload's",
            "cases": rows}


WORKLOADS = {"ktruss": bench_ktruss, "tc": bench_tc, "tc_er": bench_tc_er, "copy": bench_copy, "ewise": bench_ewise,
             "batch": bench_batch, "bc": bench_bc, "serve": bench_serve, "ktruss_one": bench_ktruss_one}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ktruss", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--queries", type=int, default=400, help="serve: queries per stream")
    ap.add_argument("--qps", type=float, default=0.0, help="serve: offered load (0 = 0.8 x served capacity)")
    ap.add_argument("--slots", type=int, default=4, help="serve: queries in flight")
    ap.add_argument("--budget-gb", type=int, default=96, help="serve: device arena size")
    ap.add_argument("--max-batch", type=int, default=16, help="serve: queued queries one run may take")
    ap.add_argument("--ktruss-scales", default="", help="ktruss_one: R-MAT scales to run (comma list; all)")
    ap.add_argument("--ktruss-variants", default="", help="ktruss_one: settings to run (comma list; all)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); sharding over several "
              "GPUs is designed (row blocks by wedges, BC by source columns) but not built", file=sys.stderr)
        sys.exit(2)
    res = WORKLOADS[args.workload](args)
    line = {
        "metric": METRIC, "value": res["value"], "unit": res["unit"], "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res["dtype"],
        "data": "synthetic (seeded generator)", "config": res["config"], "roofline": res["roofline"],
    }
    if "cpu_baseline" in res:
        line["cpu_baseline"] = res["cpu_baseline"]
    for k, v in res.items():
        if k not in line and k not in ("value", "unit", "ms_per_step", "dtype"):
            line[k] = v
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
