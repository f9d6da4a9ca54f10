"""Benchmark of the masked-SpGEMM path on one GPU.

The default run measures every workload in one process and prints one JSON
line.  Its headline (value, roofline, cpu_baseline at the top level) is the
truss decomposition (k-truss for every k until empty) of R-MAT scale 16, edge
factor 16 (a=.57 b=.19 c=.19 d=.05, symmetrised, deduplicated, loops
removed): one step = one full decomposition of the HBM-resident graph with
the truss numbers read back to the host; value = undirected edges / seconds
per step.  The line's "workloads" object holds the others, each with its own
value, roofline and cpu_baseline:

  tc_er_n1000     triangle count C<L>=L*L of the Erdos-Renyi graph n=1000,
                  m=9990 (latency-bound single launch)
  bc_rmat18_256   batched BC, 256 seeded random sources, R-MAT scale 18
  batch_tc_knn4096  triangle counts of 4096 kNN graphs in one launch
  serve_rmat_pool served stream over eleven resident R-MAT graphs of scale
                  10..20 (p50 / p99 latency and throughput, vs the plain path)
  mxm_rmat16      C<A>=A*A on R-MAT scale 16 (the reference's mxm action)

The dominant kernel's roofline entry uses algorithmic bytes per launch

  4(m+1) + 4 nnz(A) + sum_{(i,k) in A} (8 + 4 nnz(B_k)) + 4(m+1) + 4 nnz(M) + out

(values implicit ones) divided by the GPU time of its launches, measured
with HIP events on the stream the kernels run on; "traffic" is the HBM bytes
per launch rocprofv3's FETCH_SIZE + WRITE_SIZE counted (scaled by a copy
kernel of known size) and "traffic_frac" that traffic's rate against the
peak.  CPU baselines are the OpenMP oracle on the host cores, timed in a
child process of its own (OMP_PROC_BIND=spread, OMP_PLACES=cores, one
warm-up, median of several runs) that regenerates the same seeded inputs
and never touches the GPU; the GPU results are checked against the oracle
in this process.

--workload NAME runs one workload alone (also: tc, tc_wide, ktruss_one, ewise, copy).
"""
from __future__ import annotations

import os
import argparse
import gc
import json
import subprocess
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


_HOST = None
_HOST_CPUS = None


def host_info() -> dict:
    """CPU model, sockets, physical cores and the cores this process may use."""
    global _HOST, _HOST_CPUS
    if _HOST is None:
        try:
            _HOST_CPUS = os.sched_getaffinity(0)
        except (AttributeError, OSError):
            _HOST_CPUS = None
        info = {}
        try:
            out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
            for line in out.splitlines():
                k, _, v = line.partition(":")
                info[k.strip()] = v.strip()
        except (OSError, subprocess.SubprocessError):
            pass

        def num(key):
            try:
                return int(info.get(key, ""))
            except ValueError:
                return None
        sockets, cps, tpc = num("Socket(s)"), num("Core(s) per socket"), num("Thread(s) per core")
        _HOST = {"cpu_model": info.get("Model name"), "sockets": sockets,
                 "physical_cores": sockets * cps if sockets and cps else None, "threads_per_core": tpc,
                 "logical_cpus": num("CPU(s)"), "cpus_available": _cpu_cores(),
                 "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
                 "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    return _HOST


def _oracle():
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    O.lib()
    return O


def _cpu_median(fn, reps: int, warm=None) -> tuple[float, list[float]]:
    """One warm-up call (warm, or fn itself), then the median of reps timed calls."""
    (warm or fn)()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def _cpu_entry(value: float, unit: str, sample: str, ts: list[float]) -> dict:
    O = _oracle()
    return {"value": value, "unit": unit, "cores": O.num_threads(), "kind": "port", "sample": sample,
            "runs_s": [round(t, 4) for t in ts], "statistic": f"median of {len(ts)} after one warm-up",
            "host": host_info()}


def cpu_baseline(name: str, args) -> dict:
    """Runs the CPU leg `name` (a key of CPU_LEGS) in a child process with the
    OpenMP threads spread over the cores, and returns its cpu_baseline entry.
    A child of its own because binding the OpenMP threads also pins the
    thread that starts them, which would confine this process's later
    threads (graph generation, the HIP runtime's) to one core."""
    env = dict(os.environ, OMP_PROC_BIND="spread", OMP_PLACES="cores")
    cmd = [sys.executable, str(Path(__file__).resolve()), "--cpu-leg", name, "--queries", str(args.queries)]
    proc = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if proc.returncode != 0:
        raise RuntimeError(f"CPU leg {name} failed:\n{proc.stderr[-2000:]}")
    return json.loads(proc.stdout.strip().splitlines()[-1])


def _measure(step, steps: int, warmup: int):
    """Runs `warmup` untimed steps, then `steps` timed steps with no hooks
    (seconds per step: the reported value), then `steps` more with HIP events
    around every launch of the dominant kernel on the stream it runs on (the
    roofline's kernel time), then one extra step with the algorithmic byte
    accounting on (every step issues the same launches).
    Returns (seconds per step, last step result, timing, accounting)."""
    from graphalgo_amd import api
    for _ in range(warmup):
        step()
    _sync()
    # no collector pass inside the timed loops: a full collection over the
    # objects earlier workloads left (thousands of graphs) took ~20 ms
    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter()
        out = None
        for _ in range(steps):
            out = step()
        _sync()
        dt = (time.perf_counter() - t0) / steps
        api.prof_enable(api.PROF_EVENTS)
        api.prof_read(reset=True)
        t1 = time.perf_counter()
        for _ in range(steps):
            step()
        _sync()
        dt_events = (time.perf_counter() - t1) / steps
    finally:
        gc.enable()
    timing = api.prof_read(reset=True)
    api.prof_enable(api.PROF_BYTES)
    step()
    acct = api.prof_read(reset=True)
    api.prof_enable(0)
    timing["steps"] = steps
    timing["ms_per_step_with_events"] = dt_events * 1e3
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
    """achieved / frac: algorithmic bytes per launch / average launch time (the
    rate the kernel would need from HBM if every algorithmic byte came from
    there); traffic_frac: the measured HBM bytes per launch at the same launch
    time, against the same peak (below frac when the graph stays in L2/MALL)."""
    steps = max(timing.get("steps", 1), 1)
    launches = max(timing["launches"] // steps, 1)
    kernel_s = timing["kernel_ms"] / steps / 1e3
    achieved = acct["bytes_alg"] / kernel_s / 1e9 if kernel_s > 0 else 0.0
    traffic, source = _traffic(traffic_key) if traffic_key else (None, None)
    avg_s = kernel_s / launches
    traffic_gbs = traffic / avg_s / 1e9 if traffic and avg_s > 0 else None
    return {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": source,
        "frac_basis": "algorithmic bytes per launch / average launch time",
        "traffic_gbs": None if traffic_gbs is None else round(traffic_gbs, 1),
        "traffic_frac": None if traffic_gbs is None else round(traffic_gbs / HBM_PEAK_GBS, 4),
        "kernel": kernel, "launches_per_step": launches,
        "avg_launch_us": round(kernel_s * 1e6 / launches, 2),
        "bytes_alg_per_launch": int(acct["bytes_alg"] / max(acct["accounted_launches"], 1)),
        "ms_per_step_with_events": round(timing.get("ms_per_step_with_events", 0.0), 3),
        "wedges_per_step": acct["wedges"], "wedges_processed_per_step": acct["wedges_processed"],
    }


def bench_ktruss(args) -> dict:
    from graphalgo_amd import api, graphs as G
    g = G.rmat(16, 16, seed=1)
    m = g.m_undirected
    R = api.Graph(g)
    # a step ends with the truss number of every stored entry on the host
    dt, (truss, kmax, stats), timing, acct = _measure(lambda: R.truss(want_values=True), args.steps, args.warmup)
    R.close()
    digest = _sha256_i32(truss)
    out = {
        "value": m / dt, "unit": "edges/s", "ms_per_step": dt * 1e3, "dtype": "int32",
        "config": {"workload": "truss decomposition (k-truss for all k) on R-MAT scale 16, truss numbers read back",
                   "graph": "rmat(scale=16, edge_factor=16, a=.57, b=.19, c=.19, seed=1)",
                   "n": g.n, "m_undirected": m, "nnz": g.nnz, "kmax": kmax,
                   "support_passes": stats.get("passes"), "filter_steps": stats.get("steps"),
                   "truss_sha256": digest},
        "roofline": _roofline(timing, acct, "k_peel_tick (a peeling round: the triangles through the removed "
                                            "edges, supports of the survivors decremented; or the scan of the next "
                                            "threshold); the one full support product per step is the tc workload's "
                                            "kernel", "k_peel_tick/truss_rmat16"),
    }
    golden = _golden_truss_rmat16()
    if golden:
        out["config"]["truss_sha256_matches_reference"] = digest == golden
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline("ktruss", args)
    return out


def _sha256_i32(a) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int32).tobytes()).hexdigest()


def _golden_truss_rmat16():
    """SHA-256 of the reference's per-entry truss numbers of the same graph
    (tests/golden/golden.json), when the fixture is present."""
    try:
        gold = json.loads((ROOT / "tests" / "golden" / "golden.json").read_text())
        return gold["generated"]["rmat16_ef16_s1"]["truss"]["truss_sha256"]
    except (OSError, KeyError, ValueError, TypeError):
        return None


def cpu_ktruss(args) -> dict:
    from graphalgo_amd import graphs as G
    O = _oracle()
    g = G.rmat(16, 16, seed=1)
    warm = G.rmat(14, 16, seed=1)
    med, ts = _cpu_median(lambda: O.truss_decomposition(g, max_passes=1), 3,
                          warm=lambda: O.truss_decomposition(warm, max_passes=1))
    return _cpu_entry(g.m_undirected / med, "edges/s",
                      "full truss decompositions of the same R-MAT 16 graph by the OpenMP oracle (warm-up: "
                      f"R-MAT 14), median {med:.2f} s", ts)


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
        O = _oracle()
        want = O.tc(g)
        assert t == want, f"GPU triangle count {t} differs from the oracle's {want}"
        out["config"]["oracle_triangles_equal"] = True
        out["cpu_baseline"] = cpu_baseline("tc", args)
    return out


def bench_tc_wide(args) -> dict:
    """Triangle counts C<L> = L*L on R-MAT graphs of 2^18..2^20 vertices: the
    survey's R-MAT 18 probe family (edge factor 16, seed 1) and the served
    pool's scale 18, 19 and 20 graphs, resident, one call per step; each case
    with its wedges, kernel roofline and the oracle's count.  value =
    undirected edges / s of R-MAT 18 (edge factor 16)."""
    from graphalgo_amd import api, graphs as G, serving as S
    pool = {c["scale"]: c for c in S.pool_configs()}
    cases = [("rmat18_ef16_s1", dict(scale=18, edge_factor=16, seed=1))]
    for sc in (18, 19, 20):
        c = pool[sc]
        cases.append((f"rmat{sc}_ef{c['edge_factor']}_s{c['seed']}",
                      dict(scale=sc, edge_factor=c["edge_factor"], seed=c["seed"])))
    O = None if args.no_cpu else _oracle()
    rows, head = [], None
    for name, cfg in cases:
        g = G.rmat(**cfg)
        R = api.Graph(g)
        dt, t, timing, acct = _measure(R.triangle_count, args.steps, args.warmup)
        R.close()
        rl = _roofline(timing, acct, "k_mxm_items (C<L> = L*L, sum mode)")
        row = {"graph": name, "n": g.n, "m_undirected": g.m_undirected, "ms": round(dt * 1e3, 3),
               "edges_per_s": g.m_undirected / dt, "triangles": t, "wedges": acct["wedges"],
               "g_wedges_per_s": round(acct["wedges"] / dt / 1e9, 1), "kernel_us": rl["avg_launch_us"],
               "frac": rl["frac"]}
        if O is not None:
            want = O.tc(g)
            assert t == want, f"{name}: GPU triangle count {t} differs from the oracle's {want}"
            row["oracle_triangles_equal"] = True
        rows.append(row)
        _log(f"tc_wide {name}: {row['ms']} ms, {row['g_wedges_per_s']} G wedges/s, frac {row['frac']}")
        if head is None:
            head = (g.m_undirected, dt, rl)
    m, dt, rl = head
    return {"value": m / dt, "unit": "edges/s", "ms_per_step": dt * 1e3, "dtype": "int32",
            "config": {"workload": "triangle count C<L>=L*L on R-MAT graphs of scale 18..20 (headline: scale 18, "
                                   "edge factor 16, seed 1)", "cases": rows},
            "roofline": rl}


def cpu_tc(args) -> dict:
    from graphalgo_amd import graphs as G
    O = _oracle()
    g = G.rmat(16, 16, seed=1)
    med, ts = _cpu_median(lambda: O.tc(g), 5)
    return _cpu_entry(g.m_undirected / med, "edges/s",
                      f"oracle triangle counts of the same graph, median {med * 1e3:.1f} ms", ts)


def bench_tc_er(args) -> dict:
    """Triangle count C<L> = L*L on the Erdos-Renyi graph n = 1000, m = 9990
    (seed 1): a latency-bound single launch; value = edges/s of one call on the
    resident graph, with the per-call latency beside it."""
    from graphalgo_amd import api, graphs as G
    g = G.erdos_renyi(1000, 9990, seed=1)
    R = api.Graph(g)
    # a call is ~20 us: at least 200 timed calls (4 ms) for a stable mean
    calls = max(args.steps, 200)
    dt, t, timing, acct = _measure(R.triangle_count, calls, max(args.warmup, 20))
    R.close()
    out = {"value": g.m_undirected / dt, "unit": "edges/s", "ms_per_step": dt * 1e3, "dtype": "int32",
           "config": {"workload": "triangle count C<L>=L*L on Erdos-Renyi n=1000 m=9990 (seed 1)", "n": g.n,
                      "m_undirected": g.m_undirected, "triangles": t, "latency_us": round(dt * 1e6, 2),
                      "calls_timed": calls,
                      "note": "latency-bound: one call moves ~0.43 MB of algorithmic bytes (54 ns at 8 TB/s)"},
           "roofline": _roofline(timing, acct, "k_sum_wave (C<L> = L*L, sum mode)", "k_sum_wave/tc_er_n1000")}
    if not args.no_cpu:
        O = _oracle()
        want = O.tc(g)
        assert t == want, f"GPU triangle count {t} differs from the oracle's {want}"
        out["config"]["oracle_triangles_equal"] = True
        out["cpu_baseline"] = cpu_baseline("tc_er", args)
    return out


def cpu_tc_er(args) -> dict:
    from graphalgo_amd import graphs as G
    O = _oracle()
    g = G.erdos_renyi(1000, 9990, seed=1)

    def many():  # one timed sample = 1000 counts (a single count is ~50 us)
        for _ in range(1000):
            O.tc(g)
    med, ts = _cpu_median(many, 5)
    return _cpu_entry(1000 * g.m_undirected / med, "edges/s",
                      f"1000 oracle triangle counts of the same graph per run ({med * 1e3:.0f} us each)", ts)


def bench_mxm(args) -> dict:
    """C<A> = A*A on R-MAT scale 16 (the reference's mxm action,
    src/main.cpp:235-240), int32 values of ones: host-pointer entry point, so
    value = undirected edges per second of device time (HIP events around the
    product launches); the host-inclusive rate (PCIe copies, allocations) is
    beside it."""
    from graphalgo_amd import api, graphs as G
    g = G.rmat(16, 16, seed=1)
    ones = np.ones(g.nnz, dtype=np.int32)

    def step():
        return api.mxm_masked(g, g, g, dtype=np.int32, Aval=ones, Bval=ones)
    dt, C, timing, acct = _measure(step, args.steps, args.warmup)
    dev_s = timing["kernel_ms"] / timing["steps"] / 1e3
    out = {"value": g.m_undirected / dev_s, "unit": "edges/s", "ms_per_step": dev_s * 1e3, "dtype": "int32",
           "config": {"workload": "C<A>=A*A on R-MAT scale 16, int32 values", "n": g.n, "nnz": g.nnz,
                      "sum_C": int(C.astype(np.int64).sum()), "host_inclusive_ms_per_call": dt * 1e3,
                      "host_inclusive_edges_per_s": g.m_undirected / dt},
           "roofline": _roofline(timing, acct, "k_mxm_items (C<A> = A*A, int32 values per entry)", "k_mxm_items/mxm_rmat16")}
    if not args.no_cpu:
        O = _oracle()
        want = O.mxm_masked(g.n, g.n, g.rowptr, g.col, g.rowptr, g.col, g.rowptr, g.col)
        assert np.array_equal(C.astype(np.int64), want), "C<A>=A*A differs from the oracle"
        out["config"]["oracle_equal"] = True
        out["cpu_baseline"] = cpu_baseline("mxm", args)
    return out


def cpu_mxm(args) -> dict:
    from graphalgo_amd import graphs as G
    O = _oracle()
    g = G.rmat(16, 16, seed=1)
    med, ts = _cpu_median(lambda: O.mxm_masked(g.n, g.n, g.rowptr, g.col, g.rowptr, g.col, g.rowptr, g.col), 5)
    return _cpu_entry(g.m_undirected / med, "edges/s",
                      f"oracle C<A>=A*A of the same graph, median {med * 1e3:.0f} ms", ts)


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
    _, tune_ms = B.autotune()
    shape = B.shape()
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
    roof = _roofline(timing, acct, "k_sum_wave + k_mxm_items (batched)", "k_sum_wave/batched_tc_knn4096")
    out = {"value": m_total / dt, "unit": "edges/s", "ms_per_step": dt * 1e3, "dtype": "int32",
           "config": {"workload": f"batched triangle count over {count} kNN graphs (k=30) in one launch",
                      "graphs": count, "m_total": m_total, "triangles_total": int(counts.sum()),
                      "autotuned": shape,
                      "autotune_ms": {(f"wave_flops={c}" if w == "wave_flops" else f"{w}x{c}"): round(v, 3)
                                      for (w, c), v in tune_ms.items()},
                      "generation_s": round(gen_s, 1)},
           "roofline": roof,
           "plain_path": {"value": m_total / dt_plain, "unit": "edges/s", "ms": dt_plain * 1e3,
                          "speedup_batched_vs_plain": dt_plain / dt}}
    if not args.no_cpu:
        O = _oracle()
        want = [O.tc(g) for g in graphs]
        assert counts.tolist() == want, "batched counts differ from the oracle"
        out["config"]["oracle_counts_equal"] = True
        out["cpu_baseline"] = cpu_baseline("batch", args)
    return out


def cpu_batch(args) -> dict:
    from graphalgo_amd import graphs as G
    O = _oracle()
    # the generator's worker threads inherit the affinity of the thread that
    # starts them, which the OpenMP runtime bound to one core: lift it while
    # they run
    bound = os.sched_getaffinity(0)
    os.sched_setaffinity(0, _HOST_CPUS or bound)
    sample = G.knn_batch(base_seed=20240, count=512, k=30, workers=min(16, len(_HOST_CPUS or bound)))
    os.sched_setaffinity(0, bound)

    def run():
        for g in sample:
            O.tc(g)
    med, ts = _cpu_median(run, 5)
    return _cpu_entry(sum(g.m_undirected for g in sample) / med, "edges/s",
                      f"oracle triangle counts of the first 512 of the batch's kNN graphs, median {med * 1e3:.0f} ms",
                      ts)


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
    roof = _roofline(timing, acct, "k_bc_forward_items (bottom-up or top-down per level) + k_bc_sigma_items + "
                     "k_bc_backward_items", "k_bc/bc_rmat18")
    roof["launches_note"] = ("one event pair per timed segment of a pass: the forward sweep (levels queued "
                             "on the device, one host wait at its end) and the backward sweep")
    out = {"value": m * len(src) / dt, "unit": "edges*sources/s", "ms_per_step": dt * 1e3, "dtype": "f64",
           "config": {"workload": "batched BC, 256 sources, R-MAT scale 18", "n": g.n, "m_undirected": m,
                      "sources": len(src), "source_seed": 7, "bc_sum": float(bc.sum())},
           "roofline": roof}
    if not args.no_cpu:
        O = _oracle()
        want = O.bc(g, src)
        rel = float((np.abs(bc - want) / np.maximum(np.abs(want), 1.0)).max())
        assert rel <= 1e-9, f"BC differs from the FP64 oracle (max relative {rel:.3e})"
        out["config"]["oracle_max_rel_diff"] = rel
        out["cpu_baseline"] = cpu_baseline("bc", args)
    return out


def cpu_bc(args) -> dict:
    from graphalgo_amd import graphs as G
    O = _oracle()
    g = G.rmat(18, 16, seed=1)
    sample = G.pick_sources(g, 256, seed=7)[:32]
    med, ts = _cpu_median(lambda: O.bc(g, sample), 5)
    return _cpu_entry(g.m_undirected * len(sample) / med, "edges*sources/s",
                      f"oracle FP64 batched BC on the same graph with the first 32 of the 256 sources, "
                      f"median {med:.2f} s", ts)


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
    for step in range(args.serve_replays):
        t0 = time.perf_counter()
        res = srv.replay(specs, arr)
        t_served += time.perf_counter() - t0
        served.append(S.latency_summary(res, arr))
        _log(f"served replay {step}: p50 {served[-1]['p50_ms']:.2f} ms p99 {served[-1]['p99_ms']:.2f} ms")
    plain_res = api.plain_replay(graphs, specs, arr)
    plain = S.latency_summary(plain_res, arr)
    _log(f"plain replay: p50 {plain['p50_ms']:.2f} ms p99 {plain['p99_ms']:.2f} ms")
    # a second operating point that both paths sustain: 0.8 x the plain capacity
    qps_low = round(0.8 * cap_plain, 1)
    specs_low, arr_low = S.make_stream(pool, nq, qps_low, seed=4)
    low_served = S.latency_summary(srv.replay(specs_low, arr_low), arr_low)
    low_plain = S.latency_summary(api.plain_replay(graphs, specs_low, arr_low), arr_low)
    _log(f"at {qps_low} QPS: served p50 {low_served['p50_ms']:.2f} p99 {low_served['p99_ms']:.2f} ms, "
         f"plain p50 {low_plain['p50_ms']:.2f} p99 {low_plain['p99_ms']:.2f} ms")
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
    out = {"value": last["throughput_qps"], "unit": "queries/s", "ms_per_step": 1e3 * t_served / args.serve_replays,
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
                               "capacity_without_shared_runs": cap_nobatch / cap_plain,
                               "p50_latency": plain["p50_ms"] / last["p50_ms"],
                               "p99_latency": plain["p99_ms"] / last["p99_ms"]},
           "at_plain_sustainable_load": {
               "qps_offered": qps_low, "stream_seed": 4,
               "served": {k: round(v, 3) for k, v in low_served.items()},
               "plain": {k: round(v, 3) for k, v in low_plain.items()},
               "p50_latency_ratio": low_plain["p50_ms"] / low_served["p50_ms"],
               "p99_latency_ratio": low_plain["p99_ms"] / low_served["p99_ms"]},
           "roofline_note": "queries of every kind run concurrently on 'slots' streams; per-kernel roofline "
                            "is reported by the ktruss, tc, bc and batch workloads"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline("serve", args)
    return out


def cpu_serve(args) -> dict:
    """The served stream's queries (the same seeded stream: its kinds, graphs,
    k and sources do not depend on the arrival rate) over the pool graphs of
    scale <= 15, answered back to back by the oracle."""
    from graphalgo_amd import serving as S
    O = _oracle()
    pool = S.make_pool(S.pool_configs(1))
    specs, _ = S.make_stream(pool, args.queries, 1.0, seed=3)
    small = [q for q in specs if pool[q["graph"]].n <= (1 << 15)]

    def run_one(q):
        g = pool[q["graph"]]
        if q["kind"] == "tc":
            O.tc(g)
        elif q["kind"] == "ktruss":
            O.ktruss(g, q["k"])
        else:
            O.bc(g, q["sources"])
    # the sample: the first of those queries that the oracle answers within
    # about 3 s (this first pass is the warm-up), then timed as a whole
    t0, sample = time.perf_counter(), []
    for q in small:
        run_one(q)
        sample.append(q)
        if time.perf_counter() - t0 > 3.0:
            break

    def run():
        for q in sample:
            run_one(q)
    med, ts = _cpu_median(run, 5, warm=lambda: None)
    return _cpu_entry(len(sample) / med, "queries/s",
                      f"oracle on the stream's first {len(sample)} queries over pool graphs of scale <= 15, "
                      f"back to back, median {med:.2f} s", ts)


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
    variants = {"default": {}, "peel_host_steered": {"truss_loop": 4}, "host_loop": {"truss_loop": 0},
                "device_steps": {"truss_loop": 1},
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


WORKLOADS = {"ktruss": bench_ktruss, "tc": bench_tc, "tc_er": bench_tc_er, "tc_wide": bench_tc_wide,
             "copy": bench_copy, "ewise": bench_ewise,
             "batch": bench_batch, "bc": bench_bc, "serve": bench_serve, "ktruss_one": bench_ktruss_one,
             "mxm": bench_mxm}

# the workloads of the default run, after the headline truss decomposition
ALL = (("tc_er_n1000", "tc_er"), ("bc_rmat18_256", "bc"), ("batch_tc_knn4096", "batch"),
       ("serve_rmat_pool", "serve"), ("mxm_rmat16", "mxm"))


def bench_all(args) -> dict:
    """The headline (truss decomposition of R-MAT 16) plus every other
    workload, each measured with the same steps / warm-up (the served stream
    with its own replay count) and reported under "workloads"."""
    t0 = time.perf_counter()
    _log("truss decomposition, R-MAT 16")
    res = bench_ktruss(args)
    _log(f"  {res['ms_per_step']:.1f} ms/step ({time.perf_counter() - t0:.0f} s)")
    res["workloads"] = {}
    for key, name in ALL:
        _log(f"{key}")
        r = WORKLOADS[name](args)
        entry = {"value": r["value"], "unit": r["unit"], "ms_per_step": r["ms_per_step"], "dtype": r["dtype"]}
        entry.update({k: v for k, v in r.items() if k not in entry})
        res["workloads"][key] = entry
        _log(f"  {key}: {r['value']:.4g} {r['unit']} ({time.perf_counter() - t0:.0f} s)")
    res["bench_wall_s"] = round(time.perf_counter() - t0, 1)
    return res


WORKLOADS["all"] = bench_all

CPU_LEGS = {"ktruss": cpu_ktruss, "tc": cpu_tc, "tc_er": cpu_tc_er, "mxm": cpu_mxm, "batch": cpu_batch,
            "bc": cpu_bc, "serve": cpu_serve}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="all", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--queries", type=int, default=400, help="serve: queries per stream")
    ap.add_argument("--qps", type=float, default=0.0, help="serve: offered load (0 = 0.8 x served capacity)")
    ap.add_argument("--slots", type=int, default=4, help="serve: queries in flight")
    ap.add_argument("--budget-gb", type=int, default=96, help="serve: device arena size")
    ap.add_argument("--max-batch", type=int, default=16, help="serve: queued queries one run may take")
    ap.add_argument("--serve-replays", type=int, default=2, help="serve: timed replays of the stream")
    ap.add_argument("--ktruss-scales", default="", help="ktruss_one: R-MAT scales to run (comma list; all)")
    ap.add_argument("--ktruss-variants", default="", help="ktruss_one: settings to run (comma list; all)")
    ap.add_argument("--cpu-leg", default="", choices=[""] + sorted(CPU_LEGS),
                    help="internal: time one CPU baseline (oracle only, no GPU) and print its entry")
    args = ap.parse_args()
    if args.cpu_leg:
        host_info()  # before the OpenMP runtime binds this thread to one core
        print(json.dumps(CPU_LEGS[args.cpu_leg](args)), flush=True)
        return
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
