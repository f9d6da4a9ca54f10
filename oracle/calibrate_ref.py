"""ORACLE calibration -- test infrastructure only.

Times the compiled reference (oracle/_ref/ref_driver, built from the
reference tree by oracle/Makefile) and the C restatement (oracle.c) on
identical inputs with the same OpenMP thread count, and prints the ratios as
one JSON line.  The restatement is what the benchmark's cpu_baseline times on
the GPU host (the reference does not travel there); this ratio says how far
that timing sits from the reference's own code on the same machine.

  python oracle/calibrate_ref.py [threads]

Cases: triangle count C<L>=L*L (MSA, src/main.cpp:65-86) on R-MAT scale 16
and 18 and on the Erdos-Renyi graph n=1000, m=9990; k-truss for one k
(src/main.cpp:89-146) on R-MAT scale 14.  Both sides exclude graph loading
and lower-triangle extraction; the reference's figure is its own timer's.
One warm-up, then the median of 5 runs per side.  With OMP_PROC_BIND set,
the reference's process is started with every CPU this one may use (a child
inherits the affinity of the thread that starts it, which the OpenMP runtime
has bound to one core by then).
"""
from __future__ import annotations

import json
import os
import sys
import tempfile
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent))


ALL_CPUS = os.sched_getaffinity(0)


def ref_run(*args: str) -> dict:
    import subprocess
    import oracle as O
    out = subprocess.run([str(O.REF), *args], capture_output=True, text=True, check=True,
                         preexec_fn=lambda: os.sched_setaffinity(0, ALL_CPUS))
    return json.loads(out.stdout.strip().splitlines()[-1])


def main() -> None:
    threads = sys.argv[1] if len(sys.argv) > 1 else "8"
    os.environ["OMP_NUM_THREADS"] = threads
    import numpy as np
    import oracle as O
    from graphalgo_amd import graphs as G
    if not O.ref_available():
        raise SystemExit("oracle/_ref/ref_driver is not built (the reference tree is absent)")
    O.lib()
    cases = [("tc", "rmat16_ef16_s1", G.rmat(16, 16, seed=1), None),
             ("tc", "rmat18_ef16_s1", G.rmat(18, 16, seed=1), None),
             ("tc", "er_1000_9990_s1", G.erdos_renyi(1000, 9990, seed=1), None),
             ("ktruss", "rmat14_ef16_s1", G.rmat(14, 16, seed=1), 8)]
    rows = []
    with tempfile.TemporaryDirectory() as d:
        for kind, name, g, k in cases:
            path = str(Path(d) / f"{name}.csr")
            g.write_csr(path)
            if kind == "tc":
                ref = [ref_run("tc", path, "msa") for _ in range(6)][1:]
                want = ref[0]["triangles"]
                fn = lambda: O.tc(g)  # noqa: E731
            else:
                ref = [ref_run("ktruss", path, str(k)) for _ in range(6)][1:]
                want = ref[0]["nnz"]
                fn = lambda: len(O.ktruss(g, k)[1])  # noqa: E731
            got = fn()
            assert got == want, (name, got, want)
            ts = []
            for _ in range(5):
                t0 = time.perf_counter()
                fn()
                ts.append((time.perf_counter() - t0) * 1e3)
            ref_ms = float(np.median([r["ms"] for r in ref]))
            port_ms = float(np.median(ts))
            rows.append({"case": f"{kind} {name}" + (f" k={k}" if k else ""), "reference_ms": round(ref_ms, 3),
                         "restatement_ms": round(port_ms, 3), "reference_over_restatement": round(ref_ms / port_ms, 3),
                         "result": want})
    print(json.dumps({"threads": int(threads), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "cases": rows}))


if __name__ == "__main__":
    main()
