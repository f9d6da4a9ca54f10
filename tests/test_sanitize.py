"""AddressSanitizer + UBSan build of the oracle (oracle.c) and the host graph
library (graphgen.cpp), driven over generated graphs and a Matrix Market
fixture by oracle/sanitize_main.c (host code only; no GPU code is built
with a sanitizer)."""
import os
import subprocess

from conftest import GRAPH_DIR, ROOT


def test_oracle_and_host_library_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"], check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", OMP_NUM_THREADS="4",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(ROOT / "oracle" / "_build" / "sanitize"), str(GRAPH_DIR / "yeast_30NN.mtx.gz")],
                       capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert r.stdout.count("triangles=") == 4
