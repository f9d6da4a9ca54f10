"""GPU parity of batched betweenness centrality (FP64, 1e-9 relative) and of
the batched multi-graph triangle count (bit-exact)."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9


def _bc_close(got, want):
    scale = max(1.0, float(np.abs(want).max()))
    return np.allclose(got, want, rtol=BC_RTOL, atol=BC_RTOL * scale * 1e-3)


def _complete(G, n):
    u, v = zip(*itertools.combinations(range(n), 2))
    return G.CSR.from_edges(n, u, v)


@pytest.mark.parametrize("name", ["G43.mtx", "iris_dataset_30NN.mtx", "yeast_30NN.mtx.gz"])
def test_bc_fixtures_vs_oracle(gpu, small_graphs, oracle, name):
    g = small_graphs[name][0]
    for src in (np.arange(16), np.random.default_rng(1).choice(g.n, 100, replace=False)):
        print(f"case bc {name} ns={len(src)}", flush=True)
        got = gpu.betweenness_centrality_batch(g, src)
        assert _bc_close(got, oracle.bc(g, src))


def test_bc_deep_bfs_widens_depths(gpu, G):
    """A path of 600 vertices has BFS levels beyond the 8-bit depth range: the
    pass is redone with 16-bit depths and still meets the closed form."""
    n = 600
    path = G.CSR.from_edges(n, range(n - 1), range(1, n))
    print("case bc path600", flush=True)
    src = np.array([0, n // 3, n - 1], dtype=np.int32)
    bc = gpu.betweenness_centrality_batch(path, src)
    v = np.arange(n, dtype=np.float64)
    # each source s adds (number of vertices beyond v, seen from s) for v strictly between
    want = np.zeros(n)
    for s_ in src:
        want += np.where(v > s_, n - 1 - v, 0) + np.where(v < s_, v, 0)
    assert np.allclose(bc, want, rtol=1e-12)


def test_bc_closed_forms(gpu, G):
    n = 70  # more sources than one 64-lane chunk
    star = G.CSR.from_edges(n, [0] * (n - 1), range(1, n))
    bc = gpu.betweenness_centrality_batch(star, np.arange(1, n))
    assert bc[0] == pytest.approx((n - 1) * (n - 2), rel=1e-12) and np.allclose(bc[1:], 0)
    path = G.CSR.from_edges(n, range(n - 1), range(1, n))
    bc = gpu.betweenness_centrality_batch(path, np.arange(n))
    v = np.arange(n)
    assert np.allclose(bc, 2 * v * (n - 1 - v), rtol=1e-12)
    assert np.allclose(gpu.betweenness_centrality_batch(_complete(G, 9), np.arange(9)), 0)


def test_bc_isolated_and_disconnected(gpu, G, oracle):
    g = G.CSR.from_edges(10, [0, 1, 5, 6], [1, 2, 6, 7])  # two paths + isolated vertices
    src = [0, 3, 5, 9]
    assert _bc_close(gpu.betweenness_centrality_batch(g, src), oracle.bc(g, src))


def test_bc_rejects_duplicate_sources(gpu, G):
    from graphalgo_amd._native import GraphAlgoError
    g = G.CSR.from_edges(4, [0, 1], [1, 2])
    with pytest.raises(GraphAlgoError):
        gpu.betweenness_centrality_batch(g, [1, 1])


def test_bc_rmat_256_sources_resident(gpu, G, oracle):
    """256 seeded-random non-isolated sources as in the benchmark config, at scale 14."""
    g = G.rmat(14, 16, seed=14)
    deg = g.degrees()
    src = np.random.default_rng(7).choice(np.nonzero(deg)[0], 256, replace=False)
    print("case bc rmat14 ns=256", flush=True)
    R = gpu.Graph(g)
    got = R.bc(src)
    again = R.bc(src)
    R.close()
    assert np.array_equal(got, again)  # deterministic run to run
    assert _bc_close(got, oracle.bc(g, src))


def test_batch_tc_knn(gpu, G, oracle):
    graphs = G.knn_batch(base_seed=1000, count=24)
    want = np.array([oracle.tc(g) for g in graphs])
    B = gpu.Batch(graphs)
    for threads in (64, 128, 256):
        print(f"case batch tc threads={threads}", flush=True)
        assert np.array_equal(B.triangle_counts(threads), want)
    (width, claim), ms = B.autotune()
    assert width in (128, 256) and claim in (8, 16, 32, 64) and len(ms) == 12
    shape = B.shape()
    assert (shape["item_width"], shape["rows_per_claim"]) == (width, claim)
    assert shape["wave_flops"] in (2048, 16384, 65536, 1 << 20, gpu.get_option("wave_flops"))
    assert np.array_equal(B.triangle_counts(), want)
    B.close()


def test_batch_tc_mixed_and_empty(gpu, G, oracle, small_graphs):
    graphs = [small_graphs["iris_dataset_30NN.mtx"][1], G.CSR.from_edges(3, [], []), _complete(G, 30),
              small_graphs["G43.mtx"][1], G.CSR.from_edges(1, [], []), G.rmat(10, 8, seed=2)]
    B = gpu.Batch(graphs)
    got = B.triangle_counts()
    B.close()
    assert got.tolist() == [oracle.tc(g) for g in graphs]
