"""GPU parity of the masked SpGEMM, triangle counting and k-truss (through the C ABI).

Integer results must be bit-identical to the oracle and to the reference's
recorded outputs (tests/golden/golden.json); FP64 masked products within
1e-12 relative (only their summation order differs).
"""
import hashlib
import itertools

import numpy as np
import pytest

from conftest import GOLDEN_DIR

pytestmark = pytest.mark.gpu


def _csr_from_dense(D):
    rp = np.concatenate([[0], np.cumsum((D != 0).sum(1))]).astype(np.int32)
    r, c = np.nonzero(D)
    return rp, c.astype(np.int32), D[r, c]


def _complete(G, n):
    u, v = zip(*itertools.combinations(range(n), 2))
    return G.CSR.from_edges(n, u, v)


# ------------------------------------------------------------- masked SpGEMM
@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float64])
@pytest.mark.parametrize("seed", range(4))
def test_mxm_random_small_valued(gpu, G, oracle, dtype, seed):
    rng = np.random.default_rng(100 + seed)
    m, k, n = (int(x) for x in rng.integers(1, 40, size=3))
    A = (rng.random((m, k)) < 0.3) * rng.integers(-5, 6, size=(m, k))
    B = (rng.random((k, n)) < 0.3) * rng.integers(-5, 6, size=(k, n))
    Mk = (rng.random((m, n)) < 0.4).astype(np.int64)
    Arp, Acol, Aval = _csr_from_dense(A)
    Brp, Bcol, Bval = _csr_from_dense(B)
    Mrp, Mcol, _ = _csr_from_dense(Mk)
    ga = G.CSR(m, Arp, Acol)
    gb = G.CSR(k, Brp, Bcol)
    gm = G.CSR(m, Mrp, Mcol)
    print(f"case mxm seed={seed} dtype={np.dtype(dtype).name} m,k,n={m},{k},{n}", flush=True)
    got = gpu.mxm_masked(ga, gb, gm, dtype=dtype, Aval=Aval.astype(dtype), Bval=Bval.astype(dtype), shapes=(m, k, n))
    want = oracle.mxm_masked(m, n, Arp, Acol, Brp, Bcol, Mrp, Mcol, Aval=Aval, Bval=Bval,
                             dtype=np.float64 if dtype == np.float64 else np.int64)
    if dtype == np.float64:
        assert np.allclose(got, want, rtol=1e-12, atol=0)
    else:
        assert np.array_equal(got.astype(np.int64), want)


@pytest.mark.parametrize("scale,ef", [(10, 8), (14, 16)])
def test_mxm_rmat_square(gpu, G, oracle, scale, ef):
    """C<A> = A*A (the k-truss support product) on R-MAT: covers hub rows split into
    many items and mask rows longer than the LDS-staged capacity."""
    g = G.rmat(scale, ef, seed=scale)
    print(f"case mxm rmat scale={scale} nnz={g.nnz} maxdeg={g.degrees().max()}", flush=True)
    got = gpu.mxm_masked(g, g, g, dtype=np.int32)
    want = oracle.mxm_masked(g.n, g.n, g.rowptr, g.col, g.rowptr, g.col, g.rowptr, g.col)
    assert np.array_equal(got.astype(np.int64), want)


def test_mxm_closed_form_complete(gpu, G):
    n = 300  # rows longer than the 256-entry item width
    L = _complete(G, n).lower()
    got = gpu.mxm_masked(L, L, L, dtype=np.int64)
    rows = np.repeat(np.arange(n), np.diff(L.rowptr))
    assert np.array_equal(got, rows - L.col - 1)


def test_mxm_empty_and_ragged(gpu, G, oracle):
    empty = G.CSR(7, np.zeros(8, dtype=np.int32), np.zeros(0, dtype=np.int32))
    assert gpu.mxm_masked(empty, empty, empty).size == 0
    # mask present, A empty -> explicit zeros
    g = G.CSR.from_edges(6, [0, 1, 2], [1, 2, 3])
    a_empty = G.CSR(6, np.zeros(7, dtype=np.int32), np.zeros(0, dtype=np.int32))
    assert np.array_equal(gpu.mxm_masked(a_empty, g, g), np.zeros(g.nnz, dtype=np.int64))
    # rows of M empty while A is not: nothing is written
    star = G.CSR.from_edges(6, [0] * 5, [1, 2, 3, 4, 5])
    assert np.array_equal(gpu.mxm_masked(star, star, a_empty), np.zeros(0, dtype=np.int64))


def test_mxm_rejects_bad_input(gpu, G):
    from graphalgo_amd._native import GraphAlgoError
    bad = G.CSR(3, np.array([0, 2, 3, 3], dtype=np.int32), np.array([2, 1, 0], dtype=np.int32))  # unsorted row
    with pytest.raises(GraphAlgoError) as e:
        gpu.mxm_masked(bad, bad, bad)
    assert e.value.code == 3


# ------------------------------------------------------------- triangle counting
def test_tc_fixtures(gpu, small_graphs, golden):
    for name, (g, gs) in small_graphs.items():
        print(f"case tc {name}", flush=True)
        assert gpu.triangle_count(gs) == golden["graphs"][name]["tc_msa"], name


def test_tc_er_config(gpu, G, golden):
    rec = golden["generated"]["er_1000_9990_s1"]
    er = G.erdos_renyi(**rec["cfg"])
    assert er.sha256() == rec["csr_sha256"]
    assert gpu.triangle_count(er) == rec["tc"]


def test_tc_er_resident_repeats(gpu, G, golden, oracle):
    """Repeated counts of resident graphs: from the second call on, a graph
    runs its kept plan (statically split short rows) and each call's publish
    kernel leaves the sum and the claim counters zero for the next; two
    graphs' counts alternate, each with its own buffers."""
    rec = golden["generated"]["er_1000_9990_s1"]
    er = G.erdos_renyi(**rec["cfg"])
    other = G.erdos_renyi(n=700, m=9000, seed=5)
    want_other = oracle.tc(other)
    R, Ro = gpu.Graph(er), gpu.Graph(other)
    try:
        for it in range(6):
            print(f"case tc er resident call {it}", flush=True)
            assert R.triangle_count() == rec["tc"]
            assert Ro.triangle_count() == want_other
    finally:
        R.close()
        Ro.close()


@pytest.mark.parametrize("key", ["rmat16_ef16_s1", "rmat18_ef16_s1"])
def test_tc_rmat_full_size(gpu, G, golden, key):
    rec = golden["generated"][key]
    g = G.rmat(**rec["cfg"])
    assert g.sha256() == rec["csr_sha256"]
    print(f"case tc {key}", flush=True)
    assert gpu.triangle_count(g) == rec["tc"]
    R = gpu.Graph(g)
    assert R.triangle_count() == rec["tc"]
    assert R.triangle_count() == rec["tc"]  # resident: repeatable
    R.close()


def test_tc_closed_forms(gpu, G):
    for n in (3, 4, 7, 40, 300):
        assert gpu.triangle_count(_complete(G, n)) == n * (n - 1) * (n - 2) // 6
    assert gpu.triangle_count(G.CSR.from_edges(5, [], [])) == 0
    assert gpu.triangle_count(G.CSR.from_edges(1, [], [])) == 0


def test_tc_other_formulations(gpu, G, oracle):
    g = G.rmat(12, 8, seed=7)
    t = oracle.tc(g)
    assert gpu.triangle_counting_masked_lu(g) == t
    assert np.array_equal(gpu.triangle_counting_vertex(g), oracle.tc_vertex(g))


# ------------------------------------------------------------- k-truss
def test_ktruss_g43_per_k(gpu, small_graphs, golden):
    gs = small_graphs["G43.mtx"][1]
    for k, nnz, iters in golden["graphs"]["G43.mtx"]["truss"]["per_k"]:
        T, it = gpu.k_truss(gs, k)
        assert T.nnz == nnz and it == iters, k


def test_ktruss_pattern_matches_oracle(gpu, G, oracle):
    g = G.rmat(12, 16, seed=21)
    for k in (3, 5, 9):
        T, it = gpu.k_truss(g, k)
        rp, col, it_o = oracle.ktruss(g, k)
        assert np.array_equal(T.rowptr, rp) and np.array_equal(T.col, col) and it == it_o, k


@pytest.mark.parametrize("name", ["G43.mtx", "iris_dataset_30NN.mtx", "yeast_30NN.mtx.gz", "cage3.mtx"])
def test_truss_fixtures(gpu, small_graphs, golden, name):
    gs = small_graphs[name][1]
    rec = golden["graphs"][name]["truss"]
    truss, kmax, _ = gpu.truss_decomposition(gs)
    assert np.array_equal(truss, np.load(GOLDEN_DIR / rec["truss_npy"]))
    assert kmax == rec["kmax"]


def test_truss_rmat16_full_size(gpu, G, golden):
    """The bench workload: bit-exact truss numbers per edge against the reference's
    per-k k_truss loop on R-MAT scale 16 (recorded as SHA-256 + histogram)."""
    rec = golden["generated"]["rmat16_ef16_s1"]
    g = G.rmat(**rec["cfg"])
    assert g.sha256() == rec["csr_sha256"]
    print("case truss rmat16", flush=True)
    R = gpu.Graph(g)
    truss, kmax, stats = R.truss()
    R.close()
    assert kmax == rec["truss"]["kmax"]
    assert np.bincount(truss).tolist() == rec["truss"]["truss_hist"]
    assert hashlib.sha256(truss.astype(np.int32).tobytes()).hexdigest() == rec["truss"]["truss_sha256"]


def test_truss_edge_cases(gpu, G):
    truss, kmax, _ = gpu.truss_decomposition(G.CSR.from_edges(4, [], []))
    assert truss.size == 0 and kmax == 2
    truss, kmax, _ = gpu.truss_decomposition(G.CSR.from_edges(2, [0], [1]))
    assert (truss == 2).all() and kmax == 2
    truss, kmax, _ = gpu.truss_decomposition(_complete(G, 9))
    assert (truss == 9).all() and kmax == 9


def test_truss_numbers_in_page_locked_buffers(gpu, G, oracle):
    """Resident-graph truss numbers come back in page-locked buffers
    (ga_host_alloc): a buffer is reused only once the array handed out from it
    is gone, and arrays that outlive the graph stay valid."""
    g = G.rmat(10, 8, seed=3)
    want = oracle.truss_decomposition(g)[0]
    R = gpu.Graph(g)
    t1, _, _ = R.truss()
    t2, _, _ = R.truss()  # t1 still alive: a second buffer
    assert t1.ctypes.data != t2.ctypes.data
    addr = t2.ctypes.data
    del t2
    t3, _, _ = R.truss()  # t2's buffer again
    assert t3.ctypes.data == addr
    R.close()
    assert np.array_equal(t1, want) and np.array_equal(t3, want)


def _wide_graph(G, n_wide=(1 << 19) + 1000, seed=11):
    """R-MAT 13 plus a hub of 5000 neighbours, its vertices spread over
    n_wide > 2^18 ids: mask rows wider than the exact LDS bitmap (the filter
    path, with columns that alias in it) and one mask row longer than the
    LDS-staged capacity (the global search)."""
    rng = np.random.default_rng(seed)
    g = G.rmat(13, 16, seed=seed)
    rows = np.repeat(np.arange(g.n), np.diff(g.rowptr))
    ids = np.sort(rng.choice(n_wide - 1, size=g.n, replace=False))  # an order-keeping spread
    hub_nb = ids[rng.choice(g.n, size=5000, replace=False)]  # the hub is the last vertex: a long lower row
    u = np.concatenate([ids[rows], np.full(5000, n_wide - 1)])
    v = np.concatenate([ids[g.col], hub_nb])
    return G.CSR.from_edges(n_wide, u, v)


@pytest.mark.parametrize("n_wide", [1 << 18, (1 << 19) - 1000, (1 << 19) + 1000])
def test_wide_matrix_filter_paths(gpu, G, oracle, n_wide):
    """2^18 columns: the exact bitmap (sums count its bits, products with
    positions take the rank lookup); up to 2^19: the supports' 1024-entry
    items with rank bases; wider: sums on 1024-entry items over an exact
    bitmap, products with positions through the filter and a search."""
    g = _wide_graph(G, n_wide)
    print(f"case wide n={g.n} nnz={g.nnz} maxdeg={g.degrees().max()}", flush=True)
    got = gpu.mxm_masked(g, g, g, dtype=np.int32)
    want = oracle.mxm_masked(g.n, g.n, g.rowptr, g.col, g.rowptr, g.col, g.rowptr, g.col)
    assert np.array_equal(got.astype(np.int64), want)
    R = gpu.Graph(g)
    assert R.triangle_count() == oracle.tc(g)
    truss, kmax, _ = R.truss()
    want_t, want_kmax, *_ = oracle.truss_decomposition(g)
    R.close()
    assert kmax == want_kmax and np.array_equal(truss, want_t)
