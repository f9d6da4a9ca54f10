"""GPU tests: every tuning option gives the same results.  BC with top-down
forward steps only, with bottom-up steps from the first level on and with the
default switch point (bit-identical scores, and the oracle); the host-steered,
device-sized and graph-replayed BC level loops (bit-identical); triangle counts
with every short-row kernel and claim size (exact); truss numbers with tiny
work items (rows split over many items, whose partial supports meet in global
atomics) and with the selective support passes (exact); the host-driven,
device-driven and graph-replayed k-truss loops (identical counts)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9


def _bc_close(got, want):
    scale = max(1.0, float(np.abs(want).max()))
    return np.allclose(got, want, rtol=BC_RTOL, atol=BC_RTOL * scale * 1e-3)


@pytest.mark.parametrize("scale,ns", [(12, 64), (14, 256)])
def test_bc_directions_bit_identical(gpu, G, oracle, scale, ns):
    g = G.rmat(scale, 16, seed=40 + scale)
    src = G.pick_sources(g, ns, 3)
    R = gpu.Graph(g)
    print(f"case bc directions rmat{scale} ns={ns}", flush=True)
    with gpu.options(bc_bottom_up=0):
        top_down = R.bc(src)
    with gpu.options(bc_bottom_up=1, bc_alpha=1e6):  # bottom-up from level 1 on
        early = R.bc(src)
    with gpu.options(bc_bottom_up=1, bc_alpha=1.0):
        late = R.bc(src)
    default = R.bc(src)
    R.close()
    assert np.array_equal(top_down, early)
    assert np.array_equal(top_down, late)
    assert np.array_equal(top_down, default)
    assert _bc_close(default, oracle.bc(g, src))


def test_bc_directions_deep_and_disconnected(gpu, G, oracle):
    """A long path (levels of one vertex) plus a disjoint star: the bottom-up
    candidates include vertices no source reaches."""
    n_path, n_star = 300, 50
    rows = list(range(n_path - 1)) + [n_path] * (n_star - 1)
    cols = list(range(1, n_path)) + list(range(n_path + 1, n_path + n_star))
    g = G.CSR.from_edges(n_path + n_star, rows, cols)
    src = np.array([0, 7, 150, n_path + 3], dtype=np.int32)
    want = oracle.bc(g, src)
    for opts in ({"bc_bottom_up": 0}, {"bc_bottom_up": 1, "bc_alpha": 1e6}):
        with gpu.options(**opts):
            got = gpu.betweenness_centrality_batch(g, src)
        assert _bc_close(got, want), opts


@pytest.mark.parametrize("scale,ns", [(12, 64), (14, 256), (16, 200)])
def test_bc_loop_modes_bit_identical(gpu, G, oracle, scale, ns):
    """Host-steered levels, levels sized on the device, and those steps
    replayed from captured HIP graphs: the same scores bit for bit, in both
    forward directions, on repeated calls of one resident graph (the captured
    steps are reused) and against the oracle."""
    g = G.rmat(scale, 16, seed=60 + scale)
    src = G.pick_sources(g, ns, 5)
    src2 = G.pick_sources(g, ns, 6)
    R = gpu.Graph(g)
    got = {}
    for loop in (0, 1, 2):
        for bu in ({"bc_bottom_up": 0}, {"bc_bottom_up": 1}, {"bc_bottom_up": 1, "bc_alpha": 1e6}):
            print(f"case bc loop={loop} {bu} rmat{scale} ns={ns}", flush=True)
            with gpu.options(bc_loop=loop, **bu):
                got[loop, tuple(bu.items())] = (R.bc(src), R.bc(src2), R.bc(src))
    R.close()
    ref = got[0, (("bc_bottom_up", 0),)]
    for key, (a, b, c) in got.items():
        assert np.array_equal(a, ref[0]) and np.array_equal(c, ref[0]), key
        assert np.array_equal(b, ref[1]), key
    if scale <= 14:
        assert _bc_close(ref[0], oracle.bc(g, src))


def test_bc_loop_modes_deep_and_disconnected(gpu, G, oracle):
    """Levels beyond the 8-bit depth range (the pass is redone with 16-bit
    depths), a disjoint star no source reaches, isolated vertices, and a
    graph whose sources have no edges."""
    n_path, n_star = 600, 50
    rows = list(range(n_path - 1)) + [n_path] * (n_star - 1)
    cols = list(range(1, n_path)) + list(range(n_path + 1, n_path + n_star))
    g = G.CSR.from_edges(n_path + n_star + 5, rows, cols)
    src = np.array([0, 7, 300, n_path + 3], dtype=np.int32)
    lonely = np.array([n_path + n_star, n_path + n_star + 4], dtype=np.int32)
    want, want_lonely = oracle.bc(g, src), oracle.bc(g, lonely)
    for loop in (0, 1, 2):
        for opts in ({"bc_bottom_up": 0}, {"bc_bottom_up": 1, "bc_alpha": 1e6}):
            print(f"case bc deep loop={loop} {opts}", flush=True)
            with gpu.options(bc_loop=loop, **opts):
                got = gpu.betweenness_centrality_batch(g, src)
                got_lonely = gpu.betweenness_centrality_batch(g, lonely)
            assert _bc_close(got, want), (loop, opts)
            assert np.array_equal(got_lonely, want_lonely) and not got_lonely.any()


@pytest.fixture(scope="module")
def tc_cases(G, oracle):
    graphs = G.knn_batch(base_seed=3000, count=12, k=30)
    er = G.erdos_renyi(1000, 9990, seed=1)
    return graphs, np.array([oracle.tc(g) for g in graphs], dtype=np.int64), er, oracle.tc(er)


@pytest.mark.parametrize("sub", [0, 1, 2, 8, 16, 32, 64])
@pytest.mark.parametrize("claim", [0, 1, 3, 16])
def test_tc_short_row_kernels_and_claims(gpu, tc_cases, sub, claim):
    graphs, want, er, want_er = tc_cases
    print(f"case tc sum_sub={sub} row_claim={claim}", flush=True)
    with gpu.options(sum_sub=sub, row_claim=claim):
        B = gpu.Batch(graphs)
        got = B.triangle_counts()
        B.close()
        R = gpu.Graph(er)
        got_er = R.triangle_count()
        R.close()
    assert np.array_equal(got, want)
    assert got_er == want_er


@pytest.mark.parametrize("wave_flops", [-1, 0, 2048, 16384, 1 << 20])
def test_tc_short_row_threshold(gpu, tc_cases, wave_flops):
    """Every candidate of the batch autotuner's short-row product threshold
    (and the extremes: no short-row kernel, every row that fits it) gives the
    exact counts, batched and per graph."""
    graphs, want, er, want_er = tc_cases
    print(f"case tc wave_flops={wave_flops}", flush=True)
    with gpu.options(wave_flops=wave_flops):
        B = gpu.Batch(graphs)
        got = B.triangle_counts()
        B.close()
        R = gpu.Graph(er)
        got_er = R.triangle_count()
        R.close()
    assert np.array_equal(got, want)
    assert got_er == want_er


@pytest.mark.parametrize("item_wedges", [64, 4096])
def test_truss_small_items(gpu, G, oracle, item_wedges):
    g = G.rmat(12, 16, seed=8)
    want, kmax, _, _, _ = oracle.truss_decomposition(g)
    print(f"case truss item_wedges={item_wedges}", flush=True)
    with gpu.options(item_wedges=item_wedges, truss_loop=0):
        got, got_kmax, _ = gpu.truss_decomposition(g)
    assert got_kmax == kmax
    assert np.array_equal(got, want)


def test_truss_selective_support(gpu, G, oracle):
    g = G.rmat(12, 16, seed=9)
    want, kmax, _, _, _ = oracle.truss_decomposition(g)
    with gpu.options(selective_support=1, truss_loop=0):
        got, got_kmax, _ = gpu.truss_decomposition(g)
        R = gpu.Graph(g)
        nnz5 = R.ktruss(5)[0]
        R.close()
    assert got_kmax == kmax and np.array_equal(got, want)
    assert nnz5 == int((want >= 5).sum())


def test_large_graph_pass_kinds(gpu, G):
    """A graph past the sizes where the defaults change pass kind (2^20 / 2^22
    stored entries: triangle-listing passes): the same k-truss sizes and
    pass counts as with symmetric half passes and as with selective passes,
    and the sizes the decomposition's truss numbers give."""
    g = G.rmat(18, 16, seed=31)
    assert g.nnz >= 1 << 22
    ks = (3, 6, 12)
    print(f"case large graph pass kinds rmat18 nnz={g.nnz}", flush=True)
    R = gpu.Graph(g)
    with gpu.options(selective_support=0, support_tri=0, truss_loop=0):
        truss, kmax, _ = R.truss()
        half = [R.ktruss(k) for k in ks]
    with gpu.options(selective_support=1, support_tri=0, truss_loop=0):
        selective = [R.ktruss(k) for k in ks]
    with gpu.options(truss_loop=0):
        tri = [R.ktruss(k) for k in ks]
    default = [R.ktruss(k) for k in ks]
    peel_truss, peel_kmax, _ = R.truss()
    R.close()
    assert default == half == selective == tri
    assert peel_kmax == kmax and np.array_equal(peel_truss, truss)
    for k, (nnz, _) in zip(ks, default):
        assert nnz == int((truss >= k).sum())


def _truss_cases(G):
    path = G.CSR.from_edges(40, list(range(39)), list(range(1, 40)))  # a tree: every k-truss is empty
    empty = G.CSR.from_edges(16, [], [])
    return {"rmat12": G.rmat(12, 16, seed=21), "rmat14": G.rmat(14, 8, seed=22),
            "er1000": G.erdos_renyi(1000, 9990, seed=1), "path40": path, "empty16": empty}


@pytest.mark.parametrize("name", ["rmat12", "rmat14", "er1000", "path40", "empty16"])
def test_truss_loops_identical(gpu, G, oracle, name):
    """Host-driven peeling (0), device-driven steps (1), steps replayed
    from a captured HIP graph (2), peeling by support updates with its rounds
    queued on the device (3) or steered by the host (4) and the default (-1):
    the oracle's truss numbers and support
    pass count, and between the loops the same filter steps, k-truss sizes,
    pass counts and k-truss patterns (resident graph queried twice, so the
    captured step is reused)."""
    g = _truss_cases(G)[name]
    want, kmax, opasses, _, _ = oracle.truss_decomposition(g)
    ks = sorted({3, 4, 5, max(kmax, 3), kmax + 1})
    res = {}
    for loop in (0, 1, 2, 3, 4, -1):
        print(f"case truss loops {name} truss_loop={loop}", flush=True)
        with gpu.options(truss_loop=loop):
            t1, k1, s1 = gpu.truss_decomposition(g)
            R = gpu.Graph(g)
            t2, k2, s2 = R.truss()
            t3, k3, s3 = R.truss()
            kt = [R.ktruss(k) for k in ks] + [R.ktruss(k) for k in ks]
            R.close()
            pat, it = gpu.k_truss(g, 4)
        for t, k, st in ((t1, k1, s1), (t2, k2, s2), (t3, k3, s3)):
            assert k == kmax and np.array_equal(t, want)
            assert st["passes"] == opasses
        assert s1["steps"] == s2["steps"] == s3["steps"]
        for k, (nnz, _) in zip(ks + ks, kt):
            assert nnz == int((want >= k).sum()), (k, nnz)
        res[loop] = (s1["steps"], kt, it, pat.rowptr.tolist(), pat.col.tolist())
    assert res[0] == res[1] == res[2] == res[3] == res[4] == res[-1]


@pytest.mark.parametrize("name", ["rmat12", "rmat14", "er1000", "path40", "empty16"])
def test_truss_support_tri(gpu, G, oracle, name):
    """Support passes as triangle listing (each triangle found once from
    C<L> = L*L, its three edges incremented): the oracle's truss numbers and
    pass count with the host-driven and the device-driven loop, and the
    oracle's k-truss pattern and iterations."""
    g = _truss_cases(G)[name]
    want, kmax, opasses, _, _ = oracle.truss_decomposition(g)
    for loop in (0, 1):
        print(f"case truss support_tri {name} truss_loop={loop}", flush=True)
        with gpu.options(support_tri=1, truss_loop=loop, selective_support=0):
            got, got_kmax, st = gpu.truss_decomposition(g)
            T, it = gpu.k_truss(g, 4)
        assert got_kmax == kmax and np.array_equal(got, want) and st["passes"] == opasses
        rp, col, it_o = oracle.ktruss(g, 4)
        assert np.array_equal(T.rowptr, rp) and np.array_equal(T.col, col) and it == it_o


def test_truss_support_tri_rmat16(gpu, G, golden):
    """Triangle-listing support passes on the bench workload: the reference's
    truss numbers (SHA-256 and histogram)."""
    import hashlib
    rec = golden["generated"]["rmat16_ef16_s1"]
    g = G.rmat(**rec["cfg"])
    print("case truss support_tri rmat16", flush=True)
    R = gpu.Graph(g)
    with gpu.options(support_tri=1):
        truss, kmax, _ = R.truss()
    R.close()
    assert kmax == rec["truss"]["kmax"]
    assert np.bincount(truss).tolist() == rec["truss"]["truss_hist"]
    assert hashlib.sha256(truss.astype(np.int32).tobytes()).hexdigest() == rec["truss"]["truss_sha256"]


def test_option_validation(gpu):
    for name, bad in (("bc_bottom_up", 2), ("bc_alpha", 0), ("row_claim", -1), ("sum_sub", 12),
                      ("item_wedges", 0), ("selective_support", 0.5), ("truss_loop", 5), ("truss_loop", -2),
                      ("support_tri", 2), ("serve_join", 3), ("bc_loop", 3), ("bc_loop", -1),
                      ("wave_flops", -2), ("wave_flops", 1.5)):
        with pytest.raises(Exception):
            gpu.set_option(name, bad)
    before = gpu.get_option("row_claim")
    with gpu.options(row_claim=5):
        assert gpu.get_option("row_claim") == 5
    assert gpu.get_option("row_claim") == before
