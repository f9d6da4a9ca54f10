"""Pins of the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each function of the oracle is checked against (a) the reference's own outputs
recorded in tests/golden/golden.json by tests/golden/make_golden.py (the
harness compiled from the reference headers), (b) closed forms, and (c)
independent library routines (scipy, networkx), chosen so that a dropped term,
a wrong index or a transposed operand fails at least one of them.
"""
import itertools

import numpy as np
import pytest

from conftest import GOLDEN_DIR


# ---------------------------------------------------------------- helpers
def complete(G, n):
    u, v = zip(*itertools.combinations(range(n), 2))
    return G.CSR.from_edges(n, u, v)


def path(G, n):
    return G.CSR.from_edges(n, range(n - 1), range(1, n))


def cycle(G, n):
    return G.CSR.from_edges(n, range(n), [(i + 1) % n for i in range(n)])


def star(G, n):
    return G.CSR.from_edges(n, [0] * (n - 1), range(1, n))


def wheel(G, n):  # hub 0 + cycle over 1..n-1
    u = [0] * (n - 1) + list(range(1, n))
    v = list(range(1, n)) + [1 + (i % (n - 1)) for i in range(1, n)]
    return G.CSR.from_edges(n, u, v)


def random_tree(G, n, seed):
    rng = np.random.default_rng(seed)
    parent = [int(rng.integers(0, i)) for i in range(1, n)]
    return G.CSR.from_edges(n, range(1, n), parent)


def dense(g):
    D = np.zeros((g.n, g.n), dtype=np.int64)
    for i in range(g.n):
        D[i, g.col[g.rowptr[i]:g.rowptr[i + 1]]] = 1
    return D


# ---------------------------------------------------------------- masked SpGEMM
def test_mxm_masked_complete_graph_closed_form(G, oracle):
    """K_n: C<L> = L*L has C(i,j) = i-j-1 (the number of k with j<k<i)."""
    n = 12
    L = complete(G, n).lower()
    C = oracle.mxm_masked(n, n, L.rowptr, L.col, L.rowptr, L.col, L.rowptr, L.col)
    rows = np.repeat(np.arange(n), np.diff(L.rowptr))
    assert np.array_equal(C, rows - L.col - 1)


@pytest.mark.parametrize("seed", range(6))
def test_mxm_masked_brute_force_small(G, oracle, seed):
    """Exhaustive dense product on tiny random operands with values, rectangular shapes."""
    rng = np.random.default_rng(seed)
    m, k, n = rng.integers(1, 9, size=3)
    A = (rng.random((m, k)) < 0.5) * rng.integers(-3, 4, size=(m, k))
    B = (rng.random((k, n)) < 0.5) * rng.integers(-3, 4, size=(k, n))
    Mk = rng.random((m, n)) < 0.5

    def csr(D):
        rp = np.concatenate([[0], np.cumsum((D != 0).sum(1))]).astype(np.int32)
        r, c = np.nonzero(D)
        return rp, c.astype(np.int32), D[r, c]
    Arp, Acol, Aval = csr(A)
    Brp, Bcol, Bval = csr(B)
    Mrp, Mcol, _ = csr(Mk.astype(np.int64))
    got = oracle.mxm_masked(m, n, Arp, Acol, Brp, Bcol, Mrp, Mcol, Aval=Aval, Bval=Bval)
    r, c = np.nonzero(Mk)
    assert np.array_equal(got, (A @ B)[r, c])
    gotf = oracle.mxm_masked(m, n, Arp, Acol, Brp, Bcol, Mrp, Mcol, Aval=Aval * 0.5, Bval=Bval, dtype=np.float64)
    assert np.allclose(gotf, 0.5 * (A @ B)[r, c], rtol=0, atol=1e-12)


def test_mxm_masked_scipy_rmat(G, oracle):
    """C<A> = A*A on an R-MAT graph against scipy, explicit zeros included."""
    g = G.rmat(11, 8, seed=3)
    C = oracle.mxm_masked(g.n, g.n, g.rowptr, g.col, g.rowptr, g.col, g.rowptr, g.col)
    S = g.to_scipy()
    P = (S @ S).tocsr()
    rows = np.repeat(np.arange(g.n), np.diff(g.rowptr))
    exp = np.asarray(P[rows, g.col]).ravel()
    assert np.array_equal(C, exp)
    assert (C == 0).any()  # explicit zeros are part of the result


def test_mxm_matches_reference_mxm(G, oracle, small_graphs, tmp_path):
    """C<A> = A*A on G43 against the reference's own mxmm_msa (oracle/_ref), when built here."""
    if not oracle.ref_available():
        pytest.skip("reference harness not built on this host")
    g = small_graphs["G43.mtx"][1]
    out = tmp_path / "c.bin"
    oracle.ref_run("mxm", str(GOLDEN_DIR / "graphs" / "G43.mtx"), str(out))
    ref = np.fromfile(out, dtype=np.int32)
    C = oracle.mxm_masked(g.n, g.n, g.rowptr, g.col, g.rowptr, g.col, g.rowptr, g.col)
    assert np.array_equal(C, ref)


def test_wedges(G, oracle):
    g = G.rmat(10, 8, seed=5)
    assert oracle.wedges(g) == int((g.degrees() ** 2).sum())


# ---------------------------------------------------------------- triangle counting
def test_tc_reference_fixtures(small_graphs, oracle, golden):
    for name, (g, gs) in small_graphs.items():
        rec = golden["graphs"][name]
        assert rec["tc_msa"] == rec["tc_mca"]
        assert oracle.tc(gs) == rec["tc_msa"], name


def test_tc_reference_pins(small_graphs, oracle):
    # triangle counts the compiled reference printed for these four graphs
    assert oracle.tc(small_graphs["G43.mtx"][1]) == 1392
    assert oracle.tc(small_graphs["yeast_30NN.mtx.gz"][1]) == 180175
    assert oracle.tc(small_graphs["iris_dataset_30NN.mtx"][1]) == 25972
    assert oracle.tc(small_graphs["cage3.mtx"][1]) == 3


def test_tc_closed_forms(G, oracle):
    for n in (3, 4, 7, 13):
        assert oracle.tc(complete(G, n)) == n * (n - 1) * (n - 2) // 6
    for n in (4, 5, 9):
        assert oracle.tc(cycle(G, n)) == 0
    for n in (5, 8, 11):
        assert oracle.tc(wheel(G, n)) == n - 1
    assert oracle.tc(random_tree(G, 40, 1)) == 0
    assert oracle.tc(G.CSR.from_edges(5, [], [])) == 0


def test_tc_formulations_agree(G, oracle):
    g = G.rmat(12, 8, seed=7)
    t = oracle.tc(g)
    assert oracle.tc_lu(g) == t
    pv = oracle.tc_vertex(g)
    assert pv.sum() == 3 * t
    S = g.to_scipy()
    diag3 = np.asarray((S @ S).multiply(S).sum(axis=1)).ravel()  # diag(A^3)
    assert np.array_equal(pv, diag3 // 2)


def test_tc_generated_pins(G, oracle, golden):
    rec = golden["generated"]["er_1000_9990_s1"]
    er = G.erdos_renyi(**rec["cfg"])
    assert er.sha256() == rec["csr_sha256"]
    assert oracle.tc(er) == rec["tc"]
    assert er.lower().wedges() == rec["wedges_LL"]


# ---------------------------------------------------------------- k-truss
def test_ktruss_g43_sizes(small_graphs, oracle, golden):
    """Per-k sizes and iteration counts of the reference's k_truss (from scratch per k)."""
    gs = small_graphs["G43.mtx"][1]
    for k, nnz, iters in golden["graphs"]["G43.mtx"]["truss"]["per_k"]:
        rp, col, it = oracle.ktruss(gs, k)
        assert rp[-1] == nnz and it == iters, k
    assert oracle.ktruss(gs, 3)[0][-1] == 6818  # the reference's k_truss output on G43
    assert oracle.ktruss(gs, 4)[0][-1] == 36


@pytest.mark.parametrize("name", ["G43.mtx", "iris_dataset_30NN.mtx", "yeast_30NN.mtx.gz", "cage3.mtx"])
def test_truss_decomposition_reference(small_graphs, oracle, golden, name):
    """Incremental truss decomposition == the reference's per-k loop, entry by entry."""
    gs = small_graphs[name][1]
    rec = golden["graphs"][name]["truss"]
    truss, kmax, passes, _, _ = oracle.truss_decomposition(gs)
    ref = np.load(GOLDEN_DIR / rec["truss_npy"])
    assert np.array_equal(truss, ref)
    assert kmax == rec["kmax"]


def test_truss_closed_forms(G, oracle):
    for n in (3, 5, 8):
        truss, kmax, *_ = oracle.truss_decomposition(complete(G, n))
        assert (truss == n).all() and kmax == n
    truss, kmax, *_ = oracle.truss_decomposition(random_tree(G, 30, 2))
    assert (truss == 2).all() and kmax == 2
    # two K4 sharing a vertex plus a pendant edge
    u = [0, 0, 0, 1, 1, 2, 3, 3, 3, 4, 4, 5, 6]
    v = [1, 2, 3, 2, 3, 3, 4, 5, 6, 5, 6, 6, 7]
    g = G.CSR.from_edges(8, u, v)
    truss, kmax, *_ = oracle.truss_decomposition(g)
    rows = np.repeat(np.arange(8), np.diff(g.rowptr))
    pend = ((rows == 6) & (g.col == 7)) | ((rows == 7) & (g.col == 6))
    assert (truss[pend] == 2).all() and (truss[~pend] == 4).all()


def test_truss_networkx(G, oracle):
    nx = pytest.importorskip("networkx")
    g = G.rmat(10, 8, seed=11)
    truss, kmax, *_ = oracle.truss_decomposition(g)
    X = nx.Graph()
    rows = np.repeat(np.arange(g.n), np.diff(g.rowptr))
    X.add_edges_from(zip(rows.tolist(), g.col.tolist()))
    for k in range(3, kmax + 2):
        T = nx.k_truss(X, k)
        assert 2 * T.number_of_edges() == int((truss >= k).sum()), k


def test_truss_invariant_support_sum(G, oracle):
    """sum over directed entries of the 3-truss support = 6 T (every triangle counted 6x)."""
    g = G.rmat(10, 8, seed=13)
    S = oracle.mxm_masked(g.n, g.n, g.rowptr, g.col, g.rowptr, g.col, g.rowptr, g.col)
    assert S.sum() == 6 * oracle.tc(g)


# ---------------------------------------------------------------- cmask / spd
def test_cmask_scipy(G, oracle):
    g = G.rmat(9, 8, seed=17)
    rng = np.random.default_rng(0)
    ns = 7
    # B: random sparse n x ns with FP64 values, M: random sparse n x ns
    B = (rng.random((g.n, ns)) < 0.1) * rng.random((g.n, ns))
    M = rng.random((g.n, ns)) < 0.3

    def csr(D):
        rp = np.concatenate([[0], np.cumsum((D != 0).sum(1))]).astype(np.int32)
        r, c = np.nonzero(D)
        return rp, c.astype(np.int32), D[r, c]
    Brp, Bcol, Bval = csr(B)
    Mrp, Mcol, _ = csr(M.astype(float))
    rp, col, val = oracle.mxm_cmask(g.n, ns, g.rowptr, g.col, Brp, Bcol, Mrp, Mcol, Bval=Bval)
    P = dense(g).astype(float) @ B
    Pat = (dense(g) @ (B != 0)) > 0
    want = Pat & ~M
    r, c = np.nonzero(want)
    assert np.array_equal(rp, np.concatenate([[0], np.cumsum(want.sum(1))]))
    assert np.array_equal(col, c)
    assert np.allclose(val, P[r, c], rtol=1e-14, atol=0)


def test_spd_dense(G, oracle):
    g = G.rmat(8, 8, seed=19)
    rng = np.random.default_rng(1)
    ncols = 5
    B = rng.random((g.n, ncols))
    M = rng.random((g.n, ncols)) < 0.4
    Mrp = np.concatenate([[0], np.cumsum(M.sum(1))]).astype(np.int32)
    Mcol = np.nonzero(M)[1].astype(np.int32)
    got = oracle.mxm_spd(g.n, ncols, g.rowptr, g.col, B, Mrp, Mcol)
    assert np.allclose(got, (dense(g) @ B) * M, rtol=1e-14, atol=0)


# ---------------------------------------------------------------- betweenness centrality
def test_bc_closed_forms(G, oracle):
    n = 9
    s = star(G, n)
    bc = oracle.bc(s, list(range(1, n)))
    assert bc[0] == pytest.approx((n - 1) * (n - 2)) and np.allclose(bc[1:], 0)
    p = path(G, n)
    bc = oracle.bc(p, list(range(n)))
    v = np.arange(n)
    assert np.allclose(bc, 2 * v * (n - 1 - v))
    assert np.allclose(oracle.bc(complete(G, 7), list(range(7))), 0)


def test_bc_networkx(G, oracle, small_graphs):
    nx = pytest.importorskip("networkx")
    g = small_graphs["G43.mtx"][0]
    X = nx.Graph()
    X.add_nodes_from(range(g.n))
    rows = np.repeat(np.arange(g.n), np.diff(g.rowptr))
    X.add_edges_from(zip(rows.tolist(), g.col.tolist()))
    rng = np.random.default_rng(5)
    src = rng.choice(g.n, 24, replace=False)
    bc = oracle.bc(g, src)
    ref = nx.betweenness_centrality_subset(X, sources=src.tolist(), targets=list(range(g.n)), normalized=False)
    ref = 2 * np.array([ref[v] for v in range(g.n)])
    assert np.allclose(bc, ref, rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("name", ["G43.mtx", "iris_dataset_30NN.mtx", "yeast_30NN.mtx.gz"])
def test_bc_reference_float32(small_graphs, oracle, golden, name):
    """FP64 restatement vs the reference's float32 betweenness_centrality_batch
    (sources 0..15): equal up to float32 rounding and its cancellation (defect D5)."""
    g = small_graphs[name][0]
    rec = golden["graphs"][name]["bc_ref_f32"]
    ref = np.load(GOLDEN_DIR / rec["npy"]).astype(np.float64)
    bc = oracle.bc(g, np.arange(rec["batch"]))
    assert np.allclose(bc, ref, rtol=1e-3, atol=2e-3 * max(1.0, np.abs(ref).max()) * 1e-2)


def test_bc_brute_force_tiny(G, oracle):
    """Exhaustive shortest-path enumeration (BFS path counts) on a tiny random graph."""
    g = G.erdos_renyi(14, 24, seed=3)
    D = dense(g)
    n = g.n
    INF = 10 ** 9
    dist = np.full((n, n), INF)
    sig = np.zeros((n, n))
    for s in range(n):
        dist[s, s] = 0
        sig[s, s] = 1
        frontier = [s]
        while frontier:
            nxt = []
            for u in frontier:
                for w in np.nonzero(D[u])[0]:
                    if dist[s, w] == INF:
                        dist[s, w] = dist[s, u] + 1
                        nxt.append(w)
                    if dist[s, w] == dist[s, u] + 1:
                        sig[s, w] += sig[s, u]
            frontier = nxt
    src = [0, 3, 5, 9]
    want = np.zeros(n)
    for s in src:
        for t in range(n):
            if t == s or dist[s, t] >= INF:
                continue
            for v in range(n):
                if v in (s, t) or dist[s, v] >= INF or dist[v, t] >= INF:
                    continue
                if dist[s, v] + dist[v, t] == dist[s, t]:
                    want[v] += sig[s, v] * sig[v, t] / sig[s, t]
    assert np.allclose(oracle.bc(g, src), want, rtol=1e-12, atol=1e-12)


def test_planted_mistake_is_caught(G, oracle):
    """A plausible oracle mistake -- masking C<L> = L*L with U instead of L (transposed
    mask) -- changes the K_n closed form, so the closed-form pin discriminates."""
    n = 8
    K = complete(G, n)
    L = K.lower()
    rows = np.repeat(np.arange(n), np.diff(L.rowptr))
    # transposed operand: L * L^T masked by L is not the sandia product
    Lt_rp = np.zeros(n + 1, dtype=np.int32)
    cnt = np.bincount(L.col, minlength=n)
    np.cumsum(cnt, out=Lt_rp[1:])
    order = np.lexsort((rows, L.col))
    Lt_col = rows[order].astype(np.int32)
    wrong = oracle.mxm_masked(n, n, L.rowptr, L.col, Lt_rp, Lt_col, L.rowptr, L.col)
    assert not np.array_equal(wrong, rows - L.col - 1)


# ---------------------------------------------------------------- element-wise operators
def _ew_operand(rng, m, p):
    """Random pattern with random values, explicit zeros included (a stored zero is part
    of the pattern)."""
    P = rng.random((m, m)) < p
    V = rng.integers(-3, 4, size=(m, m)) * P
    rp = np.concatenate([[0], np.cumsum(P.sum(1))]).astype(np.int32)
    r, c = np.nonzero(P)
    return P, V, rp, c.astype(np.int32), V[r, c]


@pytest.mark.parametrize("seed", range(8))
def test_ewise_dense_brute_force(oracle, seed):
    """eWiseAdd / eWiseMult / add_nointersect, with and without a mask, against dense
    arrays: the result pattern and values read off boolean and value matrices."""
    rng = np.random.default_rng(100 + seed)
    m = int(rng.integers(1, 12))
    PA, VA, Arp, Acol, Aval = _ew_operand(rng, m, 0.4)
    PB, VB, Brp, Bcol, Bval = _ew_operand(rng, m, 0.4)
    PM, _, Mrp, Mcol, _ = _ew_operand(rng, m, 0.5)

    def check(got, P, D):
        rp, col, val = got
        r, c = np.nonzero(P)
        assert np.array_equal(rp, np.concatenate([[0], np.cumsum(P.sum(1))]))
        assert np.array_equal(col, c) and np.array_equal(val, D[r, c])

    S = VA + VB
    for dt in (np.int64, np.float64):
        check(oracle.ewise(oracle.EW_ADD, m, Arp, Acol, Aval, Brp, Bcol, Bval, dtype=dt), PA | PB, S)
        check(oracle.ewise(oracle.EW_MULT, m, Arp, Acol, Aval, Brp, Bcol, Bval, dtype=dt), PA & PB, VA * VB)
        check(oracle.ewise(oracle.EW_ADD, m, Arp, Acol, Aval, Brp, Bcol, Bval, Mrp, Mcol, dtype=dt),
              PM & (S != 0), S)
        check(oracle.ewise(oracle.EW_MULT, m, Arp, Acol, Aval, Brp, Bcol, Bval, Mrp, Mcol, dtype=dt),
              PM & PA & PB, VA * VB)
    # pattern operands: values are ones
    check(oracle.ewise(oracle.EW_ADD, m, Arp, Acol, None, Brp, Bcol, None), PA | PB,
          PA.astype(np.int64) + PB)
    # add_nointersect: an error on a shared position, the union otherwise
    if (PA & PB).any():
        with pytest.raises(ValueError):
            oracle.ewise(oracle.EW_ADD_DISJOINT, m, Arp, Acol, Aval, Brp, Bcol, Bval)
    PC = PB & ~PA
    Crp = np.concatenate([[0], np.cumsum(PC.sum(1))]).astype(np.int32)
    r, c = np.nonzero(PC)
    check(oracle.ewise(oracle.EW_ADD_DISJOINT, m, Arp, Acol, Aval, Crp, c.astype(np.int32), VB[r, c]),
          PA | PC, VA + VB * PC)


def test_ewise_invariants(G, oracle):
    """On an R-MAT graph: A+B = B+A, A.*B = B.*A, the masked results are the unmasked
    ones restricted to M (zeros of the masked sum dropped), A.*A has A's pattern."""
    g = G.rmat(9, 8, seed=4)
    h = G.rmat(9, 8, seed=5)
    rng = np.random.default_rng(6)
    va = rng.integers(-2, 3, size=g.nnz)
    vb = rng.integers(-2, 3, size=h.nnz)
    n = g.n
    for op in (oracle.EW_ADD, oracle.EW_MULT):
        x = oracle.ewise(op, n, g.rowptr, g.col, va, h.rowptr, h.col, vb)
        y = oracle.ewise(op, n, h.rowptr, h.col, vb, g.rowptr, g.col, va)
        assert all(np.array_equal(a, b) for a, b in zip(x, y))
    m = G.rmat(9, 4, seed=7)
    for op in (oracle.EW_ADD, oracle.EW_MULT):
        rp, col, val = oracle.ewise(op, n, g.rowptr, g.col, va, h.rowptr, h.col, vb)
        mrp, mcol, mval = oracle.ewise(op, n, g.rowptr, g.col, va, h.rowptr, h.col, vb, m.rowptr, m.col)
        full = dict(zip(np.repeat(np.arange(n), np.diff(rp)) * n + col, val))
        mk = np.repeat(np.arange(n), np.diff(m.rowptr)) * n + m.col
        want = [(k, full[k]) for k in mk if k in full and (op == oracle.EW_MULT or full[k] != 0)]
        got = list(zip(np.repeat(np.arange(n), np.diff(mrp)) * n + mcol, mval))
        assert got == want
    rp, col, val = oracle.ewise(oracle.EW_MULT, n, g.rowptr, g.col, va, g.rowptr, g.col, va)
    assert np.array_equal(rp, g.rowptr) and np.array_equal(col, g.col) and np.array_equal(val, va * va)


def test_ewise_reference(G, oracle, small_graphs, tmp_path):
    """eWiseAdd(A, B) and add_nointersect(A, B) against the reference's own operators
    (oracle/_ref, values 1), when built here: G43 with a random pattern, and G43's
    lower triangle with its strictly upper part."""
    if not oracle.ref_available():
        pytest.skip("reference harness not built on this host")
    g = small_graphs["G43.mtx"][1]
    n = g.n
    h = G.erdos_renyi(n, 3 * n, seed=11)
    rows = np.repeat(np.arange(n), np.diff(g.rowptr))
    up = g.col > rows
    U = G.CSR(n, np.concatenate([[0], np.cumsum(np.bincount(rows[up], minlength=n))]).astype(np.int32),
              g.col[up].astype(np.int32))
    L = g.lower()
    for mode, a, b in (("ewadd", g, h), ("ewdisj", L, U)):
        pa, pb, out = tmp_path / "a.csr", tmp_path / "b.csr", tmp_path / "c.bin"
        a.write_csr(str(pa))
        b.write_csr(str(pb))
        info = oracle.ref_run(mode, str(pa), str(pb), str(out))
        raw = np.fromfile(out, dtype=np.int32)
        nnz = info["nnz"]
        rp, col, val = raw[:n + 1], raw[n + 1:n + 1 + nnz], raw[n + 1 + nnz:]
        op = oracle.EW_ADD if mode == "ewadd" else oracle.EW_ADD_DISJOINT
        want = oracle.ewise(op, n, a.rowptr, a.col, None, b.rowptr, b.col, None)
        assert np.array_equal(rp, want[0]) and np.array_equal(col, want[1]) and np.array_equal(val, want[2])
    assert (want[2] == 1).all() and np.array_equal(want[1], g.col)


def test_scale(oracle):
    """multScalar: the product in FP64, converted toward zero for integer values."""
    assert np.array_equal(oracle.scale([7, -7, 3, 0], 0.5), [3, -3, 1, 0])
    assert np.array_equal(oracle.scale([7, -7], -1.5, dtype=np.int32), [-10, 10])
    assert np.allclose(oracle.scale([0.25, -2.0], 3.0, dtype=np.float64), [0.75, -6.0], rtol=0, atol=0)


# ---------------------------------------------------------------- float32 and dense operators
def test_spd_sequential_restatement(G, oracle):
    """mxm_spd_seq (numpy, any value type) against the C oracle's FP64
    mxmm_spd: bit-identical in FP64 (both sum A's row in CSR order with
    separate roundings); in FP32 exact on integer-valued operands and within
    FP32 rounding of the FP64 result otherwise."""
    g = G.rmat(9, 8, seed=23)
    rng = np.random.default_rng(2)
    ncols = 6
    M = rng.random((g.n, ncols)) < 0.5
    Mrp = np.concatenate([[0], np.cumsum(M.sum(1))]).astype(np.int32)
    Mcol = np.nonzero(M)[1].astype(np.int32)
    Aval = rng.random(g.nnz)
    B = rng.random((g.n, ncols))
    want = oracle.mxm_spd(g.n, ncols, g.rowptr, g.col, B, Mrp, Mcol, Aval=Aval)
    got = oracle.mxm_spd_seq(g.rowptr, g.col, Aval, B, Mrp, Mcol, dtype=np.float64)
    assert np.array_equal(got, want)
    Bi = rng.integers(-8, 9, size=(g.n, ncols)).astype(np.float64)
    exact = oracle.mxm_spd(g.n, ncols, g.rowptr, g.col, Bi, Mrp, Mcol)
    assert np.array_equal(oracle.mxm_spd_seq(g.rowptr, g.col, None, Bi, Mrp, Mcol, dtype=np.float32), exact)
    f32 = oracle.mxm_spd_seq(g.rowptr, g.col, Aval, B, Mrp, Mcol, dtype=np.float32)
    assert f32.dtype == np.float32
    assert np.allclose(f32, want, rtol=2e-5, atol=1e-6)
    assert not np.array_equal(f32.astype(np.float64), want)  # the FP32 rounding is really there


def test_dense_elementwise_brute_force(oracle):
    """fuseEWiseMultAdd and the masked dense eWiseMult against scalar loops
    in the value type (each product and sum rounded separately)."""
    rng = np.random.default_rng(3)
    m, n = 7, 5
    M = rng.random((m, n)) < 0.4
    Mrp = np.concatenate([[0], np.cumsum(M.sum(1))]).astype(np.int32)
    Mcol = np.nonzero(M)[1].astype(np.int32)
    for dt in (np.float32, np.float64, np.int32, np.int64):
        if np.dtype(dt).kind == "f":
            A, B, Cm = (rng.random((m, n)).astype(dt) for _ in range(3))
        else:
            A, B, Cm = (rng.integers(-50, 50, size=(m, n)).astype(dt) for _ in range(3))
        want = np.empty_like(Cm)
        wmul = np.zeros_like(Cm)
        for i in range(m):
            for j in range(n):
                p = A[i, j] * B[i, j]  # numpy scalars of type dt: rounded to dt
                want[i, j] = Cm[i, j] + p
                if M[i, j]:
                    wmul[i, j] = p
        assert np.array_equal(oracle.dense_fuse_mult_add(A, B, Cm), want), dt
        assert np.array_equal(oracle.dense_mult_masked(A, B, Mrp, Mcol), wmul), dt
