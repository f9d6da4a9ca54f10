"""GPU tests: the element-wise operators (include/matrix_la.h:272-541; eWiseAdd,
eWiseMult with and without a mask, add_nointersect, multScalar) through the
C ABI, against the oracle: bit-exact patterns and values for int32, int64 and
float64 (sums and products of small integers are exact in FP64 too), on random
sparse operands, graphs with hub rows (rows of several 64-entry chunks), empty
rows and matrices, pattern operands (values one) and the disjoint-union error;
at R-MAT 18 through properties fixed by the algebra (A + A = 2A, A .* A = A^2,
L + U = A, A + (-A) on A's pattern is empty)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DTYPES = [np.int32, np.int64, np.float64]


def _random_csr(G, n, nnz, seed):
    rng = np.random.default_rng(seed)
    keys = np.unique(rng.integers(0, n * n, size=nnz))
    rows = keys // n
    rp = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    return G.CSR(n, rp, (keys % n).astype(np.int32))


def _vals(rng, nnz, dt):
    return rng.integers(-3, 4, size=nnz).astype(dt)  # explicit zeros included


def _same(got, want):
    (gc, gv), (rp, col, val) = got, want
    return (np.array_equal(gc.rowptr, rp) and np.array_equal(gc.col, col) and gv.dtype == val.dtype
            and np.array_equal(gv, val))


def _cases(G):
    return {
        "random": (_random_csr(G, 300, 6000, 1), _random_csr(G, 300, 6000, 2), _random_csr(G, 300, 20000, 3)),
        "rmat12": (G.rmat(12, 16, seed=21), G.rmat(12, 16, seed=22), G.rmat(12, 8, seed=23)),
    }


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("case", ["random", "rmat12"])
def test_ewise_matches_oracle(gpu, G, oracle, case, dt):
    A, B, M = _cases(G)[case]
    rng = np.random.default_rng(7)
    va, vb = _vals(rng, A.nnz, dt), _vals(rng, B.nnz, dt)
    n = A.n
    print(f"case ewise {case} {np.dtype(dt).name}", flush=True)
    for op, fn in ((oracle.EW_ADD, gpu.ewise_add), (oracle.EW_MULT, gpu.ewise_mult)):
        want = oracle.ewise(op, n, A.rowptr, A.col, va, B.rowptr, B.col, vb, dtype=dt)
        assert _same(fn(A, va, B, vb, dtype=dt), want), op
        want = oracle.ewise(op, n, A.rowptr, A.col, va, B.rowptr, B.col, vb, M.rowptr, M.col, dtype=dt)
        assert _same(fn(A, va, B, vb, M, dtype=dt), want), (op, "masked")
    # pattern operands hold ones
    want = oracle.ewise(oracle.EW_ADD, n, A.rowptr, A.col, None, B.rowptr, B.col, None, dtype=dt)
    assert _same(gpu.ewise_add(A, None, B, None, dtype=dt), want)
    want = oracle.ewise(oracle.EW_MULT, n, A.rowptr, A.col, va, B.rowptr, B.col, None, M.rowptr, M.col, dtype=dt)
    assert _same(gpu.ewise_mult(A, va, B, None, M, dtype=dt), want)


def test_ewise_masked_sum_cancels(gpu, G, oracle):
    """Masked eWiseAdd drops positions whose sum is zero; the unmasked one keeps them."""
    A = G.rmat(10, 8, seed=5)
    va = np.random.default_rng(1).integers(1, 9, size=A.nnz).astype(np.int64)
    C, v = gpu.ewise_add(A, va, A, -va, A)
    assert C.nnz == 0 and np.array_equal(C.rowptr, np.zeros(A.n + 1))
    C, v = gpu.ewise_add(A, va, A, -va)
    assert np.array_equal(C.col, A.col) and not v.any()


def test_ewise_hub_rows(gpu, G, oracle):
    """Rows of 1000+ entries in both operands (several 64-entry chunks, matches in
    every chunk) next to empty rows."""
    n = 3000
    rng = np.random.default_rng(9)
    rows_a, rows_b = [], []
    for r in (0, 5, 2999):
        rows_a.append(np.sort(rng.choice(n, 1500, replace=False)))
        rows_b.append(np.sort(rng.choice(n, 1200, replace=False)))

    def csr(rows):
        rp = np.zeros(n + 1, dtype=np.int32)
        for r, c in zip((0, 5, 2999), rows):
            rp[r + 1] = len(c)
        np.cumsum(rp, out=rp)
        return G.CSR(n, rp, np.concatenate(rows).astype(np.int32))
    A, B = csr(rows_a), csr(rows_b)
    va, vb = _vals(rng, A.nnz, np.int64), _vals(rng, B.nnz, np.int64)
    for op, fn in ((oracle.EW_ADD, gpu.ewise_add), (oracle.EW_MULT, gpu.ewise_mult)):
        want = oracle.ewise(op, n, A.rowptr, A.col, va, B.rowptr, B.col, vb)
        assert _same(fn(A, va, B, vb), want)
        want = oracle.ewise(op, n, A.rowptr, A.col, va, B.rowptr, B.col, vb, B.rowptr, B.col)
        assert _same(fn(A, va, B, vb, B), want)


def test_ewise_empty(gpu, G, oracle):
    E = G.CSR(40, np.zeros(41, dtype=np.int32), np.zeros(0, dtype=np.int32))
    A = _random_csr(G, 40, 200, 4)
    va = np.arange(A.nnz, dtype=np.int64)
    C, v = gpu.ewise_add(E, None, E, None)
    assert C.nnz == 0 and len(v) == 0
    C, v = gpu.ewise_add(A, va, E, None)
    assert np.array_equal(C.col, A.col) and np.array_equal(v, va)
    C, v = gpu.ewise_mult(A, va, E, None)
    assert C.nnz == 0
    C, v = gpu.add_nointersect(E, None, A, va)
    assert np.array_equal(C.col, A.col) and np.array_equal(v, va)
    Z = G.CSR(0, np.zeros(1, dtype=np.int32), np.zeros(0, dtype=np.int32))
    C, v = gpu.ewise_add(Z, None, Z, None)
    assert C.nnz == 0 and len(C.rowptr) == 1
    assert len(gpu.mult_scalar(E, np.zeros(0, dtype=np.int64), 2.0)) == 0


@pytest.mark.parametrize("dt", DTYPES)
def test_add_nointersect(gpu, G, oracle, dt):
    A = G.rmat(11, 16, seed=31)
    n = A.n
    rows = np.repeat(np.arange(n), np.diff(A.rowptr))
    low, up = A.col < rows, A.col > rows

    def part(sel):
        rp = np.zeros(n + 1, dtype=np.int32)
        np.cumsum(np.bincount(rows[sel], minlength=n), out=rp[1:])
        return G.CSR(n, rp, A.col[sel].astype(np.int32))
    L, U = part(low), part(up)
    rng = np.random.default_rng(2)
    vl, vu = _vals(rng, L.nnz, dt), _vals(rng, U.nnz, dt)
    want = oracle.ewise(oracle.EW_ADD_DISJOINT, n, L.rowptr, L.col, vl, U.rowptr, U.col, vu, dtype=dt)
    got = gpu.add_nointersect(L, vl, U, vu, dtype=dt)
    assert _same(got, want)
    assert np.array_equal(got[0].col, A.col)
    with pytest.raises(Exception, match="share a position"):
        gpu.add_nointersect(A, None, L, None, dtype=dt)


@pytest.mark.parametrize("dt", DTYPES)
def test_mult_scalar(gpu, G, oracle, dt):
    A = _random_csr(G, 500, 20000, 6)
    v = np.random.default_rng(3).integers(-1000, 1000, size=A.nnz).astype(dt)
    for alpha in (0.5, -1.5, 3.0, 0.0, 1e-3):
        assert np.array_equal(gpu.mult_scalar(A, v, alpha, dtype=dt), oracle.scale(v, alpha, dtype=dt)), alpha


def test_ewise_errors(gpu, G):
    A = _random_csr(G, 50, 300, 8)
    B = _random_csr(G, 60, 300, 9)
    with pytest.raises(Exception):
        gpu.ewise_add(A, None, B, None)  # shapes differ
    with pytest.raises(Exception):
        gpu.ewise_mult(A, None, A, None, B)  # mask shape differs
    with pytest.raises(Exception):
        gpu.mult_scalar(A, None, 2.0)


def test_ewise_rmat18_properties(gpu, G):
    """R-MAT 18 (about 7.5M entries): A + A = 2A, A .* A = A^2 (values), L + U = A
    with add_nointersect, A + (-A) masked by A is empty."""
    A = G.rmat(18, 16, seed=18)
    n = A.n
    rng = np.random.default_rng(18)
    va = rng.integers(-1000, 1000, size=A.nnz).astype(np.int64)
    print(f"case ewise rmat18 nnz={A.nnz}", flush=True)
    C, v = gpu.ewise_add(A, va, A, va)
    assert np.array_equal(C.rowptr, A.rowptr) and np.array_equal(C.col, A.col) and np.array_equal(v, 2 * va)
    C, v = gpu.ewise_mult(A, va, A, va, A)
    assert np.array_equal(C.col, A.col) and np.array_equal(v, va * va)
    assert np.array_equal(gpu.mult_scalar(A, va, -2.0), -2 * va)
    C, v = gpu.ewise_add(A, va, A, -va, A)
    assert C.nnz == 0
    rows = np.repeat(np.arange(n), np.diff(A.rowptr))
    low = A.col < rows

    def part(sel):
        rp = np.zeros(n + 1, dtype=np.int32)
        np.cumsum(np.bincount(rows[sel], minlength=n), out=rp[1:])
        return G.CSR(n, rp, A.col[sel].astype(np.int32))
    C, v = gpu.add_nointersect(part(low), va[low], part(~low), va[~low])
    assert np.array_equal(C.rowptr, A.rowptr) and np.array_equal(C.col, A.col) and np.array_equal(v, va)
