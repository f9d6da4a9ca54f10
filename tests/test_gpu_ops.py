"""GPU parity of the complement-masked SpGEMM and the masked sparse x dense
product, and the reference-style C++ client of the drop-in headers."""
import ctypes as C
import json
import subprocess

import numpy as np
import pytest

from conftest import GRAPH_DIR, ROOT

pytestmark = pytest.mark.gpu


def _rand_csr(rng, m, n, density, values=True):
    D = (rng.random((m, n)) < density)
    rp = np.concatenate([[0], np.cumsum(D.sum(1))]).astype(np.int32)
    r, c = np.nonzero(D)
    val = rng.integers(1, 5, size=len(c)).astype(np.float64) if values else None
    return rp, c.astype(np.int32), val


def _view(N, rp, col, nrows, ncols, val=None):
    return N.GaCsr(nrows, ncols, int(rp[-1]), rp.ctypes.data_as(N.I32P), col.ctypes.data_as(N.I32P),
                   None if val is None else val.ctypes.data_as(C.c_void_p))


def _cmask(N, A, B, M, dtype):
    rp, col, val, nnz = N.I32P(), N.I32P(), C.c_void_p(), C.c_int64()
    N.check(N.galib().ga_mxm_cmask(C.byref(A), C.byref(B), C.byref(M), dtype, C.byref(rp), C.byref(col),
                                   C.byref(val), C.byref(nnz)))
    nz = nnz.value
    m = A.nrows
    r = np.ctypeslib.as_array(rp, shape=(m + 1,)).copy()
    c = np.ctypeslib.as_array(col, shape=(max(nz, 1),))[:nz].copy()
    vp = C.cast(val, C.POINTER(C.c_double if dtype == N.GA_FLOAT64 else C.c_int64))
    v = np.ctypeslib.as_array(vp, shape=(max(nz, 1),))[:nz].copy()
    for p in (rp, col, val):
        N.galib().ga_free_host(C.cast(p, C.c_void_p))
    return r, c, v


@pytest.mark.parametrize("seed", range(3))
def test_cmask_random_vs_oracle(gpu, oracle, seed):
    from graphalgo_amd import _native as N
    rng = np.random.default_rng(seed)
    m, k, n = 200, 150, 300
    Arp, Acol, Aval = _rand_csr(rng, m, k, 0.05)
    Brp, Bcol, Bval = _rand_csr(rng, k, n, 0.05)
    Mrp, Mcol, _ = _rand_csr(rng, m, n, 0.2, values=False)
    want = oracle.mxm_cmask(m, n, Arp, Acol, Brp, Bcol, Mrp, Mcol, Aval=Aval, Bval=Bval)
    got = _cmask(N, _view(N, Arp, Acol, m, k, Aval), _view(N, Brp, Bcol, k, n, Bval), _view(N, Mrp, Mcol, m, n),
                 N.GA_FLOAT64)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert np.allclose(got[2], want[2], rtol=1e-13, atol=0)


def test_cmask_long_rows_vs_oracle(gpu, G, oracle):
    """Rows with more than 2048 products take the dense-accumulator path."""
    from graphalgo_amd import _native as N
    g = G.rmat(12, 16, seed=9)
    rng = np.random.default_rng(3)
    ns = 64
    Brp, Bcol, Bval = _rand_csr(rng, g.n, ns, 0.3)
    Mrp, Mcol, _ = _rand_csr(rng, g.n, ns, 0.3, values=False)
    want = oracle.mxm_cmask(g.n, ns, g.rowptr, g.col, Brp, Bcol, Mrp, Mcol, Bval=Bval)
    got = _cmask(N, _view(N, g.rowptr, g.col, g.n, g.n), _view(N, Brp, Bcol, g.n, ns, Bval),
                 _view(N, Mrp, Mcol, g.n, ns), N.GA_FLOAT64)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert np.allclose(got[2], want[2], rtol=1e-12, atol=0)


def test_spd_vs_oracle_bit_exact(gpu, G, oracle):
    from graphalgo_amd import _native as N
    g = G.rmat(11, 8, seed=5)
    rng = np.random.default_rng(4)
    ncols = 96
    B = rng.random((g.n, ncols))
    Mrp, Mcol, _ = _rand_csr(rng, g.n, ncols, 0.3, values=False)
    want = oracle.mxm_spd(g.n, ncols, g.rowptr, g.col, B, Mrp, Mcol)
    out = np.zeros((g.n, ncols))
    a = _view(N, g.rowptr, g.col, g.n, g.n)
    mm = _view(N, Mrp, Mcol, g.n, ncols)
    N.check(N.galib().ga_mxm_spd(C.byref(a), B.ctypes.data_as(N.F64P), ncols, C.byref(mm),
                                 out.ctypes.data_as(N.F64P)))
    assert np.array_equal(out, want)  # same accumulation order as the reference


def test_cpp_dropin_client(gpu, golden, small_graphs, oracle):
    exe = ROOT / "tests" / "cpp" / "_build" / "dropin"
    if not exe.exists():
        pytest.fail("tests/cpp/_build/dropin not built (run __graft_entry__.build())")
    r = subprocess.run([str(exe), str(GRAPH_DIR / "G43.mtx")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    rec = golden["graphs"]["G43.mtx"]
    per_k = {k: (nnz, it) for k, nnz, it in rec["truss"]["per_k"]}
    assert out["tc"] == out["tc_lu"] == rec["tc_msa"] == 1392
    assert (out["k3_nnz"], out["k3_it"]) == per_k[3]
    assert (out["k4_nnz"], out["k4_it"]) == per_k[4]
    g = small_graphs["G43.mtx"][1]
    assert out["sq_sum"] == oracle.mxm_masked(g.n, g.n, g.rowptr, g.col, g.rowptr, g.col, g.rowptr, g.col).sum()
    bc = oracle.bc(small_graphs["G43.mtx"][0], np.arange(16))
    assert out["bc_sum"] == pytest.approx(bc.astype(np.float32).astype(np.float64).sum(), rel=1e-6)
    assert out["kmax"] == rec["truss"]["kmax"]
    # element-wise operators through the C++ headers, against the oracle
    sq = oracle.mxm_masked(g.n, g.n, g.rowptr, g.col, g.rowptr, g.col, g.rowptr, g.col)
    half = oracle.scale(sq, 0.5, dtype=np.int32)
    add = oracle.ewise(oracle.EW_ADD, g.n, g.rowptr, g.col, None, g.rowptr, g.col, half)
    mult = oracle.ewise(oracle.EW_MULT, g.n, g.rowptr, g.col, sq, g.rowptr, g.col, sq, g.rowptr, g.col)
    assert out["half_sum"] == half.sum()
    assert (out["add_nnz"], out["add_sum"]) == (len(add[1]), add[2].sum())
    assert (out["mult_nnz"], out["mult_sum"]) == (len(mult[1]), mult[2].sum())
    assert out["cancel_nnz"] == 0
    assert (out["disj_nnz"], out["disj_same"], out["disj_err"]) == (g.nnz, 1, 1)
    # full-vertex BC (every vertex a source, 64 per block): FP64, 1e-9 relative to the oracle
    g0 = small_graphs["G43.mtx"][0]
    want = oracle.bc(g0, np.arange(g0.n))
    got = np.array(out["bc_full"])
    assert got.shape == want.shape
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9 * max(1.0, float(want.max())) * 1e-3)
