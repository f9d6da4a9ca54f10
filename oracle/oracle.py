"""ORACLE -- test infrastructure only.

ctypes wrapper of oracle/_build/liboracle.so (the C restatement in oracle.c)
and of the reference harness oracle/_ref/ref_driver, plus a numpy restatement
of the element-wise operators (ewise, scale; pinned in tests/test_oracle.py by
dense brute force, invariants and the reference's own eWiseAdd and
add_nointersect).  Only tests/, the smoke
check and bench.py's cpu_baseline leg import this module, and only as the
checker or the timed CPU baseline; the product (graphalgo_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
REF = HERE / "_ref" / "ref_driver"

I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
F64P = C.POINTER(C.c_double)

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.orc_mxm_masked_i64.argtypes = [C.c_int64, C.c_int64, I32P, I32P, I32P, I32P, I32P, I32P, I32P, I32P,
                                            I64P]
        _lib.orc_mxm_masked_f64.argtypes = [C.c_int64, C.c_int64, I32P, I32P, F64P, I32P, I32P, F64P, I32P, I32P,
                                            F64P]
        _lib.orc_wedges.argtypes = [C.c_int64, I32P, I32P, I32P]
        _lib.orc_wedges.restype = C.c_int64
        _lib.orc_tc_sandia.argtypes = [C.c_int64, I32P, I32P]
        _lib.orc_tc_sandia.restype = C.c_int64
        _lib.orc_tc_lu.argtypes = [C.c_int64, I32P, I32P]
        _lib.orc_tc_lu.restype = C.c_int64
        _lib.orc_tc_vertex.argtypes = [C.c_int64, I32P, I32P, I64P]
        _lib.orc_ktruss.argtypes = [C.c_int64, I32P, I32P, C.c_int, C.POINTER(I32P), C.POINTER(I32P), I64P,
                                    C.POINTER(C.c_int)]
        _lib.orc_truss_decomposition.argtypes = [C.c_int64, I32P, I32P, I32P, I32P, I64P, I64P, I64P, C.c_int64]
        _lib.orc_mxm_cmask_f64.argtypes = [C.c_int64, C.c_int64, I32P, I32P, F64P, I32P, I32P, F64P, I32P, I32P,
                                           C.POINTER(I32P), C.POINTER(I32P), C.POINTER(F64P), I64P]
        _lib.orc_mxm_spd_f64.argtypes = [C.c_int64, C.c_int64, I32P, I32P, F64P, F64P, I32P, I32P, F64P]
        _lib.orc_bc_batch.argtypes = [C.c_int64, I32P, I32P, I32P, C.c_int64, F64P, I32P]
        _lib.orc_free.argtypes = [C.c_void_p]
        _lib.orc_num_threads.restype = C.c_int
    return _lib


def _p(a, t=I32P):
    return None if a is None else a.ctypes.data_as(t)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def mxm_masked(m, bn, Arp, Acol, Brp, Bcol, Mrp, Mcol, Aval=None, Bval=None, dtype=np.int64):
    """C<M> = A*B (include/matrix_la.h:715-749); values aligned with Mcol."""
    Arp, Acol, Brp, Bcol, Mrp, Mcol = map(_i32, (Arp, Acol, Brp, Bcol, Mrp, Mcol))
    out = np.zeros(len(Mcol), dtype=np.float64 if dtype == np.float64 else np.int64)
    if dtype == np.float64:
        Av = None if Aval is None else np.ascontiguousarray(Aval, dtype=np.float64)
        Bv = None if Bval is None else np.ascontiguousarray(Bval, dtype=np.float64)
        rc = lib().orc_mxm_masked_f64(m, bn, _p(Arp), _p(Acol), _p(Av, F64P), _p(Brp), _p(Bcol), _p(Bv, F64P),
                                      _p(Mrp), _p(Mcol), _p(out, F64P))
    else:
        Av, Bv = _i32(Aval), _i32(Bval)
        rc = lib().orc_mxm_masked_i64(m, bn, _p(Arp), _p(Acol), _p(Av), _p(Brp), _p(Bcol), _p(Bv), _p(Mrp),
                                      _p(Mcol), _p(out, I64P))
    assert rc == 0, rc
    return out


def wedges(g, b=None) -> int:
    b = g if b is None else b
    return lib().orc_wedges(g.n, _p(_i32(g.rowptr)), _p(_i32(g.col)), _p(_i32(b.rowptr)))


def tc(g) -> int:
    """triangle_counting_masked_sandia, src/main.cpp:65-86."""
    return lib().orc_tc_sandia(g.n, _p(_i32(g.rowptr)), _p(_i32(g.col)))


def tc_lu(g) -> int:
    return lib().orc_tc_lu(g.n, _p(_i32(g.rowptr)), _p(_i32(g.col)))


def tc_vertex(g) -> np.ndarray:
    out = np.zeros(g.n, dtype=np.int64)
    assert lib().orc_tc_vertex(g.n, _p(_i32(g.rowptr)), _p(_i32(g.col)), _p(out, I64P)) == 0
    return out


def ktruss(g, k: int):
    """k_truss, src/main.cpp:89-146: (rowptr, col, iterations)."""
    rp, col, nnz, it = I32P(), I32P(), C.c_int64(), C.c_int()
    assert lib().orc_ktruss(g.n, _p(_i32(g.rowptr)), _p(_i32(g.col)), k, C.byref(rp), C.byref(col), C.byref(nnz),
                            C.byref(it)) == 0
    r = np.ctypeslib.as_array(rp, shape=(g.n + 1,)).copy()
    c = np.ctypeslib.as_array(col, shape=(max(nnz.value, 1),))[: nnz.value].copy()
    lib().orc_free(C.cast(rp, C.c_void_p))
    lib().orc_free(C.cast(col, C.c_void_p))
    return r, c, it.value


def truss_decomposition(g, max_passes: int = 100000):
    """(truss[nnz], kmax, passes, pass_nnz, pass_wedges)."""
    truss = np.zeros(max(g.nnz, 1), dtype=np.int32)
    kmax = C.c_int32()
    passes = C.c_int64()
    pn = np.zeros(max_passes, dtype=np.int64)
    pw = np.zeros(max_passes, dtype=np.int64)
    assert lib().orc_truss_decomposition(g.n, _p(_i32(g.rowptr)), _p(_i32(g.col)), _p(truss), C.byref(kmax),
                                         C.byref(passes), _p(pn, I64P), _p(pw, I64P), max_passes) == 0
    np_ = min(passes.value, max_passes)
    return truss[: g.nnz], kmax.value, passes.value, pn[:np_], pw[:np_]


def mxm_cmask(m, bn, Arp, Acol, Brp, Bcol, Mrp, Mcol, Aval=None, Bval=None):
    """C<!M> = A*B (include/matrix_la.h:800-891): (rowptr, col, val)."""
    Arp, Acol, Brp, Bcol, Mrp, Mcol = map(_i32, (Arp, Acol, Brp, Bcol, Mrp, Mcol))
    Av = None if Aval is None else np.ascontiguousarray(Aval, dtype=np.float64)
    Bv = None if Bval is None else np.ascontiguousarray(Bval, dtype=np.float64)
    rp, col, val, nnz = I32P(), I32P(), F64P(), C.c_int64()
    assert lib().orc_mxm_cmask_f64(m, bn, _p(Arp), _p(Acol), _p(Av, F64P), _p(Brp), _p(Bcol), _p(Bv, F64P),
                                   _p(Mrp), _p(Mcol), C.byref(rp), C.byref(col), C.byref(val), C.byref(nnz)) == 0
    nz = nnz.value
    r = np.ctypeslib.as_array(rp, shape=(m + 1,)).copy()
    c = np.ctypeslib.as_array(col, shape=(max(nz, 1),))[:nz].copy()
    v = np.ctypeslib.as_array(val, shape=(max(nz, 1),))[:nz].copy()
    for p in (rp, col, val):
        lib().orc_free(C.cast(p, C.c_void_p))
    return r, c, v


def mxm_spd(m, ncols, Arp, Acol, B, Mrp, Mcol, Aval=None):
    """mxmm_spd (include/matrix_la.h:239-270) with FP64 values."""
    Arp, Acol, Mrp, Mcol = map(_i32, (Arp, Acol, Mrp, Mcol))
    Av = None if Aval is None else np.ascontiguousarray(Aval, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    out = np.zeros((m, ncols), dtype=np.float64)
    lib().orc_mxm_spd_f64(m, ncols, _p(Arp), _p(Acol), _p(Av, F64P), _p(B, F64P), _p(Mrp), _p(Mcol),
                          _p(out, F64P))
    return out


def bc(g, sources):
    """FP64 restatement of betweenness_centrality_batch (include/betweenness_centrality.h:125-237)."""
    src = np.ascontiguousarray(sources, dtype=np.int32)
    out = np.zeros(g.n, dtype=np.float64)
    depth = C.c_int32()
    rc = lib().orc_bc_batch(g.n, _p(_i32(g.rowptr)), _p(_i32(g.col)), _p(src), len(src), _p(out, F64P),
                            C.byref(depth))
    assert rc == 0, rc
    return out


def num_threads() -> int:
    return lib().orc_num_threads()


# ---- the compiled reference (oracle/_ref), present where it was built ------
def ref_available() -> bool:
    return REF.exists()


def ref_run(*args: str) -> dict:
    out = subprocess.run([str(REF), *args], capture_output=True, text=True, check=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


def ref_truss(graph_path: str, nnz: int) -> tuple[np.ndarray, dict]:
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "truss.bin")
        info = ref_run("ktall", graph_path, out)
        return np.fromfile(out, dtype=np.int32, count=nnz), info


def ref_bc(graph_path: str, batch: int, n: int) -> np.ndarray:
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "bc.bin")
        ref_run("bc", graph_path, str(batch), out)
        return np.fromfile(out, dtype=np.float32, count=n)


# ---- element-wise operators (numpy restatement) -----------------------------
# include/matrix_la.h:272-541 with the reference's defects fixed (the masked
# eWiseAdd's second loop reads A's columns for B, :294/:321; the unmasked
# eWiseMult's fill loop advances the row index instead of the output
# position, :472; the masked eWiseMult's count and fill walks disagree,
# :497-530).  Keys are row * n + col over sorted rows, so every result comes
# out in CSR order.
EW_ADD, EW_MULT, EW_ADD_DISJOINT = 0, 1, 2


def _keys(n, rp, col):
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(np.asarray(rp, dtype=np.int64)))
    return rows * n + np.asarray(col, dtype=np.int64)


def _csr_from_keys(n, keys):
    rows = keys // n
    rp = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    return rp, (keys % n).astype(np.int32)


def ewise(op, n, Arp, Acol, Aval, Brp, Bcol, Bval, Mrp=None, Mcol=None, dtype=np.int64):
    """(rowptr, col, val) of eWiseAdd / eWiseMult / add_nointersect on n x n
    operands; None values are ones; M (rowptr, col) optional."""
    dt = np.dtype(dtype)
    ka, kb = _keys(n, Arp, Acol), _keys(n, Brp, Bcol)
    va = np.ones(len(ka), dt) if Aval is None else np.asarray(Aval, dtype=dt)
    vb = np.ones(len(kb), dt) if Bval is None else np.asarray(Bval, dtype=dt)

    def lookup(keys, k, v):
        p = np.searchsorted(k, keys)
        hit = (p < len(k)) & (k[np.minimum(p, max(len(k) - 1, 0))] == keys) if len(k) else np.zeros(len(keys), bool)
        out = np.zeros(len(keys), dt)
        out[hit] = v[p[hit]]
        return hit, out

    if Mrp is not None:
        km = _keys(n, Mrp, Mcol)
        ha, xa = lookup(km, ka, va)
        hb, xb = lookup(km, kb, vb)
        if op == EW_MULT:
            keep = ha & hb
            val = (xa * xb)[keep]
        else:
            s = xa + xb
            keep = s != 0
            val = s[keep]
        rp, col = _csr_from_keys(n, km[keep])
        return rp, col, val
    if op == EW_MULT:
        k = np.intersect1d(ka, kb)
        _, xa = lookup(k, ka, va)
        _, xb = lookup(k, kb, vb)
        rp, col = _csr_from_keys(n, k)
        return rp, col, xa * xb
    if op == EW_ADD_DISJOINT and len(np.intersect1d(ka, kb)):
        raise ValueError("add_nointersect: the operands share a position")
    k = np.union1d(ka, kb)
    ha, xa = lookup(k, ka, va)
    hb, xb = lookup(k, kb, vb)
    rp, col = _csr_from_keys(n, k)
    return rp, col, xa + xb


def mxm_spd_seq(Arp, Acol, Aval, B, Mrp, Mcol, dtype=np.float32):
    """mxmm_spd (include/matrix_la.h:239-270) in the value type `dtype`, in
    the reference's order: row i of C starts at zero and, for each stored
    A(i,k) in CSR order, C(i, M's columns) += A(i,k) * B(k, M's columns),
    every product and sum rounded to dtype (numpy rounds each operation).
    Columns of row i outside M stay zero.  Aval None = ones."""
    dt = np.dtype(dtype)
    Bd = np.asarray(B, dtype=dt)
    m = len(Arp) - 1
    out = np.zeros((m, Bd.shape[1]), dtype=dt)
    for i in range(m):
        cols = np.asarray(Mcol[Mrp[i]:Mrp[i + 1]], dtype=np.int64)
        if not len(cols):
            continue
        acc = np.zeros(len(cols), dtype=dt)
        for t in range(Arp[i], Arp[i + 1]):
            a = dt.type(1) if Aval is None else dt.type(Aval[t])
            acc = acc + a * Bd[Acol[t], cols]
        out[i, cols] = acc
    return out


def dense_fuse_mult_add(A, B, Cm):
    """fuseEWiseMultAdd (include/matrix_la.h:192-198): C + A .* B, the
    product and the sum each rounded to C's type."""
    dt = np.asarray(Cm).dtype
    return np.asarray(Cm, dtype=dt) + np.asarray(A, dtype=dt) * np.asarray(B, dtype=dt)


def dense_mult_masked(A, B, Mrp, Mcol):
    """The dense eWiseMult (include/matrix_la.h:200-219): C zeroed, then
    C(i,j) = A(i,j) * B(i,j) at M's entries."""
    A = np.asarray(A)
    out = np.zeros_like(A)
    rows = np.repeat(np.arange(len(Mrp) - 1), np.diff(np.asarray(Mrp, dtype=np.int64)))
    cols = np.asarray(Mcol, dtype=np.int64)
    out[rows, cols] = A[rows, cols] * np.asarray(B, dtype=A.dtype)[rows, cols]
    return out


def scale(val, alpha, dtype=np.int64):
    """multScalar (include/matrix_la.h:535-541): v * alpha in FP64, converted
    back to the value type toward zero."""
    dt = np.dtype(dtype)
    return (np.asarray(val, dtype=dt).astype(np.float64) * float(alpha)).astype(dt)
