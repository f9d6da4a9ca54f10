"""Graph substrate: CSR container, deterministic generators and loaders.

Every generator is a thin wrapper over ``libgagraph.so`` (host C++), whose
random draws are a pure function of (seed, stream, counter), so the same call
returns the same graph on any x86-64 host.  All graphs come out symmetric,
deduplicated, loop-free, with sorted rows -- the input contract of the GPU
kernels and of the reference's TC/k-truss path (which symmetrises the lower
triangle, src/main.cpp:221-222).
"""
from __future__ import annotations

import ctypes as C
import hashlib
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from ._native import GaCsr, I32P, I64P, hostlib

RMAT_ABCD = (0.57, 0.19, 0.19)  # d = 0.05


@dataclass
class CSR:
    """Square pattern matrix in CSR form (int32 row pointers and columns)."""

    n: int
    rowptr: np.ndarray
    col: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    @property
    def m_undirected(self) -> int:
        return self.nnz // 2

    def ga(self) -> GaCsr:
        """A ga_csr view (keeps references to the arrays on the object)."""
        self.rowptr = np.ascontiguousarray(self.rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(self.col, dtype=np.int32)
        return GaCsr(self.n, self.n, self.nnz, self.rowptr.ctypes.data_as(I32P),
                     self.col.ctypes.data_as(I32P), None)

    def sha256(self) -> str:
        h = hashlib.sha256()
        h.update(np.int64(self.n).tobytes())
        h.update(np.ascontiguousarray(self.rowptr, dtype=np.int32).tobytes())
        h.update(np.ascontiguousarray(self.col, dtype=np.int32).tobytes())
        return h.hexdigest()

    def degrees(self) -> np.ndarray:
        return np.diff(self.rowptr).astype(np.int64)

    def wedges(self, other: "CSR | None" = None) -> int:
        """W = sum_{(i,k) in self} nnz(other_k): Gustavson intermediate products."""
        b = self if other is None else other
        return int(np.diff(b.rowptr).astype(np.int64)[self.col].sum())

    def lower(self) -> "CSR":
        """Strict lower triangle (include/matrix.h:736-763)."""
        rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.rowptr))
        keep = self.col < rows
        cnt = np.bincount(rows[keep], minlength=self.n)
        rp = np.zeros(self.n + 1, dtype=np.int32)
        np.cumsum(cnt, out=rp[1:])
        return CSR(self.n, rp, self.col[keep].astype(np.int32))

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((np.ones(self.nnz, dtype=np.int64), self.col, self.rowptr), shape=(self.n, self.n))

    def write_csr(self, path: str) -> None:
        """Binary CSR: int64 n, int64 nnz, int32 rowptr[n+1], int32 col[nnz]."""
        with open(path, "wb") as f:
            f.write(np.int64(self.n).tobytes())
            f.write(np.int64(self.nnz).tobytes())
            f.write(np.ascontiguousarray(self.rowptr, dtype=np.int32).tobytes())
            f.write(np.ascontiguousarray(self.col, dtype=np.int32).tobytes())

    @staticmethod
    def from_edges(n: int, u, v) -> "CSR":
        """Symmetric CSR from an undirected edge list (loops dropped, deduplicated)."""
        u = np.asarray(u, dtype=np.int64)
        v = np.asarray(v, dtype=np.int64)
        keep = u != v
        u, v = u[keep], v[keep]
        a = np.concatenate([u, v])
        b = np.concatenate([v, u])
        key = np.unique(a * n + b)
        rows, cols = key // n, key % n
        rp = np.zeros(n + 1, dtype=np.int32)
        np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
        return CSR(n, rp, cols.astype(np.int32))


def _take(rp_p, col_p, n: int, nnz: int) -> CSR:
    lib = hostlib()
    rp = np.ctypeslib.as_array(rp_p, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(col_p, shape=(max(nnz, 1),))[:nnz].copy()
    lib.gah_free(C.cast(rp_p, C.c_void_p))
    lib.gah_free(C.cast(col_p, C.c_void_p))
    return CSR(n, rp, col)


def _hcheck(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(f"graph generator failed ({rc}): {hostlib().gah_last_error().decode()}")


def erdos_renyi(n: int, m: int, seed: int) -> CSR:
    """G(n, m): m distinct uniform unordered pairs, no self-loops."""
    rp, col, nnz = I32P(), I32P(), C.c_int64()
    _hcheck(hostlib().gah_gen_er(n, m, seed, C.byref(rp), C.byref(col), C.byref(nnz)))
    return _take(rp, col, n, nnz.value)


def rmat(scale: int, edge_factor: int, seed: int, abc=RMAT_ABCD) -> CSR:
    """R-MAT(scale, edge_factor, a, b, c), symmetrised, deduplicated, loops removed."""
    rp, col, n, nnz = I32P(), I32P(), C.c_int64(), C.c_int64()
    a, b, c = abc
    _hcheck(hostlib().gah_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(rp), C.byref(col),
                                   C.byref(n), C.byref(nnz)))
    return _take(rp, col, n.value, nnz.value)


def splitmix64(x) -> np.ndarray:
    """SplitMix64 finaliser (vectorised, uint64 wrap-around)."""
    z = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def pick_sources(g: CSR, count: int, seed: int) -> np.ndarray:
    """`count` distinct non-isolated vertices, ordered by a SplitMix64 hash of
    (seed, vertex): a seeded uniform choice that does not depend on numpy's RNG."""
    v = np.nonzero(g.degrees() > 0)[0].astype(np.uint64)
    with np.errstate(over="ignore"):
        z = splitmix64(v + np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) - np.uint64(0x9E3779B97F4A7C15))
    return v[np.argsort(z, kind="stable")[:count]].astype(np.int32)


def knn_params(base_seed: int, g: int) -> tuple[int, int, int]:
    n, d, c = C.c_int64(), C.c_int(), C.c_int()
    hostlib().gah_knn_params(base_seed, g, C.byref(n), C.byref(d), C.byref(c))
    return n.value, d.value, c.value


def knn(seed: int, n: int, dim: int, ncomp: int, k: int = 30, return_points: bool = False):
    """Exact FP64 kNN graph of a seeded Gaussian mixture, symmetrised by union."""
    rp, col, nnz = I32P(), I32P(), C.c_int64()
    pts = np.zeros((n, dim), dtype=np.float64) if return_points else None
    _hcheck(hostlib().gah_gen_knn(seed, n, dim, ncomp, k, C.byref(rp), C.byref(col), C.byref(nnz),
                                  pts.ctypes.data_as(C.POINTER(C.c_double)) if pts is not None else None))
    g = _take(rp, col, n, nnz.value)
    return (g, pts) if return_points else g


def knn_batch(base_seed: int, count: int, k: int = 30, workers: int = 8) -> list[CSR]:
    """`count` kNN graphs; graph g uses seed base_seed + g and knn_params(base_seed, g)."""
    def one(g):
        n, d, c = knn_params(base_seed, g)
        return knn(base_seed + g, n, d, c, k)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, range(count)))


def read_mtx(path: str) -> CSR:
    """Matrix Market reader with the reference's semantics (include/matrix.h:280-417):
    diagonal dropped, symmetric entries mirrored, values ignored; rows sorted."""
    rp, col, n, nnz, sym = I32P(), I32P(), C.c_int64(), C.c_int64(), C.c_int()
    _hcheck(hostlib().gah_read_mtx(str(path).encode(), C.byref(n), C.byref(rp), C.byref(col), C.byref(nnz),
                                   C.byref(sym)))
    return _take(rp, col, n.value, nnz.value)


def symmetrize(g: CSR) -> CSR:
    """A + A^T, loops dropped (the reference's build_symm_from_lower(extract_lower_triangle(A))
    equals this for symmetric inputs; for general inputs the reference keeps only
    the lower triangle, see `symm_from_lower`)."""
    rp, col, nnz = I32P(), I32P(), C.c_int64()
    _hcheck(hostlib().gah_symmetrize(g.n, g.rowptr.ctypes.data_as(I32P), g.col.ctypes.data_as(I32P),
                                     C.byref(rp), C.byref(col), C.byref(nnz)))
    return _take(rp, col, g.n, nnz.value)


def symm_from_lower(g: CSR) -> CSR:
    """build_symm_from_lower(extract_lower_triangle(g)) (include/matrix.h:700-763),
    which is what the reference's CLI feeds to TC, k-truss and mxm (src/main.cpp:221-222)."""
    L = g.lower()
    rows = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(L.rowptr))
    return CSR.from_edges(g.n, rows, L.col)
