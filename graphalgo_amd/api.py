"""Python mirror of the reference's operator interface over the gfx950 library.

Names follow the reference (include/matrix_la.h, src/main.cpp,
include/betweenness_centrality.h); every call goes through the C ABI of
``libgraphalgo.so`` and runs on the GPU.  There is no CPU path: without the
built library or a GPU these functions raise.
"""
from __future__ import annotations

import ctypes as C

import weakref

import numpy as np

from . import _native as N
from .graphs import CSR


def _ga(A: CSR, val: np.ndarray | None = None, ncols: int | None = None) -> N.GaCsr:
    g = A.ga()
    if ncols is not None:
        g.ncols = ncols
    if val is not None:
        g.val = val.ctypes.data_as(C.c_void_p)
    return g


_DT = {np.dtype(np.int32): N.GA_INT32, np.dtype(np.int64): N.GA_INT64, np.dtype(np.float64): N.GA_FLOAT64,
       np.dtype(np.float32): N.GA_FLOAT32}


def mxm_masked(A: CSR, B: CSR, M: CSR, dtype=np.int64, Aval=None, Bval=None,
               shapes: tuple[int, int, int] | None = None) -> np.ndarray:
    """C<M> = A*B with a structural mask; returns C's values aligned with M.col
    (mxmm_msa / mxmm_mca / mxmm_heap, include/matrix_la.h:700-713).
    shapes = (m, k, n) for non-square operands."""
    dt = np.dtype(dtype)
    N.init()
    m, k, n = shapes if shapes else (A.n, A.n, A.n)
    # the value arrays must outlive the call: a ctypes struct field does not keep them alive
    av = None if Aval is None else np.ascontiguousarray(Aval, dtype=dt)
    bv = None if Bval is None else np.ascontiguousarray(Bval, dtype=dt)
    a = _ga(A, av, k)
    a.nrows = m
    b = _ga(B, bv, n)
    b.nrows = k
    mm = _ga(M, None, n)
    mm.nrows = m
    out = np.zeros(M.nnz, dtype=dt)
    N.check(N.galib().ga_mxm_masked(C.byref(a), C.byref(b), C.byref(mm), _DT[dt], out.ctypes.data_as(C.c_void_p)))
    return out


mxmm_msa = mxmm_mca = mxmm_heap = mxm_masked


def _ewise(op: int, A: CSR, Aval, B: CSR, Bval, M: CSR | None, dtype) -> tuple[CSR, np.ndarray]:
    dt = np.dtype(dtype)
    N.init()
    av = None if Aval is None else np.ascontiguousarray(Aval, dtype=dt)
    bv = None if Bval is None else np.ascontiguousarray(Bval, dtype=dt)
    a, b = _ga(A, av), _ga(B, bv)
    mm = None if M is None else _ga(M)
    rp, col, val, nnz = N.I32P(), N.I32P(), C.c_void_p(), C.c_int64()
    N.check(N.galib().ga_ewise(op, C.byref(a), C.byref(b), C.byref(mm) if mm is not None else None, _DT[dt],
                               C.byref(rp), C.byref(col), C.byref(val), C.byref(nnz)))
    nz = nnz.value
    lib = N.galib()
    try:
        r = np.ctypeslib.as_array(rp, shape=(A.n + 1,)).copy()
        c = np.ctypeslib.as_array(col, shape=(max(nz, 1),))[:nz].copy()
        v = np.ctypeslib.as_array(C.cast(val, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(max(nz, 1),))[:nz].copy()
    finally:
        for p in (rp, col, val):
            lib.ga_free_host(C.cast(p, C.c_void_p))
    return CSR(A.n, r, c), v


def ewise_add(A: CSR, Aval, B: CSR, Bval, M: CSR | None = None, dtype=np.int64) -> tuple[CSR, np.ndarray]:
    """eWiseAdd (include/matrix_la.h:272-335 with a mask, :375-436 without):
    the union of the patterns with values added; with M, C<M> = A + B on M's
    pattern without the positions whose sum is zero.  Returns (pattern, values)."""
    return _ewise(N.GA_EW_ADD, A, Aval, B, Bval, M, dtype)


def ewise_mult(A: CSR, Aval, B: CSR, Bval, M: CSR | None = None, dtype=np.int64) -> tuple[CSR, np.ndarray]:
    """eWiseMult (include/matrix_la.h:438-533): the intersection with values
    multiplied, restricted to M's pattern when M is given."""
    return _ewise(N.GA_EW_MULT, A, Aval, B, Bval, M, dtype)


def add_nointersect(A: CSR, Aval, B: CSR, Bval, dtype=np.int64) -> tuple[CSR, np.ndarray]:
    """add_nointersect (include/matrix_la.h:337-373): the union of two
    patterns that share no position (raises when they do)."""
    return _ewise(N.GA_EW_ADD_DISJOINT, A, Aval, B, Bval, None, dtype)


def mult_scalar(A: CSR, Aval, alpha: float, dtype=np.int64) -> np.ndarray:
    """multScalar (include/matrix_la.h:535-541): A's values times alpha in the
    value type (integers converted back toward zero)."""
    dt = np.dtype(dtype)
    N.init()
    av = None if Aval is None else np.ascontiguousarray(Aval, dtype=dt)
    out = np.zeros(max(A.nnz, 1), dtype=dt)
    a = _ga(A, av)
    N.check(N.galib().ga_scale(C.byref(a), _DT[dt], float(alpha), out.ctypes.data_as(C.c_void_p)))
    return out[: A.nnz]


def mxm_spd(A: CSR, B: np.ndarray, M: CSR, Aval=None, dtype=np.float64) -> np.ndarray:
    """mxmm_spd (include/matrix_la.h:239-270): C<M> = A*B with B (A.ncols x n)
    and C (A.nrows x n) dense row-major; C is zero outside M's pattern.
    dtype float64 or float32 (the reference's float instantiation,
    include/betweenness_centrality.h:210)."""
    dt = np.dtype(dtype)
    N.init()
    Bd = np.ascontiguousarray(B, dtype=dt)
    ncols = Bd.shape[1]
    av = None if Aval is None else np.ascontiguousarray(Aval, dtype=dt)
    a = _ga(A, av, Bd.shape[0])
    mm = _ga(M, None, ncols)
    out = np.zeros((A.n, ncols), dtype=dt)
    if dt == np.float64:
        N.check(N.galib().ga_mxm_spd(C.byref(a), Bd.ctypes.data_as(N.F64P), ncols, C.byref(mm),
                                     out.ctypes.data_as(N.F64P)))
    elif dt == np.float32:
        N.check(N.galib().ga_mxm_spd_f32(C.byref(a), Bd.ctypes.data_as(N.F32P), ncols, C.byref(mm),
                                         out.ctypes.data_as(N.F32P)))
    else:
        raise TypeError("mxm_spd runs in float64 or float32")
    return out


def fuse_ewise_mult_add(A: np.ndarray, B: np.ndarray, Cm: np.ndarray) -> np.ndarray:
    """fuseEWiseMultAdd (include/matrix_la.h:192-198): returns C + A .* B for
    dense arrays of one shape and value type (C is not modified)."""
    N.init()
    dt = np.asarray(Cm).dtype
    a, b = (np.ascontiguousarray(x, dtype=dt) for x in (A, B))
    out = np.array(Cm, dtype=dt, order="C", copy=True)
    if not (a.shape == b.shape == out.shape):
        raise ValueError("shapes differ")
    N.check(N.galib().ga_dense_fuse_mult_add(out.size, _DT[dt], a.ctypes.data_as(C.c_void_p),
                                             b.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)))
    return out


def ewise_mult_dense_masked(A: np.ndarray, B: np.ndarray, M: CSR) -> np.ndarray:
    """The dense eWiseMult (include/matrix_la.h:200-219): C<M> = A .* B, zero
    outside M's pattern; A, B dense (M.n x ncols) of one value type."""
    N.init()
    dt = np.asarray(A).dtype
    a, b = (np.ascontiguousarray(x, dtype=dt) for x in (A, B))
    if a.shape != b.shape or a.ndim != 2 or a.shape[0] != M.n:
        raise ValueError("shapes differ")
    mm = _ga(M, None, a.shape[1])
    out = np.empty_like(a)
    N.check(N.galib().ga_dense_mult_masked(C.byref(mm), _DT[dt], a.ctypes.data_as(C.c_void_p),
                                           b.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)))
    return out


def triangle_count(A: CSR) -> int:
    """triangle_counting_masked_sandia (src/main.cpp:65-86): sum of C<L> = L*L."""
    N.init()
    t = C.c_int64()
    g = A.ga()
    N.check(N.galib().ga_triangle_count(C.byref(g), C.byref(t)))
    return t.value


triangle_counting_masked_sandia = triangle_count


def triangle_counting_masked_lu(A: CSR) -> int:
    """src/main.cpp:40-62: sum(C<A> = L*L^T) / 2."""
    N.init()
    t = C.c_int64()
    g = A.ga()
    N.check(N.galib().ga_triangle_count_lu(C.byref(g), C.byref(t)))
    return t.value


def triangle_counting_vertex(A: CSR) -> np.ndarray:
    """src/main.cpp:14-37: per-vertex triangles (row sums of (A*A)<A>) / 2."""
    N.init()
    out = np.zeros(A.n, dtype=np.int64)
    g = A.ga()
    N.check(N.galib().ga_triangle_count_vertex(C.byref(g), out.ctypes.data_as(N.I64P)))
    return out


def k_truss(A: CSR, k: int) -> tuple[CSR, int]:
    """k_truss (src/main.cpp:89-146): the k-truss pattern and the pass count."""
    N.init()
    rp = np.zeros(A.n + 1, dtype=np.int32)
    col = np.zeros(max(A.nnz, 1), dtype=np.int32)
    nnz = C.c_int64()
    it = C.c_int32()
    g = A.ga()
    N.check(N.galib().ga_ktruss(C.byref(g), k, rp.ctypes.data_as(N.I32P), col.ctypes.data_as(N.I32P),
                                C.byref(nnz), C.byref(it)))
    return CSR(A.n, rp, col[: nnz.value].copy()), it.value


def truss_decomposition(A: CSR) -> tuple[np.ndarray, int, dict]:
    """Truss number per stored entry of A (max k with the entry in the k-truss, 2 if none)."""
    N.init()
    truss = np.zeros(max(A.nnz, 1), dtype=np.int32)
    kmax = C.c_int32()
    st = N.GaTrussStats()
    g = A.ga()
    N.check(N.galib().ga_truss_decomposition(C.byref(g), truss.ctypes.data_as(N.I32P), C.byref(kmax), C.byref(st)))
    stats = {f: getattr(st, f) for f, _ in N.GaTrussStats._fields_}
    return truss[: A.nnz], kmax.value, stats


def betweenness_centrality_batch(A: CSR, sources) -> np.ndarray:
    """betweenness_centrality_batch (include/betweenness_centrality.h:125-237) with an
    explicit source list; FP64 sigma and delta; bc[v] = sum_s delta_s(v)."""
    N.init()
    src = np.ascontiguousarray(sources, dtype=np.int32)
    bc = np.zeros(A.n, dtype=np.float64)
    g = A.ga()
    N.check(N.galib().ga_bc_batch(C.byref(g), src.ctypes.data_as(N.I32P), len(src), bc.ctypes.data_as(N.F64P)))
    return bc


def _host_free(ptr: int) -> None:
    try:
        N.galib().ga_host_free(C.c_void_p(ptr))
    except Exception:
        pass


class _PinnedPool:
    """Page-locked result buffers (ga_host_alloc): a device-to-host copy into
    one runs at DMA speed with no page faults.  A buffer handed out as an
    array is reused once that array (and every view of it) is gone; a buffer
    whose array outlives the pool is freed with the array."""

    def __init__(self):
        self._bufs = []  # [address, bytes, weak reference to the array handed out or None]

    @staticmethod
    def _view(ptr: int, n: int, dtype) -> np.ndarray:
        raw = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
        return np.frombuffer(raw, dtype=dtype, count=n)

    def array(self, n: int, dtype) -> np.ndarray:
        nbytes = max(n, 1) * np.dtype(dtype).itemsize
        for b in self._bufs:
            if b[1] >= nbytes and (b[2] is None or b[2]() is None):
                arr = self._view(b[0], n, dtype)
                b[2] = weakref.ref(arr)
                return arr
        p = C.c_void_p()
        N.check(N.galib().ga_host_alloc(nbytes, C.byref(p)))
        arr = self._view(p.value, n, dtype)
        self._bufs.append([p.value, nbytes, weakref.ref(arr)])
        return arr

    def close(self) -> None:
        for ptr, _, ref in self._bufs:
            arr = ref() if ref is not None else None
            if arr is None:
                _host_free(ptr)
            else:
                weakref.finalize(arr, _host_free, ptr)
        self._bufs = []


class Graph:
    """A graph resident in HBM (ga_graph_*): upload once, query many times."""

    def __init__(self, A: CSR):
        N.init()
        self.A = A
        self._h = C.c_void_p()
        self._pinned = _PinnedPool()
        g = A.ga()
        N.check(N.galib().ga_graph_create(C.byref(g), C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            N.check(N.galib().ga_graph_destroy(self._h))
            self._h = C.c_void_p()
        if getattr(self, "_pinned", None) is not None:
            self._pinned.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def triangle_count(self) -> int:
        t = C.c_int64()
        N.check(N.galib().ga_graph_triangle_count(self._h, C.byref(t)))
        return t.value

    def truss(self, want_values: bool = True) -> tuple[np.ndarray | None, int, dict]:
        # the truss numbers land in page-locked memory (7 MB for R-MAT 16)
        truss = self._pinned.array(max(self.A.nnz, 1), np.int32) if want_values else None
        kmax = C.c_int32()
        st = N.GaTrussStats()
        N.check(N.galib().ga_graph_truss(self._h, truss.ctypes.data_as(N.I32P) if truss is not None else None,
                                         C.byref(kmax), C.byref(st)))
        stats = {f: getattr(st, f) for f, _ in N.GaTrussStats._fields_}
        return (truss[: self.A.nnz] if truss is not None else None), kmax.value, stats

    def bc(self, sources) -> np.ndarray:
        src = np.ascontiguousarray(sources, dtype=np.int32)
        bc = np.zeros(self.A.n, dtype=np.float64)
        N.check(N.galib().ga_graph_bc(self._h, src.ctypes.data_as(N.I32P), len(src), bc.ctypes.data_as(N.F64P)))
        return bc

    def ktruss(self, k: int) -> tuple[int, int]:
        """k-truss for one k: (stored entries of the k-truss, support passes)."""
        nnz = C.c_int64()
        passes = C.c_int32()
        N.check(N.galib().ga_graph_ktruss(self._h, int(k), C.byref(nnz), C.byref(passes)))
        return nnz.value, passes.value


class Batch:
    """Many small graphs resident together; triangle counts in one persistent launch."""

    def __init__(self, graphs: list[CSR]):
        N.init()
        self.graphs = graphs
        self._views = (N.GaCsr * max(len(graphs), 1))()
        for i, g in enumerate(graphs):
            self._views[i] = g.ga()
        self._h = C.c_void_p()
        N.check(N.galib().ga_batch_create(len(graphs), self._views, C.byref(self._h)))
        self.threads = 0

    def close(self) -> None:
        if self._h:
            N.check(N.galib().ga_batch_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def triangle_counts(self, threads: int = 0) -> np.ndarray:
        out = np.zeros(max(len(self.graphs), 1), dtype=np.int64)
        N.check(N.galib().ga_batch_triangle_count(self._h, threads, out.ctypes.data_as(N.I64P)))
        return out[: len(self.graphs)]

    TUNE_WIDTHS = (128, 256)
    TUNE_CLAIMS = (8, 16, 32, 64)
    TUNE_WAVE_FLOPS = (2048, 16384, 65536, 1 << 20)

    def autotune(self) -> tuple[tuple[int, int], dict]:
        """Times every (item width, rows per claim) pair on the resident batch,
        then the short-row product threshold at the best pair, and keeps the
        fastest; returns the (width, claim) kept and the time of every
        candidate in ms -- keys (width, claim) and ("wave_flops", threshold).
        `shape()` gives the threshold kept."""
        best_t, best_c = C.c_int(), C.c_int()
        ms = np.zeros(N.GA_BATCH_TUNE_CANDIDATES, dtype=np.float64)
        N.check(N.galib().ga_batch_autotune(self._h, C.byref(best_t), C.byref(best_c), ms.ctypes.data_as(N.F64P)))
        self.threads = best_t.value
        keys = [(t, c) for t in self.TUNE_WIDTHS for c in self.TUNE_CLAIMS]
        keys += [("wave_flops", f) for f in self.TUNE_WAVE_FLOPS]
        return (best_t.value, best_c.value), dict(zip(keys, ms.tolist()))

    def shape(self) -> dict:
        """The launch shape the batch uses: item width, rows per claim, short-row threshold."""
        t, c, f = C.c_int(), C.c_int(), C.c_int64()
        N.check(N.galib().ga_batch_shape(self._h, C.byref(t), C.byref(c), C.byref(f)))
        return {"item_width": t.value, "rows_per_claim": c.value, "wave_flops": f.value}


_KINDS = {"tc": N.GA_Q_TRIANGLES, "ktruss": N.GA_Q_KTRUSS, "bc": N.GA_Q_BC}


class _QueryArray:
    """ctypes array of ga_query built from dicts {"kind": "tc"|"ktruss"|"bc",
    "graph": id, "k": int, "sources": ints}; owns the source and output arrays."""

    def __init__(self, specs: list[dict], n_of_graph):
        self.specs = specs
        self.arr = (N.GaQuery * max(len(specs), 1))()
        self.src: list[np.ndarray | None] = []
        self.bc: list[np.ndarray | None] = []
        for i, q in enumerate(specs):
            kind = _KINDS[q["kind"]]
            self.arr[i].kind = kind
            self.arr[i].graph = int(q["graph"])
            self.arr[i].k = int(q.get("k", 0))
            src = bc = None
            if kind == N.GA_Q_BC:
                src = np.ascontiguousarray(q["sources"], dtype=np.int32)
                bc = np.zeros(max(n_of_graph(q["graph"]), 1), dtype=np.float64)
                self.arr[i].nsources = len(src)
                self.arr[i].sources = src.ctypes.data_as(N.I32P)
                self.arr[i].bc = bc.ctypes.data_as(N.F64P)
            self.src.append(src)
            self.bc.append(bc)

    def result(self, i: int, r: N.GaQueryResult) -> dict:
        d = {f: getattr(r, f) for f, _ in N.GaQueryResult._fields_}
        d["kind"] = self.specs[i]["kind"]
        d["graph"] = int(self.specs[i]["graph"])
        if self.bc[i] is not None:
            d["bc"] = self.bc[i]
        return d


class Server:
    """Request loop over graphs resident in HBM (ga_server_*): one device arena
    of hbm_budget bytes holds the graphs and every admitted query's workspace;
    `slots` workers run queries concurrently, each on a stream of its own;
    queued k-truss / BC queries on the same graph share a run (up to
    max_batch; 1 = no batching, 0 = the library default)."""

    def __init__(self, hbm_budget: int, slots: int = 4, max_batch: int = 0):
        N.init()
        cfg = N.GaServerConfig(int(hbm_budget), int(slots), int(max_batch))
        self._h = C.c_void_p()
        N.check(N.galib().ga_server_create(C.byref(cfg), C.byref(self._h)))
        self.graphs: list[CSR] = []
        self._pending: dict[int, tuple[_QueryArray, int]] = {}

    def close(self) -> None:
        if self._h:
            N.check(N.galib().ga_server_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_graph(self, A: CSR) -> int:
        gid = C.c_int32()
        g = A.ga()
        N.check(N.galib().ga_server_add_graph(self._h, C.byref(g), C.byref(gid)))
        self.graphs.append(A)
        return gid.value

    def _n(self, gid) -> int:
        return self.graphs[int(gid)].n if 0 <= int(gid) < len(self.graphs) else 1

    def footprint(self, spec: dict) -> int:
        qa = _QueryArray([spec], self._n)
        b = C.c_int64()
        N.check(N.galib().ga_server_query_footprint(self._h, C.byref(qa.arr[0]), C.byref(b)))
        return b.value

    def submit(self, spec: dict) -> int:
        qa = _QueryArray([spec], self._n)
        t = C.c_int64()
        N.check(N.galib().ga_server_submit(self._h, C.byref(qa.arr[0]), C.byref(t)))
        self._pending[t.value] = (qa, 0)
        return t.value

    def wait(self, ticket: int) -> dict:
        qa, i = self._pending.pop(ticket)
        r = N.GaQueryResult()
        N.check(N.galib().ga_server_wait(self._h, int(ticket), C.byref(r)))
        return qa.result(i, r)

    def replay(self, specs: list[dict], arrival_ms) -> list[dict]:
        """Submits query i at arrival_ms[i] after the call starts, returns every result."""
        qa = _QueryArray(specs, self._n)
        arr = np.ascontiguousarray(arrival_ms, dtype=np.float64)
        res = (N.GaQueryResult * max(len(specs), 1))()
        N.check(N.galib().ga_server_replay(self._h, len(specs), qa.arr, arr.ctypes.data_as(N.F64P), res))
        return [qa.result(i, res[i]) for i in range(len(specs))]

    def stats(self) -> dict:
        st = N.GaServerStats()
        N.check(N.galib().ga_server_stats_read(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in N.GaServerStats._fields_}


def plain_replay(graphs: list[Graph], specs: list[dict], arrival_ms) -> list[dict]:
    """The same stream on the one-graph-per-call path: one query at a time in
    arrival order on the library stream (ga_plain_replay)."""
    qa = _QueryArray(specs, lambda g: graphs[int(g)].A.n)
    hs = (C.c_void_p * max(len(graphs), 1))(*[g._h.value for g in graphs])
    arr = np.ascontiguousarray(arrival_ms, dtype=np.float64)
    res = (N.GaQueryResult * max(len(specs), 1))()
    N.check(N.galib().ga_plain_replay(len(graphs), hs, len(specs), qa.arr, arr.ctypes.data_as(N.F64P), res))
    return [qa.result(i, res[i]) for i in range(len(specs))]


PROF_EVENTS = 1  # HIP events around every masked-SpGEMM launch
PROF_BYTES = 2   # algorithmic wedge / byte accounting (extra kernels: not for timed runs)


def prof_enable(flags: int = PROF_EVENTS) -> None:
    N.init()
    N.check(N.galib().ga_prof_enable(int(flags)))


def prof_read(reset: bool = True) -> dict:
    p = N.GaKernelProf()
    N.check(N.galib().ga_prof_read(C.byref(p), 1 if reset else 0))
    return {f: getattr(p, f) for f, _ in N.GaKernelProf._fields_}


def device_info() -> dict:
    N.init()
    name = C.create_string_buffer(128)
    cu = C.c_int()
    hbm = C.c_int64()
    N.check(N.galib().ga_device_info(name, 128, C.byref(cu), C.byref(hbm)))
    return {"name": name.value.decode(), "compute_units": cu.value, "hbm_bytes": hbm.value}


# process-wide tuning options (include/graphalgo.h, GA_OPT_*): every value
# gives the same results, only the speed differs
OPTIONS = {"bc_bottom_up": 1, "bc_alpha": 2, "row_claim": 3, "sum_sub": 4, "item_wedges": 5,
           "selective_support": 6, "truss_loop": 7, "support_tri": 8, "serve_join": 9, "bc_loop": 10,
           "wave_flops": 11}


def set_option(name: str, value: float) -> None:
    N.init()
    N.check(N.galib().ga_set_option(OPTIONS[name], float(value)))


def get_option(name: str) -> float:
    N.init()
    v = C.c_double()
    N.check(N.galib().ga_get_option(OPTIONS[name], C.byref(v)))
    return v.value


class options:
    """Context manager that sets options and restores the previous values:
    ``with api.options(bc_bottom_up=0): ...``."""

    def __init__(self, **values):
        self.values = values
        self.saved = {}

    def __enter__(self):
        for k, v in self.values.items():
            self.saved[k] = get_option(k)
            set_option(k, v)
        return self

    def __exit__(self, *exc):
        for k, v in self.saved.items():
            set_option(k, v)
        return False
