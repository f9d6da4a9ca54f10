"""The allocation-free triangle count (ga_triangle_count_noalloc) through the
C ABI: device-resident operands (torch device tensors as plain device
memory), work buffers from the library's fixed scratch array.  Its result
must equal ga_triangle_count's on the same graph, call after call, with a
smaller graph second (the array keeps the previous call's contents)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GA_ERR_LIMIT = 6


def _noalloc(lib, N, torch, L, stream=None, nrows=None):
    rp = torch.from_numpy(np.ascontiguousarray(L.rowptr, dtype=np.int32)).cuda()
    col = torch.from_numpy(np.ascontiguousarray(L.col if L.nnz else np.zeros(1), dtype=np.int32)).cuda()
    out = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    n = L.n if nrows is None else nrows
    view = N.GaCsr(n, n, L.nnz, C.cast(rp.data_ptr(), N.I32P), C.cast(col.data_ptr(), N.I32P), None)
    max_row = int(np.diff(L.rowptr).max()) if L.n else 0
    torch.cuda.synchronize()
    st = lib.ga_triangle_count_noalloc(C.byref(view), max_row, C.c_void_p(out.data_ptr()),
                                       C.c_void_p(stream) if stream else None)
    if st != 0:
        return st, None
    if stream:
        torch.cuda.synchronize()
    else:
        N.check(lib.ga_synchronize())
    return 0, int(out.item())


def test_noalloc_equals_triangle_count(gpu, G, small_graphs):
    import torch
    from graphalgo_amd import _native as N
    lib = N.galib()
    cases = [("rmat16", G.rmat(16, 16, seed=1)), ("rmat12", G.rmat(12, 16, seed=2)),
             ("er_1000_9990", G.erdos_renyi(1000, 9990, seed=1)), ("G43", small_graphs["G43.mtx"][1]),
             ("empty", G.CSR.from_edges(5, [], []))]
    stream = torch.cuda.current_stream().cuda_stream
    for name, g in cases:  # each graph smaller than the one before it
        print(f"case noalloc tc {name}: n={g.n} nnz={g.nnz}", flush=True)
        want = gpu.triangle_count(g)
        L = g.lower()
        for s in (stream, None):  # the caller's stream, then the library's
            st, got = _noalloc(lib, N, torch, L, stream=s)
            assert st == 0, N.galib().ga_last_error()
            assert got == want, (name, got, want)


def test_noalloc_refuses_graphs_beyond_the_scratch(gpu, G):
    """The size check runs on the host before anything is launched: a row
    count whose work buffers exceed the array is refused."""
    import torch
    from graphalgo_amd import _native as N
    lib = N.galib()
    L = G.rmat(8, 4, seed=1).lower()
    st, _ = _noalloc(lib, N, torch, L, nrows=1 << 23)
    assert st == GA_ERR_LIMIT
    assert b"scratch" in lib.ga_last_error()
    # and the entry point still works afterwards
    st, got = _noalloc(lib, N, torch, L)
    assert st == 0 and got == gpu.triangle_count(G.rmat(8, 4, seed=1))
