"""ctypes bindings of the two native libraries.

``libgraphalgo.so`` is the gfx950 HIP library (C ABI in ``include/graphalgo.h``);
``libgagraph.so`` is the host graph substrate.  Both are built in-tree by
``graphalgo_amd.build``.  There is no fallback: if a library is missing the
import of the function that needs it raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"

GA_OK = 0
GA_ERRORS = {
    1: "GA_ERR_ARG", 2: "GA_ERR_DIM", 3: "GA_ERR_UNSORTED", 4: "GA_ERR_OOM",
    5: "GA_ERR_HIP", 6: "GA_ERR_LIMIT", 7: "GA_ERR_STATE",
}
GA_INT32, GA_INT64, GA_FLOAT64, GA_FLOAT32 = 0, 1, 2, 3

I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
F64P = C.POINTER(C.c_double)
F32P = C.POINTER(C.c_float)


class GaCsr(C.Structure):
    _fields_ = [
        ("nrows", C.c_int64), ("ncols", C.c_int64), ("nnz", C.c_int64),
        ("rowptr", I32P), ("col", I32P), ("val", C.c_void_p),
    ]


class GaTrussStats(C.Structure):
    _fields_ = [
        ("passes", C.c_int64), ("steps", C.c_int64), ("wedges", C.c_int64),
        ("bytes_alg", C.c_int64), ("support_ms", C.c_double), ("total_ms", C.c_double),
    ]


class GaKernelProf(C.Structure):
    _fields_ = [
        ("launches", C.c_int64), ("kernel_ms", C.c_double),
        ("wedges", C.c_int64), ("bytes_alg", C.c_int64), ("wedges_processed", C.c_int64),
        ("accounted_launches", C.c_int64),
    ]


class GaServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("slots", C.c_int32), ("max_batch", C.c_int32)]


GA_Q_TRIANGLES, GA_Q_KTRUSS, GA_Q_BC = 0, 1, 2
GA_EW_ADD, GA_EW_MULT, GA_EW_ADD_DISJOINT = 0, 1, 2
GA_BATCH_TUNE_CANDIDATES = 12  # include/graphalgo.h


class GaQuery(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("graph", C.c_int32), ("k", C.c_int32), ("nsources", C.c_int32),
        ("sources", I32P), ("bc", F64P),
    ]


class GaQueryResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("passes", C.c_int32), ("value", C.c_int64), ("footprint", C.c_int64),
        ("t_submit_ms", C.c_double), ("t_start_ms", C.c_double), ("t_end_ms", C.c_double),
        ("slot", C.c_int32), ("batched", C.c_int32),
    ]


class GaServerStats(C.Structure):
    _fields_ = [
        ("arena_bytes", C.c_int64), ("resident_bytes", C.c_int64), ("peak_bytes", C.c_int64),
        ("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
        ("batches", C.c_int64), ("peak_inflight", C.c_int32), ("slots", C.c_int32),
        ("joined", C.c_int64),
    ]


class GraphAlgoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{GA_ERRORS.get(code, code)}: {msg}")
        self.code = code


# name -> (restype, argtypes); every symbol declared in include/graphalgo.h
GA_SIGNATURES = {
    "ga_last_error": (C.c_char_p, []),
    "ga_init": (C.c_int, [C.c_int]),
    "ga_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), I64P]),
    "ga_synchronize": (C.c_int, []),
    "ga_check_csr": (C.c_int, [C.POINTER(GaCsr)]),
    "ga_mxm_masked": (C.c_int, [C.POINTER(GaCsr), C.POINTER(GaCsr), C.POINTER(GaCsr), C.c_int, C.c_void_p]),
    "ga_mxm_cmask": (C.c_int, [C.POINTER(GaCsr), C.POINTER(GaCsr), C.POINTER(GaCsr), C.c_int,
                               C.POINTER(I32P), C.POINTER(I32P), C.POINTER(C.c_void_p), I64P]),
    "ga_mxm_spd": (C.c_int, [C.POINTER(GaCsr), F64P, C.c_int64, C.POINTER(GaCsr), F64P]),
    "ga_mxm_spd_f32": (C.c_int, [C.POINTER(GaCsr), F32P, C.c_int64, C.POINTER(GaCsr), F32P]),
    "ga_dense_fuse_mult_add": (C.c_int, [C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ga_dense_mult_masked": (C.c_int, [C.POINTER(GaCsr), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ga_ewise": (C.c_int, [C.c_int, C.POINTER(GaCsr), C.POINTER(GaCsr), C.POINTER(GaCsr), C.c_int,
                           C.POINTER(I32P), C.POINTER(I32P), C.POINTER(C.c_void_p), I64P]),
    "ga_scale": (C.c_int, [C.POINTER(GaCsr), C.c_int, C.c_double, C.c_void_p]),
    "ga_free_host": (None, [C.c_void_p]),
    "ga_host_alloc": (C.c_int, [C.c_int64, C.POINTER(C.c_void_p)]),
    "ga_host_free": (C.c_int, [C.c_void_p]),
    "ga_triangle_count": (C.c_int, [C.POINTER(GaCsr), I64P]),
    "ga_triangle_count_lu": (C.c_int, [C.POINTER(GaCsr), I64P]),
    "ga_triangle_count_vertex": (C.c_int, [C.POINTER(GaCsr), I64P]),
    "ga_triangle_count_noalloc": (C.c_int, [C.POINTER(GaCsr), C.c_int64, C.c_void_p, C.c_void_p]),
    "ga_ktruss": (C.c_int, [C.POINTER(GaCsr), C.c_int, I32P, I32P, I64P, I32P]),
    "ga_truss_decomposition": (C.c_int, [C.POINTER(GaCsr), I32P, I32P, C.POINTER(GaTrussStats)]),
    "ga_bc_batch": (C.c_int, [C.POINTER(GaCsr), I32P, C.c_int64, F64P]),
    "ga_graph_create": (C.c_int, [C.POINTER(GaCsr), C.POINTER(C.c_void_p)]),
    "ga_graph_destroy": (C.c_int, [C.c_void_p]),
    "ga_graph_triangle_count": (C.c_int, [C.c_void_p, I64P]),
    "ga_graph_truss": (C.c_int, [C.c_void_p, I32P, I32P, C.POINTER(GaTrussStats)]),
    "ga_graph_bc": (C.c_int, [C.c_void_p, I32P, C.c_int64, F64P]),
    "ga_graph_ktruss": (C.c_int, [C.c_void_p, C.c_int, I64P, I32P]),
    "ga_server_create": (C.c_int, [C.POINTER(GaServerConfig), C.POINTER(C.c_void_p)]),
    "ga_server_destroy": (C.c_int, [C.c_void_p]),
    "ga_server_add_graph": (C.c_int, [C.c_void_p, C.POINTER(GaCsr), I32P]),
    "ga_server_query_footprint": (C.c_int, [C.c_void_p, C.POINTER(GaQuery), I64P]),
    "ga_server_submit": (C.c_int, [C.c_void_p, C.POINTER(GaQuery), I64P]),
    "ga_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(GaQueryResult)]),
    "ga_server_stats_read": (C.c_int, [C.c_void_p, C.POINTER(GaServerStats)]),
    "ga_server_replay": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(GaQuery), F64P, C.POINTER(GaQueryResult)]),
    "ga_plain_replay": (C.c_int, [C.c_int64, C.POINTER(C.c_void_p), C.c_int64, C.POINTER(GaQuery), F64P,
                                  C.POINTER(GaQueryResult)]),
    "ga_batch_create": (C.c_int, [C.c_int64, C.POINTER(GaCsr), C.POINTER(C.c_void_p)]),
    "ga_batch_destroy": (C.c_int, [C.c_void_p]),
    "ga_batch_triangle_count": (C.c_int, [C.c_void_p, C.c_int, I64P]),
    "ga_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), F64P]),
    "ga_batch_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "ga_set_option": (C.c_int, [C.c_int, C.c_double]),
    "ga_get_option": (C.c_int, [C.c_int, C.POINTER(C.c_double)]),
    "ga_prof_enable": (C.c_int, [C.c_int]),
    "ga_prof_read": (C.c_int, [C.POINTER(GaKernelProf), C.c_int]),
    "ga_stream_copy": (C.c_int, [C.c_int64, C.c_int, F64P]),
}

GAH_SIGNATURES = {
    "gah_last_error": (C.c_char_p, []),
    "gah_free": (None, [C.c_void_p]),
    "gah_rng_bits": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
    "gah_gen_er": (C.c_int, [C.c_int64, C.c_int64, C.c_uint64, C.POINTER(I32P), C.POINTER(I32P), I64P]),
    "gah_gen_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64,
                               C.POINTER(I32P), C.POINTER(I32P), I64P, I64P]),
    "gah_knn_params": (None, [C.c_uint64, C.c_int64, I64P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gah_gen_knn": (C.c_int, [C.c_uint64, C.c_int64, C.c_int, C.c_int, C.c_int,
                              C.POINTER(I32P), C.POINTER(I32P), I64P, F64P]),
    "gah_read_mtx": (C.c_int, [C.c_char_p, I64P, C.POINTER(I32P), C.POINTER(I32P), I64P, C.POINTER(C.c_int)]),
    "gah_symmetrize": (C.c_int, [C.c_int64, I32P, I32P, C.POINTER(I32P), C.POINTER(I32P), I64P]),
}

_libs: dict[str, C.CDLL] = {}


def _load(name: str, sigs: dict) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = LIB_DIR / name
    if not path.exists():
        raise ImportError(f"{path} is not built; run `python -m graphalgo_amd.build`")
    lib = C.CDLL(str(path), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    for fn, (res, args) in sigs.items():
        f = getattr(lib, fn, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _libs[name] = lib
    return lib


def galib() -> C.CDLL:
    """The gfx950 HIP library."""
    return _load("libgraphalgo.so", GA_SIGNATURES)


def hostlib() -> C.CDLL:
    """The host graph library."""
    return _load("libgagraph.so", GAH_SIGNATURES)


def check(status: int) -> None:
    if status != GA_OK:
        msg = galib().ga_last_error()
        raise GraphAlgoError(status, msg.decode() if msg else "")


_initialised_device: int | None = None


def init(device: int = 0) -> None:
    """Selects the GPU and creates the library stream (idempotent)."""
    global _initialised_device
    if _initialised_device == device:
        return
    check(galib().ga_init(device))
    _initialised_device = device
