"""ctypes binding of ``libcumetis_amd.so`` (the C ABI in include/cumetis_amd.h).

The library is the product: nothing here computes on the CPU.  Loading fails
loudly when the shared library has not been built.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libcumetis_amd.so"

CM_OK = 0
CM_ERR_INVALID = -1
CM_ERR_HIP = -2
CM_ERR_NOMEM = -3
CM_ERR_UNSUPPORTED = -4
CM_ERR_INPUT = -5
CM_OFF32, CM_OFF64 = 0, 1
CM_ROW_FIRST_SEEN, CM_ROW_ASCENDING = 0, 1
CM_FLAG_TRUSTED_INPUT = 1
CM_FLAG_NO_BALANCE = 2
CM_FLAG_EXACT_INITPART = 4
CM_FLAG_HOST_INIT_COARSEN = 8
CM_DEFAULT_TRIALS = 4


class GraphView(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("offset_type", C.c_int32),
                ("on_device", C.c_int32), ("xadj", C.c_void_p), ("adjncy", C.c_void_p),
                ("adjwgt", C.c_void_p), ("vwgt", C.c_void_p)]


class CoarsenParams(C.Structure):
    _fields_ = [("nparts", C.c_int32), ("threshold", C.c_int32), ("maxvwgt", C.c_int32),
                ("row_order", C.c_int32), ("max_levels", C.c_int32), ("max_rounds", C.c_int32),
                ("seed", C.c_uint64), ("flags", C.c_int32), ("trials", C.c_int32)]


class LevelInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("offset_type", C.c_int32),
                ("rounds", C.c_int32), ("xadj", C.c_void_p), ("adjncy", C.c_void_p),
                ("adjwgt", C.c_void_p), ("vwgt", C.c_void_p), ("match", C.c_void_p),
                ("cmap", C.c_void_p), ("tvwgt", C.c_int64), ("vertex_base", C.c_int64),
                ("edge_base", C.c_int64), ("coarse_base", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("first_round_batch", C.c_int32), ("level_batch_margin", C.c_int32),
                ("large_level_batch", C.c_int32), ("compact_round", C.c_int32),
                ("compact_min_n", C.c_int64), ("speculate_edges", C.c_int64),
                ("number_one_tile_max", C.c_int32), ("seg_wave_bins", C.c_int32),
                ("graph_rounds", C.c_int32), ("reserved", C.c_int32 * 7)]


class PartReport(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("coarsen_ms", C.c_double), ("init_ms", C.c_double),
                ("uncoarsen_ms", C.c_double), ("levels", C.c_int32), ("coarsest_n", C.c_int32),
                ("edgecut", C.c_int64), ("trials", C.c_int32), ("best_trial", C.c_int32),
                ("init_device_steps", C.c_int32), ("reserved", C.c_int32)]


# name -> (restype, argtypes); every symbol declared in include/cumetis_amd.h
_P = C.c_void_p
_PROTOS = {
    "cm_ctx_create": (C.c_int, [C.c_int, _P, C.POINTER(_P)]),
    "cm_ctx_destroy": (C.c_int, [_P]),
    "cm_ctx_last_error": (C.c_char_p, [_P]),
    "cm_ctx_sync": (C.c_int, [_P]),
    "cm_ctx_set_timing": (C.c_int, [_P, C.c_int]),
    "cm_ctx_reset_stats": (C.c_int, [_P]),
    "cm_ctx_kernel_stat": (C.c_int, [_P, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int64),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "cm_ctx_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "cm_ctx_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "cm_ctx_pool_bytes": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "cm_dmalloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "cm_dfree": (C.c_int, [_P, _P]),
    "cm_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "cm_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "cm_memcpy_d2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "cm_coarsen": (C.c_int, [_P, C.POINTER(GraphView), C.POINTER(CoarsenParams), C.POINTER(_P)]),
    "cm_coarsen_batch": (C.c_int, [_P, C.c_int32, C.POINTER(GraphView), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64), C.POINTER(CoarsenParams), C.POINTER(_P)]),
    "cm_hier_num_levels": (C.c_int, [_P]),
    "cm_hier_level": (C.c_int, [_P, C.c_int, C.POINTER(LevelInfo)]),
    "cm_hier_copy_level": (C.c_int, [_P, _P, C.c_int, _P, _P, _P, _P, _P, _P]),
    "cm_project": (C.c_int, [_P, _P, C.c_int, _P, _P]),
    "cm_hier_release": (C.c_int, [_P, _P]),
    "cm_match": (C.c_int, [_P, C.POINTER(GraphView), C.c_int32, C.c_uint64, C.c_int32, _P,
                           C.POINTER(C.c_int32)]),
    "cm_number": (C.c_int, [_P, C.c_int64, _P, C.c_int32, _P, _P, C.POINTER(C.c_int64)]),
    "cm_contract": (C.c_int, [_P, C.POINTER(GraphView), _P, _P, C.c_int64, C.c_int32, _P, _P, _P,
                              _P, C.POINTER(C.c_int64)]),
    "cm_segsort_i32": (C.c_int, [_P, _P, _P, C.c_int32, _P, C.c_int32]),
    "cm_segsort_i32_noalloc": (C.c_int, [_P, _P, _P, C.c_int32, _P, C.c_int32]),
    "cm_project_raw": (C.c_int, [_P, C.c_int64, _P, _P, _P]),
    "cuMetis_PartGraph": (C.c_int, [C.POINTER(C.c_int)] * 5 + [C.POINTER(C.c_int),
                                    C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cm_partition": (C.c_int, [_P, C.POINTER(GraphView), C.POINTER(CoarsenParams), _P,
                               C.POINTER(PartReport)]),
    "cm_server_create": (C.c_int, [C.c_int, C.c_int64, C.c_int32, C.POINTER(_P)]),
    "cm_server_submit": (C.c_int, [_P, C.POINTER(GraphView), C.POINTER(CoarsenParams),
                                   C.POINTER(C.c_uint64)]),
    "cm_server_wait": (C.c_int, [_P, C.c_uint64, _P, C.POINTER(PartReport)]),
    "cm_server_stats": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_int32)]),
    "cm_server_destroy": (C.c_int, [_P]),
    "cm_read_metis": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                                C.c_char_p, C.c_int]),
    "cm_free_host": (None, [_P]),
    "cm_write_partition": (C.c_int, [C.c_char_p, _P, C.c_int64]),
    "cm_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                              C.POINTER(C.c_int64), C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "cm_gen_rgg2d": (C.c_int, [C.c_int64, C.c_double, C.c_int32, C.c_uint64, C.POINTER(_P),
                               C.POINTER(_P), C.POINTER(_P)]),
    "cm_initial_partition": (C.c_int, [C.c_int64, _P, _P, _P, _P, C.c_int32, _P,
                                       C.POINTER(C.c_int64)]),
    "cm_initial_partition_par": (C.c_int, [C.c_int64, _P, _P, _P, _P, C.c_int32, _P,
                                           C.POINTER(C.c_int64)]),
}

_lib = None


class EngineError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load(path: Path | None = None) -> C.CDLL:
    """Load the engine library; raises FileNotFoundError when it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is missing: build the HIP engine first (python -m cumetis_amd.build)")
    lib = C.CDLL(str(p))
    missing = []
    for name, (res, args) in _PROTOS.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    lib.cm_missing_symbols = missing
    if path is None:
        _lib = lib
    return lib


def declared_symbols() -> list[str]:
    return list(_PROTOS)
