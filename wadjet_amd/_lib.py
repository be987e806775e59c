"""ctypes binding of the C-ABI in include/wadjet_gpu.h and include/wadjet_gen.h.

The library is loaded from wadjet_amd/lib/libwadjet_gpu.so (built in-tree by
wadjet_amd/build.py). There is no fallback: if the library is missing or does
not load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libwadjet_gpu.so"

# --- enums (mirror the header) ---------------------------------------------
WG_OK, WG_ERR_INVALID, WG_ERR_UNSUPPORTED, WG_ERR_DEVICE, WG_ERR_OOM, WG_ERR_STATE, WG_ERR_NULL_KEY = range(7)
INT32, INT64, FLOAT64, DECIMAL, DECIMAL128, IPV4, UINT32, UINT16 = range(1, 9)
SUM, COUNT, MIN, MAX, AVG = range(5)
EQ, NE, LT, LE, GT, GE = range(6)
MEM_HOST, MEM_DEVICE = 0, 1

GEN_UNIFORM_I64, GEN_UNIFORM_I32, GEN_UNIFORM_U16, GEN_ZIPF_I32, GEN_ZIPF_U32 = 1, 2, 3, 4, 5
GEN_NORMAL_F64, GEN_LOGNORMAL_I64, GEN_PERM_KEY_I64, GEN_FK_I64, GEN_SEQ_I64, GEN_RANK_U32 = 6, 7, 8, 9, 10, 11

TYPE_BYTES = {INT32: 4, INT64: 8, FLOAT64: 8, DECIMAL: 8, DECIMAL128: 16, IPV4: 8, UINT32: 4, UINT16: 2}


class Column(C.Structure):
    _fields_ = [("data", C.c_void_p), ("validity", C.c_void_p), ("type", C.c_int32), ("scale", C.c_int32)]


class AggSpecC(C.Structure):
    _fields_ = [("func", C.c_int32), ("input_col", C.c_int32), ("input_col2", C.c_int32), ("reserved", C.c_int32)]


class PredicateC(C.Structure):
    _fields_ = [("col", C.c_int32), ("op", C.c_int32), ("i64", C.c_int64), ("f64", C.c_double)]


class ProbeCursor(C.Structure):
    _fields_ = [("pos", C.c_uint64), ("chain", C.c_int64), ("done", C.c_int32), ("reserved", C.c_int32)]


class PqRuns(C.Structure):
    _fields_ = [("runs", C.c_void_p), ("n_runs", C.c_uint32), ("bit_width", C.c_int32), ("data", C.c_void_p)]


class PqPage(C.Structure):
    _fields_ = [("num_rows", C.c_uint64), ("physical_type", C.c_int32), ("encoding", C.c_int32), ("def_", PqRuns),
                ("plain", C.c_void_p), ("idx", PqRuns), ("dict", C.c_void_p), ("dict_n", C.c_uint64)]


class PqChunk(C.Structure):
    _fields_ = [("offset", C.c_uint64), ("bytes", C.c_uint64), ("num_rows", C.c_uint64), ("codec", C.c_int32),
                ("reserved", C.c_int32)]


class BatchQuery(C.Structure):
    _fields_ = [("key", Column), ("val", Column), ("n_rows", C.c_uint64), ("est_groups", C.c_uint64),
                ("row_begin", C.c_uint64), ("range_col", Column), ("range_lo", C.c_int64), ("range_hi", C.c_int64),
                ("dim", C.c_void_p), ("dim_payload", C.c_int32), ("reserved", C.c_int32)]


class ServeQuery(C.Structure):
    _fields_ = [("kind", C.c_int32), ("key_col", C.c_int32), ("val_col", C.c_int32), ("dim", C.c_int32),
                ("dim_payload", C.c_int32), ("reserved", C.c_int32), ("ts_lo", C.c_int64), ("ts_hi", C.c_int64),
                ("est_groups", C.c_uint64)]


class ServeStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("batches", C.c_uint64), ("max_in_flight", C.c_uint64),
                ("peak_footprint", C.c_uint64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double), ("wall_s", C.c_double), ("qps", C.c_double),
                ("device_busy_s", C.c_double), ("lane_busy_s", C.c_double), ("host_submit_s", C.c_double),
                ("host_retire_s", C.c_double), ("host_deliver_s", C.c_double)]


P = C.c_void_p
U64 = C.c_uint64
I64 = C.c_int64
I32 = C.c_int32
PP = C.POINTER(C.c_void_p)

# name -> argtypes; every function returns wg_status (int)
SIGNATURES: dict[str, list] = {
    "wg_ctx_create": [C.c_int, C.POINTER(P)],
    "wg_ctx_destroy": [P],
    "wg_ctx_sync": [P],
    "wg_ctx_info": [P, C.POINTER(I32), C.POINTER(U64), C.POINTER(I32)],
    "wg_ctx_timer_start": [P],
    "wg_ctx_timer_stop": [P, C.POINTER(C.c_float)],
    "wg_malloc": [P, U64, C.POINTER(P)],
    "wg_free": [P, P],
    "wg_memcpy_h2d": [P, P, P, U64],
    "wg_memcpy_d2h": [P, P, P, U64],
    "wg_memset": [P, P, C.c_int, U64],
    "wg_memcpy_d2d": [P, P, P, U64],
    "wg_agg_create": [P, I32, C.POINTER(I32), C.POINTER(I32), C.POINTER(C.c_uint8), I32, C.POINTER(I32), I32,
                      C.POINTER(AggSpecC), U64, C.POINTER(P)],
    "wg_agg_destroy": [P],
    "wg_agg_consume": [P, C.POINTER(Column), U64, P, U64, C.POINTER(PredicateC), I32, I32],
    "wg_agg_merge": [P, P],
    "wg_agg_finalize": [P],
    "wg_agg_num_groups": [P, C.POINTER(U64)],
    "wg_agg_fetch": [P, U64, U64, PP, PP, PP, I32],
    "wg_agg_fetch_state": [P, I32, U64, U64, P, P, P, P, I32],
    "wg_agg_fetch_key_valid": [P, U64, U64, P, I32],
    "wg_agg_reset": [P],
    "wg_agg_footprint": [P, C.POINTER(U64)],
    "wg_agg_set_tuning": [P, I32, I32],
    "wg_agg_stats": [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(U64), C.POINTER(U64)],
    "wg_agg_kernel_time": [P, C.POINTER(C.c_float), C.POINTER(C.c_uint32)],
    "wg_agg_pass_times": [P, C.POINTER(C.c_float)],
    "wg_agg_set_strategy": [P, I32, I32, I32, U64],
    "wg_agg_set_heavy_hitters": [P, I32],
    "wg_agg_spill_run": [P, P, P, P, U64, C.POINTER(U64)],
    "wg_agg_state_bytes": [P, C.POINTER(U64)],
    "wg_agg_merge_run": [P, P, U64],
    "wg_filter": [P, C.POINTER(Column), U64, P, U64, I32, I64, C.c_double, P, C.POINTER(U64)],
    "wg_filter_noalloc": [P, C.POINTER(Column), U64, P, U64, I32, I64, C.c_double, P, C.POINTER(U64)],
    "wg_join_create": [P, I32, C.POINTER(I32), C.POINTER(I32), I32, I32, C.POINTER(I32), U64, C.POINTER(P)],
    "wg_join_destroy": [P],
    "wg_join_create_dual": [P, I32, C.POINTER(I32), C.POINTER(I32), I32, I32, I32, C.POINTER(I32), U64, C.POINTER(P)],
    "wg_join_dual_probe_key": [P, C.POINTER(Column), C.POINTER(Column), U64, P, U64, P, P],
    "wg_join_build_append": [P, C.POINTER(Column), U64, I32],
    "wg_join_build_finish": [P],
    "wg_join_build_rows": [P, C.POINTER(U64)],
    "wg_join_semi": [P, C.POINTER(Column), U64, P, U64, I32, P, C.POINTER(U64)],
    "wg_join_probe": [P, C.POINTER(Column), U64, P, U64, P, P, U64, C.POINTER(ProbeCursor), C.POINTER(U64)],
    "wg_join_gather_build": [P, I32, P, U64, P],
    "wg_gather": [P, C.POINTER(Column), P, U64, P],
    "wg_join_probe_agg": [P, C.POINTER(Column), I32, I32, U64, P, U64, P],
    "wg_join_set_tuning": [P, I32, I32, I32],
    "wg_join_bloom_build": [P],
    "wg_join_bloom_filter": [P, C.POINTER(Column), U64, P, U64, P, C.POINTER(U64)],
    "wg_join_bloom_words": [P, C.POINTER(U64), P, U64],
    "wg_join_digest": [P, C.POINTER(Column), P, P, U64, C.POINTER(I32), I32, C.POINTER(I32), I32, C.POINTER(U64)],
    "wg_pq_scan_runs": [P, U64, I32, U64, P, U64, C.POINTER(U64), C.POINTER(U64), C.POINTER(U64)],
    "wg_parquet_decode_page": [P, P, P, P, U64],
    "wg_parquet_decode_pages": [P, P, U64, P, P, U64],
    "wg_ctx_inject_alloc_failure": [P, I64, I32],
    "wg_pq_stage": [P, C.c_char_p, P, U64, I32, I32, I32, C.POINTER(C.c_void_p)],
    "wg_pq_staged_info": [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(U64), C.POINTER(U64)],
    "wg_pq_staged_pages": [P, P, U64],
    "wg_pq_staged_free": [P],
    "wg_pq_decompress": [I32, P, U64, P, U64],
    "wg_view_flatten": [P, P, U64, P, P, U64, P, P],
    "wg_join_build_column": [P, I32, P],
    "wg_shuffle_partition": [P, C.POINTER(Column), I32, C.POINTER(I32), I32, U64, P, U64, I32, P, P, P],
    "wg_batch_create": [P, I32, U64, C.POINTER(P)],
    "wg_batch_destroy": [P],
    "wg_batch_run": [P, C.POINTER(BatchQuery), I32, I32, I32],
    "wg_batch_autotune": [P, C.POINTER(BatchQuery), I32],
    "wg_batch_result": [P, I32, C.POINTER(U64), P, P, P, P, U64],
    "wg_batch_choice": [P, I32, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)],
    "wg_batch_tuning": [P, C.POINTER(I32), C.POINTER(I32), C.POINTER(U64), C.POINTER(I32), C.POINTER(I32)],
    "wg_batch_set_tuning": [P, I32, I32, U64, I32, I32],
    "wg_serve_create": [P, C.POINTER(Column), I32, U64, I32, I32, U64, I32, I32, C.POINTER(P)],
    "wg_serve_destroy": [P],
    "wg_serve_add_dim": [P, P, C.POINTER(I32)],
    "wg_serve_submit": [P, C.POINTER(ServeQuery), C.POINTER(U64)],
    "wg_serve_wait": [P, U64, C.POINTER(U64), C.POINTER(C.c_double)],
    "wg_serve_result": [P, U64, P, P, P, P, U64],
    "wg_serve_replay": [P, C.POINTER(ServeQuery), I32, C.c_double, U64, P, P],
    "wg_serve_stats_get": [P, C.POINTER(ServeStats)],
    "wg_serve_reset": [P],
    "wg_gen": [P, I32, P, U64, U64, U64, I64, I64, C.c_double, C.c_double],
}
STRING_FNS = ("wg_last_error", "wg_version")


class WgError(RuntimeError):
    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} -> status {status}: {msg}")
        self.status = status
        self.fn = fn


class NullKeyError(WgError):
    pass


class UnsupportedError(WgError):
    pass


_lib = None


def lib():
    """Load the library once; raise loudly if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `python -m wadjet_amd.build` (no CPU fallback exists)")
    L = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = C.c_int
    for name in STRING_FNS:
        fn = getattr(L, name)
        fn.argtypes = []
        fn.restype = C.c_char_p
    _lib = L
    return L


def call(name: str, *args) -> None:
    L = lib()
    st = getattr(L, name)(*args)
    if st != WG_OK:
        msg = (L.wg_last_error() or b"").decode()
        cls = NullKeyError if st == WG_ERR_NULL_KEY else UnsupportedError if st == WG_ERR_UNSUPPORTED else WgError
        raise cls(name, st, msg)


def exported_symbols() -> list[str]:
    return sorted(list(SIGNATURES) + list(STRING_FNS))
