"""ctypes binding of libgpma.so (the C ABI declared in include/gpma.h).

The library is the only compute path: if it is missing or fails to load,
importing this module raises, and nothing falls back to a CPU path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# GPMA_LIB selects another build of the library (e.g. a variant under test)
LIB_PATH = os.environ.get("GPMA_LIB") or os.path.join(HERE, "libgpma.so")

GPMA_OK = 0
GPMA_ERR_ARG = 1
GPMA_ERR_HIP = 2
GPMA_ERR_OOM = 3
GPMA_ERR_CAPACITY = 4
GPMA_ERR_KEY_RANGE = 5
GPMA_ERR_NO_DEVICE = 6
GPMA_ERR_TIMEOUT = 7
GPMA_ERR_STOPPED = 8
GPMA_SERVE_STRICT = 0
GPMA_SERVE_SNAPSHOT = 1

ERROR_NAMES = {
    GPMA_ERR_ARG: "GPMA_ERR_ARG",
    GPMA_ERR_HIP: "GPMA_ERR_HIP",
    GPMA_ERR_OOM: "GPMA_ERR_OOM",
    GPMA_ERR_CAPACITY: "GPMA_ERR_CAPACITY",
    GPMA_ERR_KEY_RANGE: "GPMA_ERR_KEY_RANGE",
    GPMA_ERR_NO_DEVICE: "GPMA_ERR_NO_DEVICE",
    GPMA_ERR_TIMEOUT: "GPMA_ERR_TIMEOUT",
    GPMA_ERR_STOPPED: "GPMA_ERR_STOPPED",
}

KEY_NONE = 0xFFFFFFFFFFFFFFFF
KEY_MAX = 0xFFFFFFFFFFFFFFFE
COL_IDX_NONE = 0xFFFFFFFF


class GpmaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


class GpmaConfig(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("bulk_threshold", ctypes.c_uint32),
        ("pull_permille", ctypes.c_uint32),
        ("segment_length", ctypes.c_uint32),
        ("device_levels", ctypes.c_uint32),
        ("sort_bucket_bits", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("density_permille", ctypes.c_uint16 * 4),
        ("sort_items_per_lane", ctypes.c_uint32),
        ("push_tile", ctypes.c_uint32),
        ("ms_tile", ctypes.c_uint32),
        ("reserved2", ctypes.c_uint32),
    ]


class GpmaCsr(ctypes.Structure):
    _fields_ = [
        ("keys", ctypes.c_void_p),
        ("values", ctypes.c_void_p),
        ("row_offset", ctypes.c_void_p),
        ("capacity", ctypes.c_uint64),
        ("segment_length", ctypes.c_uint32),
        ("tree_height", ctypes.c_uint32),
        ("row_num", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class GpmaStats(ctypes.Structure):
    _fields_ = [
        ("capacity", ctypes.c_uint64),
        ("segment_length", ctypes.c_uint32),
        ("tree_height", ctypes.c_uint32),
        ("row_num", ctypes.c_uint32),
        ("last_levels", ctypes.c_uint32),
        ("updates_applied", ctypes.c_uint64),
        ("invalid_skipped", ctypes.c_uint64),
        ("dropped", ctypes.c_uint64),
        ("resizes", ctypes.c_uint64),
        ("bulk_loads", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
    ]


class GpmaStepStats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_uint64),
        ("replays", ctypes.c_uint64),
        ("captures", ctypes.c_uint64),
        ("host_steps", ctypes.c_uint64),
        ("update_tails", ctypes.c_uint64),
        ("bfs_tails", ctypes.c_uint64),
        ("graphs", ctypes.c_uint32),
        ("bfs_levels", ctypes.c_uint32),
    ]


class GpmaServerConfig(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("max_bfs_group", ctypes.c_uint32),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("max_fused_records", ctypes.c_uint64),
        ("max_queued_records", ctypes.c_uint64),
        ("order", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 3),
    ]


class GpmaRequestResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("kind", ctypes.c_uint32),
        ("reachable", ctypes.c_uint32),
        ("group", ctypes.c_uint32),
        ("queue_us", ctypes.c_double),
        ("service_us", ctypes.c_double),
        ("latency_us", ctypes.c_double),
        ("state_ticket", ctypes.c_uint64),
    ]


class GpmaServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("update_launches", ctypes.c_uint64),
        ("update_records", ctypes.c_uint64),
        ("bfs_launches", ctypes.c_uint64),
        ("bfs_requests", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
        ("max_fused_records", ctypes.c_uint64),
        ("max_bfs_group", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("deferred", ctypes.c_uint64),
        ("staged_ahead", ctypes.c_uint64),
        ("budget_bytes", ctypes.c_uint64),
        ("growth_reserve_bytes", ctypes.c_uint64),
    ]


# name -> (restype, argtypes); mirrors include/gpma.h and include/gpma_serve.h one to one
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_I = ctypes.c_int
SIGNATURES = {
    "gpma_create": (_I, [_U32, _P, _P]),
    "gpma_destroy": (_I, [_P]),
    "gpma_update": (_I, [_P, _P, _P, _U64, _I]),
    "gpma_update_batches": (_I, [_P, _P, _P, _P, _U32, _I]),
    "gpma_bfs": (_I, [_P, _U32, _P, _I, _P]),
    "gpma_bfs_multi": (_I, [_P, _P, _U32, _P, _I, _P]),
    "gpma_step": (_I, [_P, _P, _P, _P, _U32, _I, _P, _U32, _P, _I, _P]),
    "gpma_get_step_stats": (_I, [_P, _P]),
    "gpma_step_prepare": (_I, [_P, _U64, _U32, _P, _U32]),
    "gpma_csr_bfs": (_I, [_P, _P, _P, _U32, _U32, _U32, _P]),
    "gpma_csr_bfs_noalloc": (_I, [_P, _P, _P, _U32, _U32, _U32, _P, _P, _P]),
    "gpma_get_csr": (_I, [_P, _P]),
    "gpma_copy_layout": (_I, [_P, _P, _P, _P]),
    "gpma_export_live": (_I, [_P, _P, _U64, _P]),
    "gpma_get_stats": (_I, [_P, _P]),
    "gpma_live_digest": (_I, [_P, _P, _P, _P]),
    "gpma_test_fail_allocs": (_I, [_P, _U32, _U32]),
    "gpma_sync": (_I, [_P]),
    "gpma_stream": (_P, [_P]),
    "gpma_last_error": (ctypes.c_char_p, []),
    "gpma_device_count": (_I, []),
    "gpma_profile": (_I, [_P, _I]),
    "gpma_profile_phases": (_I, [_P, _U32]),
    "gpma_get_profile": (_I, [_P, _P, _P, _U32]),
    "gpma_device_malloc": (_I, [_P, _U64]),
    "gpma_device_free": (_I, [_P]),
    "gpma_memcpy": (_I, [_P, _P, _U64, _I]),
    "gpma_mem_info": (_I, [ctypes.c_int32, _P, _P]),
    "gpma_copy_async": (_I, [_P, _P, _P, _U64, _I]),
    "gpma_device_of": (_I, [_P]),
    "gpma_server_create": (_I, [_P, _P, _P]),
    "gpma_server_destroy": (_I, [_P]),
    "gpma_server_submit_update": (_I, [_P, _P, _P, _U64, _I, _P]),
    "gpma_server_submit_bfs": (_I, [_P, _U32, _P, _P]),
    "gpma_server_wait": (_I, [_P, _U64, ctypes.c_int64, _P]),
    "gpma_server_get_stats": (_I, [_P, _P]),
    "gpma_gen_rmat": (_I, [_U32, _U64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _U64, _P, _P, _U64, _P]),
    "gpma_gen_powerlaw": (_I, [_U32, _U64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _U64, _P, _P]),
    "gpma_read_edge_text": (_I, [ctypes.c_char_p, _P, _P, _P, _P, _U64]),
    "gpma_write_edge_text": (_I, [ctypes.c_char_p, _U32, _U64, _P, _P]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run gpma_bfs_amd/build.py (or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error():
    return (lib.gpma_last_error() or b"").decode(errors="replace")


def check(rc, what=""):
    if rc != GPMA_OK:
        raise GpmaError(rc, f"{what}: {last_error()}")
    return rc
