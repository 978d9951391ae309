"""ctypes binding of libhive.so (the C-ABI declared in include/hive.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
import ctypes as C
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libhive.so")
HEADER = os.path.join(ROOT, "include", "hive.h")

HIVE_OK, HIVE_EINVAL, HIVE_ENOMEM, HIVE_EDEVICE, HIVE_ECAPACITY, HIVE_ERANGE = range(6)
R_FAILED, R_INSERTED, R_UPDATED, R_STASHED, R_FOUND, R_NOT_FOUND, R_ERASED, R_INVALID_KEY, R_NOOP = range(9)
OP_INSERT, OP_LOOKUP, OP_DELETE, OP_ATOMIC_INC, OP_NONE = range(5)
STREAM_LAUNCHES, STREAM_GRAPH, STREAM_PERSISTENT, STREAM_WINDOWED = range(4)
SEQUENCE_PLAIN, SEQUENCE_GRAPH = range(2)
LAYOUTS = {"bucket8": 0, "bucket16": 1, "inline8": 2}  # hive_layout_t

POLICIES = {
    "Default2Hash": 0,
    "MurmurCityHash": 1,
    "TripleHash": 2,
    "MurmurCityBitHash": 3,
    "Lookup2Hash": 4,
    "Lookup3Hash": 5,
}


class HiveError(RuntimeError):
    def __init__(self, status, where, detail=""):
        self.status = status
        super().__init__(f"{where}: status {status} ({detail})")


class Stats(C.Structure):
    _fields_ = [
        ("num_buckets", C.c_uint64),
        ("slots_per_bucket", C.c_uint64),
        ("capacity", C.c_uint64),
        ("table_live", C.c_uint64),
        ("stash_live", C.c_uint64),
        ("load_factor", C.c_double),
        ("device_bytes", C.c_uint64),
        ("policy", C.c_int),
        ("slow_path_ops", C.c_uint64),
        ("layout", C.c_int),
    ]


class Launch(C.Structure):
    _fields_ = [("block_threads", C.c_uint32), ("keys_per_lane", C.c_uint32), ("inline_slow_path", C.c_uint32),
                ("combine_duplicates", C.c_uint32), ("split_by_type", C.c_uint32), ("lanes_per_key", C.c_uint32)]


class Tune(C.Structure):
    _fields_ = [("block_threads", C.c_uint32), ("inline_slow_path", C.c_uint32), ("split_by_type", C.c_uint32),
                ("tile_ops", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("stream_inline", C.c_uint32), ("best_mode", C.c_int),
                ("launches_ms", C.c_double), ("graph_ms", C.c_double), ("persistent_ms", C.c_double),
                ("candidates", C.c_uint32), ("window_ops", C.c_uint64), ("window_batch_ops", C.c_uint64),
                ("windowed_ms", C.c_double), ("windowed_p99_us", C.c_double), ("launches_p99_us", C.c_double),
                ("graph_p99_us", C.c_double), ("persistent_p99_us", C.c_double), ("p99_bound_us", C.c_double),
                ("bounded_window_ops", C.c_uint64), ("bounded_window_batch_ops", C.c_uint64),
                ("bounded_windowed_ms", C.c_double), ("bounded_windowed_p99_us", C.c_double),
                ("bounded_split_by_type", C.c_uint32), ("windowed_ms_by_size", C.c_double * 8),
                ("windowed_p99_us_by_size", C.c_double * 8)]


class TuneOpts(C.Structure):
    _fields_ = [("try_persistent", C.c_uint32), ("max_p99_us", C.c_double), ("p99_slack", C.c_double)]


class BulkTune(C.Structure):
    _fields_ = [("insert_block_threads", C.c_uint32), ("insert_inline_slow_path", C.c_uint32),
                ("lookup_block_threads", C.c_uint32), ("lookup_lanes_per_key", C.c_uint32),
                ("insert_ms", C.c_double), ("lookup_ms", C.c_double), ("lookup_ms_by_lanes", C.c_double * 3),
                ("load_factor", C.c_double), ("candidates", C.c_uint32), ("lookup_keys_per_lane", C.c_uint32),
                ("lookup_ms_by_keys", C.c_double * 3), ("insert_ms_by_slow_path", C.c_double * 3)]


class LayoutTune(C.Structure):
    _fields_ = [("layout", C.c_int), ("insert_ms", C.c_double * 3), ("lookup_ms", C.c_double * 3),
                ("candidates", C.c_uint32)]


class ServeConfig(C.Structure):
    _fields_ = [("max_batch_ops", C.c_uint64), ("max_inflight", C.c_uint32), ("max_load_factor", C.c_double),
                ("growth", C.c_double), ("hbm_reserve_bytes", C.c_uint64), ("stash_headroom", C.c_double),
                ("ordered", C.c_uint32), ("min_load_factor", C.c_double), ("shrink_load_factor", C.c_double)]


class ServeReport(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("batches", C.c_uint64), ("ops", C.c_uint64),
                ("rejected_requests", C.c_uint64), ("resizes", C.c_uint64), ("final_slots", C.c_uint64),
                ("span_ms", C.c_double), ("mops", C.c_double), ("p50_us", C.c_double), ("p99_us", C.c_double),
                ("max_us", C.c_double), ("mean_batch_ops", C.c_double), ("resize_ms", C.c_double),
                ("shrinks", C.c_uint64), ("shrink_ms", C.c_double)]


class SequenceReport(C.Structure):
    _fields_ = [("elapsed_ms", C.c_double), ("setup_ms", C.c_double), ("resizes", C.c_uint32),
                ("final_slots", C.c_uint64), ("rehash_failed", C.c_uint64)]


class Operation(C.Structure):
    """The reference's 16-byte Operation record (include/utils.h:18-21)."""
    _fields_ = [("type", C.c_int32), ("pad", C.c_uint32), ("key", C.c_uint64)]


_lib = None


def declared_symbols():
    """Names of every function declared in include/hive.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:hive_status_t|const char\*|uint64_t)\s+(hive_\w+)\s*\(", text, re.M)))


def lib():
    """Load libhive.so (built in-tree by hivehashtable_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m hivehashtable_amd.build` (needs hipcc)")
    L = C.CDLL(LIB_PATH)
    st, p, sz, u32, u64, i32 = C.c_int, C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint64, C.c_int
    proto = {
        "hive_create": [u64, i32, C.POINTER(p)],
        "hive_create_ex": [u64, i32, i32, C.POINTER(p)],
        "hive_set_layout": [p, i32, p],
        "hive_ceiling_gather_bytes": [p, u64, u64, u32, u32, p, p],
        "hive_destroy": [p],
        "hive_clear": [p, p],
        "hive_policy_from_name": [C.c_char_p, C.POINTER(i32)],
        "hive_insert": [p, p, p, sz, p, p],
        "hive_lookup": [p, p, sz, p, p, p],
        "hive_erase": [p, p, sz, p, p],
        "hive_mixed": [p, p, p, p, sz, p, p, p],
        "hive_mixed_noalloc": [p, p, p, p, sz, p, p, p],
        "hive_insert_ops": [p, p, sz, p, p],
        "hive_lookup_ops": [p, p, sz, p, p],
        "hive_mixed_ops": [p, p, sz, p, p],
        "hive_resize": [p, u64, p],
        "hive_stash_drain": [p, p, sz, C.POINTER(sz), p],
        "hive_dump": [p, p, sz, C.POINTER(sz), p],
        "hive_stash_reinsert": [p, C.POINTER(sz), p],
        "hive_stats": [p, C.POINTER(Stats), p],
        "hive_count_exact": [p, C.POINTER(u64), C.POINTER(u64), p],
        "hive_probe_stats": [p, p, sz, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), p],
        "hive_export_raw": [p, p, p, p, p],
        "hive_import_raw": [p, p, p, p, p],
        "hive_set_launch": [p, i32, C.POINTER(Launch)],
        "hive_get_launch": [p, i32, C.POINTER(Launch)],
        "hive_hash_table_kernel_dispatch": [p, sz, sz, sz, sz, sz, C.POINTER(C.c_double), p, i32, i32, C.c_char_p],
        "hive_hash_table_kernel_dispatch_ycsb": [p, p, sz, sz, sz, sz, sz, C.POINTER(C.c_double), p, i32, i32],
        "hive_mixed_stream": [p, p, p, p, p, sz, p, p, i32, p, C.POINTER(C.c_double), p],
        "hive_set_stream_shape": [p, u32, u32, u32],
        "hive_set_stream_window": [p, C.c_uint64, C.c_uint64],
        "hive_autotune_stream": [p, p, p, p, p, sz, C.POINTER(Tune), p],
        "hive_autotune_stream_ex": [p, p, p, p, p, sz, C.POINTER(TuneOpts), C.POINTER(Tune), p],
        "hive_autotune_bulk": [p, p, p, sz, p, sz, C.POINTER(BulkTune), p],
        "hive_autotune_layout": [p, p, p, sz, p, sz, C.POINTER(LayoutTune), p],
        "hive_serve_trace": [p, p, p, p, p, p, sz, p, p, C.POINTER(ServeConfig), p, C.POINTER(ServeReport), p],
        "hive_insert_sequence": [p, p, p, sz, sz, p, sz, C.c_double, C.c_double, i32, p, p,
                                 C.POINTER(SequenceReport), p],
        "hive_ceiling_gather": [p, u64, u64, u32, p, p],
        "hive_ceiling_cas": [p, u64, u64, u32, p],
        "hive_ceiling_gather_lanes": [p, u64, u64, u32, u32, p, p],
    }
    for name, args in proto.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = st
    L.hive_num_buckets.argtypes = [p]
    L.hive_num_buckets.restype = u64
    L.hive_status_string.argtypes = [st]
    L.hive_status_string.restype = C.c_char_p
    L.hive_last_error.argtypes = []
    L.hive_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(status, where):
    if status != HIVE_OK:
        L = lib()
        detail = (L.hive_last_error() or b"").decode(errors="replace")
        raise HiveError(status, where, detail)
