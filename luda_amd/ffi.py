"""ctypes mirror of include/luda.h and luda_amd/workload/workload.h.

The engine's boundary is the C-ABI in include/luda.h; this module only
declares its structs for Python callers (tests, bench.py).  It loads
libluda.so (the HIP engine) and libluda_workload.so (the host-only job
generator) from this package directory and never falls back to anything else:
a missing library raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ENGINE_LIB = os.path.join(HERE, "libluda.so")
WORKLOAD_LIB = os.path.join(HERE, "libluda_workload.so")

LUDA_OK = 0
LUDA_NOT_FOUND = 1
LUDA_CORRUPTION = 2
LUDA_NOT_SUPPORTED = 3
LUDA_INVALID_ARGUMENT = 4
LUDA_IO_ERROR = 5

LUDA_JOB_INPUTS_ON_DEVICE = 0x1
LUDA_JOB_KEEP_ON_DEVICE = 0x2
LUDA_JOB_KEEP_DECODED = 0x8

U64_MAX = (1 << 64) - 1


class luda_options(C.Structure):
    _fields_ = [
        ("block_size", C.c_uint64),
        ("max_file_size", C.c_uint64),
        ("block_restart_interval", C.c_uint32),
        ("bloom_bits_per_key", C.c_uint32),
        ("compression", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


def default_options(**kw):
    """leveldb::Options defaults (include/leveldb/options.h:100-142) with the
    10-bit bloom filter every BASELINE workload uses."""
    o = luda_options(block_size=4096, max_file_size=2 << 20, block_restart_interval=16,
                     bloom_bits_per_key=10, compression=0, reserved=0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class luda_file(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_uint64), ("number", C.c_uint64)]


class luda_slice(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_uint64)]


class luda_grandparent(C.Structure):
    _fields_ = [("largest", luda_slice), ("file_size", C.c_uint64)]


class luda_level_file(C.Structure):
    _fields_ = [
        ("smallest_user_key", luda_slice),
        ("largest_user_key", luda_slice),
        ("level", C.c_int32),
        ("reserved", C.c_uint32),
    ]


class luda_job(C.Structure):
    _fields_ = [
        ("level", C.c_int32),
        ("flags", C.c_uint32),
        ("smallest_snapshot", C.c_uint64),
        ("options", luda_options),
        ("inputs", C.POINTER(luda_file) * 2),
        ("num_inputs", C.c_uint32 * 2),
        ("grandparents", C.POINTER(luda_grandparent)),
        ("num_grandparents", C.c_uint32),
        ("num_deeper", C.c_uint32),
        ("deeper", C.POINTER(luda_level_file)),
    ]


class luda_output(C.Structure):
    _fields_ = [
        ("offset", C.c_uint64),
        ("file_size", C.c_uint64),
        ("num_entries", C.c_uint64),
        ("smallest_offset", C.c_uint64),
        ("largest_offset", C.c_uint64),
        ("smallest_size", C.c_uint32),
        ("largest_size", C.c_uint32),
    ]


class luda_stats(C.Structure):
    _fields_ = [
        ("device_ms", C.c_double),
        ("total_ms", C.c_double),
        ("h2d_ms", C.c_double),
        ("d2h_ms", C.c_double),
        ("top_kernel_ms", C.c_double),
        ("top_kernel_launches", C.c_uint64),
        ("top_kernel", C.c_char * 48),
        ("bytes_read", C.c_uint64),
        ("bytes_written", C.c_uint64),
        ("entries_in", C.c_uint64),
        ("entries_kept", C.c_uint64),
        ("blocks_in", C.c_uint64),
        ("blocks_out", C.c_uint64),
        ("attempts", C.c_uint32),
        ("key_width", C.c_uint32),
        ("data_block_bytes", C.c_uint64),
        ("top_kernel_bytes", C.c_uint64),
        ("end_clock_ms", C.c_double),
        ("decode_ahead", C.c_uint32),
        ("reserved_stats", C.c_uint32),
    ]


class luda_result(C.Structure):
    _fields_ = [
        ("num_outputs", C.c_uint32),
        ("reserved", C.c_uint32),
        ("outputs", C.POINTER(luda_output)),
        ("keys", C.POINTER(C.c_uint8)),
        ("data", C.POINTER(C.c_uint8)),
        ("device_data", C.c_void_p),
        ("total_bytes", C.c_uint64),
        ("stats", luda_stats),
    ]


class lw_group(C.Structure):
    _fields_ = [
        ("which", C.c_int32),
        ("reserved", C.c_uint32),
        ("max_file_size", C.c_uint64),
        ("n", C.c_uint64),
        ("keys", C.POINTER(C.c_uint8)),
        ("key_off", C.POINTER(C.c_uint64)),
        ("vals", C.POINTER(C.c_uint8)),
        ("val_off", C.POINTER(C.c_uint64)),
    ]


class lw_workload(C.Structure):
    _fields_ = [
        ("level", C.c_int32),
        ("num_groups", C.c_uint32),
        ("smallest_snapshot", C.c_uint64),
        ("groups", C.POINTER(lw_group)),
        ("num_grandparents", C.c_uint32),
        ("num_deeper", C.c_uint32),
        ("grandparents", C.POINTER(luda_grandparent)),
        ("deeper", C.POINTER(luda_level_file)),
        ("meta_keys", C.POINTER(C.c_uint8)),
        ("total_entries", C.c_uint64),
    ]


RESULT_P = C.POINTER(luda_result)


class luda_file_set(C.Structure):
    _fields_ = [("n", C.c_uint32), ("reserved", C.c_uint32), ("files", C.POINTER(luda_file)),
                ("device_data", C.c_void_p), ("total_bytes", C.c_uint64), ("read_ms", C.c_double)]


FILESET_P = C.POINTER(luda_file_set)

LUDA_READ_VERIFY = 0x4


class luda_tuning(C.Structure):
    _fields_ = [("write_blocks_groups", C.c_uint32), ("decode_groups", C.c_uint32), ("reserved", C.c_uint32 * 2)]


class luda_tuning_rule(C.Structure):
    _fields_ = [("min_bytes", C.c_uint64), ("max_bytes", C.c_uint64), ("min_key_bytes", C.c_uint32),
                ("max_key_bytes", C.c_uint32), ("tuning", luda_tuning)]


class luda_entries(C.Structure):
    _fields_ = [("n", C.c_uint64), ("num_tables", C.c_uint32), ("reserved", C.c_uint32),
                ("keys", C.POINTER(C.c_uint8)), ("key_off", C.POINTER(C.c_uint64)),
                ("vals", C.POINTER(C.c_uint8)), ("val_off", C.POINTER(C.c_uint64)),
                ("table_first", C.POINTER(C.c_uint64))]


ENTRIES_P = C.POINTER(luda_entries)

LUDA_EDIT_HAS_COMPARATOR = 0x1
LUDA_EDIT_HAS_LOG_NUMBER = 0x2
LUDA_EDIT_HAS_PREV_LOG_NUMBER = 0x4
LUDA_EDIT_HAS_NEXT_FILE = 0x8
LUDA_EDIT_HAS_LAST_SEQUENCE = 0x10


class luda_edit_pointer(C.Structure):
    _fields_ = [("level", C.c_int32), ("reserved", C.c_uint32), ("key", luda_slice)]


class luda_edit_deleted(C.Structure):
    _fields_ = [("level", C.c_int32), ("reserved", C.c_uint32), ("number", C.c_uint64)]


class luda_edit_file(C.Structure):
    _fields_ = [("level", C.c_int32), ("reserved", C.c_uint32), ("number", C.c_uint64), ("file_size", C.c_uint64),
                ("smallest", luda_slice), ("largest", luda_slice)]


class luda_edit(C.Structure):
    _fields_ = [("has", C.c_uint32), ("num_compact_pointers", C.c_uint32), ("comparator", luda_slice),
                ("log_number", C.c_uint64), ("prev_log_number", C.c_uint64), ("next_file_number", C.c_uint64),
                ("last_sequence", C.c_uint64), ("compact_pointers", C.POINTER(luda_edit_pointer)),
                ("deleted", C.POINTER(luda_edit_deleted)), ("num_deleted", C.c_uint32),
                ("num_new_files", C.c_uint32), ("new_files", C.POINTER(luda_edit_file))]

BUILD_ARGTYPES = [C.POINTER(luda_options), C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]

# Every symbol include/luda.h declares, with its ctypes signature.
ENGINE_SIGNATURES = {
    "luda_init": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]),
    "luda_shutdown": (None, [C.c_void_p]),
    "luda_compact": (C.c_int, [C.c_void_p, C.POINTER(luda_job), C.POINTER(RESULT_P), C.c_char_p, C.c_size_t]),
    "luda_build_tables": (C.c_int, [C.c_void_p, C.POINTER(luda_options), C.c_uint64, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(RESULT_P), C.c_char_p,
                                    C.c_size_t]),
    "luda_free_result": (None, [C.c_void_p, RESULT_P]),
    "luda_verify_tables": (C.c_int, [C.c_void_p, C.POINTER(luda_file), C.c_uint32, C.c_uint32,
                                     C.POINTER(C.c_uint32), C.c_char_p, C.c_size_t]),
    "luda_verify_tables_noalloc": (C.c_int, [C.c_void_p, C.POINTER(luda_file), C.c_uint32, C.c_uint32,
                                             C.POINTER(C.c_uint32), C.c_char_p, C.c_size_t]),
    "luda_device_alloc": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "luda_device_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "luda_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "luda_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "luda_version": (C.c_char_p, []),
    "luda_set_slots": (C.c_int, [C.c_void_p, C.c_uint32]),
    "luda_submit": (C.c_int, [C.c_void_p, C.POINTER(luda_job), C.c_void_p, C.c_void_p]),
    "luda_submit_build": (C.c_int, [C.c_void_p, C.POINTER(luda_options), C.c_uint64, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "luda_read_files": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64), C.c_uint32,
                                  C.POINTER(FILESET_P), C.c_char_p, C.c_size_t]),
    "luda_free_files": (None, [C.c_void_p, FILESET_P]),
    "luda_write_outputs": (C.c_int, [C.c_void_p, RESULT_P, C.POINTER(C.c_char_p), C.c_int, C.c_char_p, C.c_size_t]),
    "luda_encode_edit": (C.c_int, [C.POINTER(luda_edit), C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "luda_encode_compaction_edit": (C.c_int, [C.POINTER(luda_job), RESULT_P, C.POINTER(C.c_uint64),
                                              C.POINTER(luda_edit), C.c_void_p, C.c_uint64,
                                              C.POINTER(C.c_uint64)]),
    "luda_log_record": (C.c_int, [C.POINTER(C.c_uint64), C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                  C.POINTER(C.c_uint64)]),
    "luda_read_tables": (C.c_int, [C.c_void_p, C.POINTER(luda_file), C.c_uint32, C.c_uint32, C.POINTER(ENTRIES_P),
                                   C.c_char_p, C.c_size_t]),
    "luda_free_entries": (None, [ENTRIES_P]),
    "luda_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(luda_tuning_rule), C.c_uint32]),
    "luda_poll": (C.c_int, [C.c_void_p]),
    "luda_drain": (C.c_int, [C.c_void_p]),
}

# void (*luda_done_fn)(void* user, int status, luda_result* result, const char* err)
DONE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, RESULT_P, C.c_char_p)


def _bind(lib, table):
    for name, (res, args) in table.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_engine = None
_workload = None


def engine():
    """The HIP engine.  Raises if libluda.so is missing: there is no other
    implementation to fall back to."""
    global _engine
    if _engine is None:
        if not os.path.exists(ENGINE_LIB):
            raise RuntimeError(f"{ENGINE_LIB} is not built; run __graft_entry__.build()")
        _engine = _bind(C.CDLL(ENGINE_LIB), ENGINE_SIGNATURES)
    return _engine


def workload_lib():
    global _workload
    if _workload is None:
        if not os.path.exists(WORKLOAD_LIB):
            raise RuntimeError(f"{WORKLOAD_LIB} is not built; run __graft_entry__.build()")
        lib = C.CDLL(WORKLOAD_LIB)
        lib.lw_make.restype = C.POINTER(lw_workload)
        lib.lw_make.argtypes = [C.c_int, C.c_uint64, C.c_double]
        lib.lw_free.restype = None
        lib.lw_free.argtypes = [C.POINTER(lw_workload)]
        _workload = lib
    return _workload
