"""Loader for the oracle (TEST INFRASTRUCTURE ONLY).

liblsm_oracle.so is the C restatement of the reference's CPU table format and
compaction; oracle/_ref/libluda_ref.so (when it was built from the reference
sources) is the reference's own code.  Only tests/, __graft_entry__.smoke()
and the CPU-baseline leg of bench.py load these; the engine never does.
"""
import ctypes as C
import os

from luda_amd import ffi
from luda_amd.jobs import Table, result_tables

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liblsm_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libluda_ref.so")

_oracle = None
_ref = None


def _bind_common(lib, prefix):
    f = getattr(lib, prefix + "build_tables")
    f.restype = C.c_int
    f.argtypes = ffi.BUILD_ARGTYPES + [C.POINTER(ffi.RESULT_P)]
    f = getattr(lib, prefix + "compact")
    f.restype = C.c_int
    f.argtypes = [C.POINTER(ffi.luda_job), C.POINTER(ffi.RESULT_P)]
    f = getattr(lib, prefix + "free_result")
    f.restype = None
    f.argtypes = [ffi.RESULT_P]


def load():
    global _oracle
    if _oracle is None:
        lib = C.CDLL(ORACLE_LIB)
        _bind_common(lib, "or_")
        u32, u64, sz, vp = C.c_uint32, C.c_uint64, C.c_size_t, C.c_void_p
        sigs = {
            "or_encode_varint64": (sz, [C.c_char_p, u64]),
            "or_varint_length": (sz, [u64]),
            "or_crc32c_extend": (u32, [u32, C.c_char_p, sz]),
            "or_crc32c_value": (u32, [C.c_char_p, sz]),
            "or_crc32c_mask": (u32, [u32]),
            "or_crc32c_unmask": (u32, [u32]),
            "or_hash": (u32, [C.c_char_p, sz, u32]),
            "or_bloom_create": (sz, [C.c_char_p, C.POINTER(u64), u32, u32, C.c_char_p, sz]),
            "or_bloom_may_match": (C.c_int, [C.c_char_p, sz, C.c_char_p, sz]),
            "or_ikey_compare": (C.c_int, [C.c_char_p, sz, C.c_char_p, sz]),
            "or_shortest_separator": (sz, [C.c_char_p, sz, C.c_char_p, sz, C.c_char_p]),
            "or_short_successor": (sz, [C.c_char_p, sz, C.c_char_p]),
            "or_filter_new": (vp, [C.c_int, u32]),
            "or_filter_start_block": (None, [vp, u64]),
            "or_filter_add_key": (None, [vp, C.c_char_p, sz]),
            "or_filter_finish": (C.POINTER(C.c_uint8), [vp, C.POINTER(sz)]),
            "or_filter_free": (None, [vp]),
            "or_filter_key_may_match": (C.c_int, [C.c_int, C.c_char_p, sz, u64, C.c_char_p, sz]),
            "or_table_iterate": (C.c_int, [C.c_char_p, u64, C.c_int, vp, vp, C.POINTER(u64)]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_LIB)


def load_ref():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_LIB)
        _bind_common(lib, "ref_")
        lib.ref_encode_edit.restype = C.c_int
        lib.ref_encode_edit.argtypes = [C.POINTER(ffi.luda_edit), C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]
        lib.ref_log_record.restype = C.c_int
        lib.ref_log_record.argtypes = [C.POINTER(C.c_uint64), C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                       C.POINTER(C.c_uint64)]
        _ref = lib
    return _ref


def _run(lib, prefix, fn, *args):
    res = ffi.RESULT_P()
    rc = getattr(lib, prefix + fn)(*args, C.byref(res))
    if rc != ffi.LUDA_OK:
        return rc, None
    try:
        return rc, result_tables(res)
    finally:
        getattr(lib, prefix + "free_result")(res)


def builder(which="oracle"):
    """A table builder for luda_amd.jobs.build_group_tables backed by the
    oracle ("oracle") or the compiled reference ("ref")."""
    lib, prefix = (load(), "or_") if which == "oracle" else (load_ref(), "ref_")

    def build(opts_p, n, keys, key_off, vals, val_off):
        rc, tables = _run(lib, prefix, "build_tables", opts_p, n, keys, key_off, vals, val_off)
        if rc != ffi.LUDA_OK:
            raise RuntimeError(f"{prefix}build_tables failed: {rc}")
        return tables

    return build


def compact(job, which="oracle"):
    """Runs a luda_job (host inputs) through the oracle or the compiled
    reference; returns (status, [Table])."""
    lib, prefix = (load(), "or_") if which == "oracle" else (load_ref(), "ref_")
    return _run(lib, prefix, "compact", C.byref(job))


ENTRY_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_uint8), C.c_size_t, C.POINTER(C.c_uint8), C.c_size_t)


def table_entries(data, verify_crc=True):
    """(status, [(key, value)]) of one table, with block CRCs checked."""
    lib = load()
    out = []

    def cb(_u, k, kl, v, vl):
        out.append((C.string_at(k, kl), C.string_at(v, vl) if vl else b""))

    fn = ENTRY_FN(cb)
    n = C.c_uint64()
    rc = lib.or_table_iterate(bytes(data), len(data), 1 if verify_crc else 0, C.cast(fn, C.c_void_p), None,
                              C.byref(n))
    return rc, out


def build_tables(entries, opts, which="oracle"):
    """Tables built from a Python list of (internal key, value) by the oracle
    or the compiled reference, cut at opts.max_file_size."""
    keys = b"".join(k for k, _ in entries)
    vals = b"".join(v for _, v in entries)
    n = len(entries)
    ko = (C.c_uint64 * (n + 1))()
    vo = (C.c_uint64 * (n + 1))()
    a = b = 0
    for i, (k, v) in enumerate(entries):
        ko[i], vo[i] = a, b
        a += len(k)
        b += len(v)
    ko[n], vo[n] = a, b
    kb = C.create_string_buffer(keys, len(keys) + 1)
    vb = C.create_string_buffer(vals, len(vals) + 1)
    return builder(which)(C.byref(opts), n, C.cast(kb, C.c_void_p), C.cast(ko, C.c_void_p), C.cast(vb, C.c_void_p),
                          C.cast(vo, C.c_void_p))
