"""CPU checks of the C-ABI boundary: the engine library loads and exports
every function include/luda.h declares, and the ctypes mirror of the structs
has the C layout."""
import ctypes as C
import os
import re

from luda_amd import ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "luda.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(luda_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(ffi.ENGINE_LIB)
    names = declared_functions()
    assert "luda_compact" in names and "luda_build_tables" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(ffi.ENGINE_SIGNATURES)


def test_version_string_without_a_gpu():
    assert b"gfx950" in ffi.engine().luda_version()


def test_struct_layouts():
    # sizes fixed by include/luda.h on LP64
    assert C.sizeof(ffi.luda_options) == 32
    assert C.sizeof(ffi.luda_file) == 24
    assert C.sizeof(ffi.luda_grandparent) == 24
    assert C.sizeof(ffi.luda_level_file) == 40
    assert C.sizeof(ffi.luda_output) == 48
    assert ffi.luda_job.inputs.offset == 48 and ffi.luda_job.deeper.offset == 88
    assert C.sizeof(ffi.luda_job) == 96


def test_init_fails_cleanly_without_device():
    lib = ffi.engine()
    ctx = C.c_void_p()
    err = C.create_string_buffer(256)
    rc = lib.luda_init(99, 0, C.byref(ctx), err, 256)
    assert rc != ffi.LUDA_OK and not ctx.value


def test_queued_entry_points_reject_a_missing_context():
    lib = ffi.engine()
    opts = ffi.default_options()
    assert lib.luda_submit(None, None, None, None) == ffi.LUDA_INVALID_ARGUMENT
    assert lib.luda_submit_build(None, C.byref(opts), 0, None, None, None, None, 0, None,
                                 None) == ffi.LUDA_INVALID_ARGUMENT
    assert lib.luda_poll(None) == 0 and lib.luda_drain(None) == 0


def test_stats_layout():
    # luda_stats ends with end_clock_ms, decode_ahead and a reserved word (include/luda.h)
    assert ffi.luda_stats.end_clock_ms.offset == C.sizeof(ffi.luda_stats) - 16
    assert ffi.luda_stats.decode_ahead.offset == C.sizeof(ffi.luda_stats) - 8


def _init_with_fault(kind):
    lib = ffi.engine()
    ctx = C.c_void_p()
    err = C.create_string_buffer(256)
    os.environ["LUDA_FAULT"] = "init:" + kind
    try:
        rc = lib.luda_init(0, 0, C.byref(ctx), err, 256)
    finally:
        del os.environ["LUDA_FAULT"]
    return rc, ctx.value, err.value.decode()


def test_exceptions_never_cross_the_abi():
    """A C++ exception other than the engine's own (host allocation failure,
    a standard exception, anything thrown) comes back as LUDA_IO_ERROR with a
    message, not as std::terminate in the caller's thread."""
    rc, ctx, msg = _init_with_fault("bad_alloc")
    assert rc == ffi.LUDA_IO_ERROR and not ctx and "memory" in msg
    rc, ctx, msg = _init_with_fault("runtime")
    assert rc == ffi.LUDA_IO_ERROR and not ctx and "injected" in msg
    rc, ctx, msg = _init_with_fault("other")
    assert rc == ffi.LUDA_IO_ERROR and not ctx and msg == "internal error"


def test_edit_layouts():
    assert C.sizeof(ffi.luda_edit) == 88
    assert C.sizeof(ffi.luda_edit_file) == 56
    assert C.sizeof(ffi.luda_file_set) == 40


def test_tuning_layout():
    assert C.sizeof(ffi.luda_tuning) == 16 and C.sizeof(ffi.luda_tuning_rule) == 40
