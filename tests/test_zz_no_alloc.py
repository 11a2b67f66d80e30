"""luda_verify_tables_noalloc, the entry point for callers that cannot
allocate, against luda_verify_tables on the same device-resident tables:
two calls in a row (the second on fewer tables), a corrupted table, and the
limits it rejects."""
import ctypes as C

import pytest

from luda_amd import ffi, jobs
from tests.tiny_jobs import ikey

pytestmark = pytest.mark.gpu


def make_tables(eng, n_entries, seed):
    import random
    rng = random.Random(seed)
    ents = [(ikey(b"nk%07d" % i, 1 + i, 1), bytes(rng.randrange(256) for _ in range(rng.randint(0, 120))))
            for i in range(n_entries)]
    opts = ffi.default_options(max_file_size=64 << 10)
    keys = b"".join(k for k, _ in ents)
    vals = b"".join(v for _, v in ents)
    ko = (C.c_uint64 * (n_entries + 1))()
    vo = (C.c_uint64 * (n_entries + 1))()
    a = b = 0
    for i, (k, v) in enumerate(ents):
        ko[i], vo[i] = a, b
        a += len(k)
        b += len(v)
    ko[n_entries], vo[n_entries] = a, b
    kb = C.create_string_buffer(keys, len(keys) + 1)
    vb = C.create_string_buffer(vals, len(vals) + 1)
    rc, out, msg = eng.build_tables(C.byref(opts), n_entries, C.cast(kb, C.c_void_p), C.cast(ko, C.c_void_p),
                                    C.cast(vb, C.c_void_p), C.cast(vo, C.c_void_p), keep_on_device=True)
    assert rc == 0, msg
    res = out[2]
    r = res.contents
    tabs = [(r.device_data + r.outputs[i].offset, r.outputs[i].file_size) for i in range(r.num_outputs)]
    return res, tabs


def both(eng, tabs):
    arr = (ffi.luda_file * max(1, len(tabs)))()
    for i, (p, n) in enumerate(tabs):
        arr[i].data, arr[i].size, arr[i].number = p, n, i + 1
    out = []
    for fn in (eng.lib.luda_verify_tables, eng.lib.luda_verify_tables_noalloc):
        bad = C.c_uint32()
        err = C.create_string_buffer(256)
        rc = fn(eng.ctx, arr, len(tabs), ffi.LUDA_JOB_INPUTS_ON_DEVICE, C.byref(bad), err, 256)
        out.append((rc, bad.value))
    return out


def test_noalloc_verify_equals_verify():
    eng = jobs.Engine(0)
    try:
        print("noalloc: 40000 entries", flush=True)
        res1, tabs1 = make_tables(eng, 40000, 1)
        assert len(tabs1) > 20
        a, b = both(eng, tabs1)
        assert a == b == (ffi.LUDA_OK, 0xffffffff)
        # second call, fewer tables: the array keeps the last call's contents
        print("noalloc: 5 tables", flush=True)
        a, b = both(eng, tabs1[3:8])
        assert a == b == (ffi.LUDA_OK, 0xffffffff)
        # a flipped byte inside table 4's first data block
        p, n = tabs1[4]
        byte = eng.d2h(p + 100, 1)
        eng.h2d(p + 100, bytes([byte[0] ^ 0x40]))
        print("noalloc: corrupted table 4", flush=True)
        a, b = both(eng, tabs1)
        assert a == b == (ffi.LUDA_CORRUPTION, 4)
        a, b = both(eng, tabs1[2:6])
        assert a == b == (ffi.LUDA_CORRUPTION, 2)
        eng.h2d(p + 100, byte)
        a, b = both(eng, tabs1[:2])
        assert a == b == (ffi.LUDA_OK, 0xffffffff)
        # limits: host tables, too many bytes
        arr = (ffi.luda_file * 1)()
        arr[0].data, arr[0].size = tabs1[0]
        bad = C.c_uint32()
        assert eng.lib.luda_verify_tables_noalloc(eng.ctx, arr, 1, 0, C.byref(bad), None, 0) == \
            ffi.LUDA_INVALID_ARGUMENT
        arr[0].size = (128 << 20) + 1
        assert eng.lib.luda_verify_tables_noalloc(eng.ctx, arr, 1, ffi.LUDA_JOB_INPUTS_ON_DEVICE, C.byref(bad),
                                                  None, 0) == ffi.LUDA_NOT_SUPPORTED
        eng.free(res1)
    finally:
        eng.close()
