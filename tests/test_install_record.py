"""The install record: VersionEdit serialisation (luda_encode_edit,
luda_encode_compaction_edit) and MANIFEST log framing (luda_log_record)
against the reference's own VersionEdit::EncodeTo and log::Writer
(oracle/_ref).  Host code only: runs without a GPU."""
import ctypes as C
import random
import struct

import pytest

import oracle
from luda_amd import ffi
from tests.tiny_jobs import TinyJob, ikey

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")


def _keep(keep, b):
    buf = C.create_string_buffer(b, len(b))
    keep.append(buf)
    return ffi.luda_slice(C.cast(buf, C.c_void_p), len(b))


def random_edit(rng, keep):
    e = ffi.luda_edit()
    e.has = rng.randrange(32)
    e.comparator = _keep(keep, b"leveldb.BytewiseComparator")
    e.log_number, e.prev_log_number = rng.randrange(1 << 40), rng.randrange(1 << 20)
    e.next_file_number, e.last_sequence = rng.randrange(1 << 30), rng.randrange(1 << 56)
    cps = (ffi.luda_edit_pointer * 4)()
    for i in range(4):
        cps[i].level = rng.randrange(7)
        cps[i].key = _keep(keep, ikey(bytes(rng.randrange(256) for _ in range(rng.randint(0, 40))),
                                      rng.randrange(1 << 56), rng.randrange(2)))
    e.compact_pointers, e.num_compact_pointers = cps, rng.randint(0, 4)
    dl = (ffi.luda_edit_deleted * 30)()
    for i in range(30):  # repeats and any order: the edit keeps a sorted set
        dl[i].level, dl[i].number = rng.randrange(7), rng.choice([rng.randrange(50), rng.randrange(1 << 40)])
    e.deleted, e.num_deleted = dl, rng.randint(0, 30)
    nf = (ffi.luda_edit_file * 12)()
    for i in range(12):
        nf[i].level, nf[i].number, nf[i].file_size = rng.randrange(7), rng.randrange(1 << 40), rng.randrange(1 << 33)
        nf[i].smallest = _keep(keep, ikey(b"s%d" % rng.randrange(10 ** 9), rng.randrange(1 << 56), 1))
        nf[i].largest = _keep(keep, ikey(b"z" * rng.randint(0, 300), rng.randrange(1 << 56), 0))
    e.new_files, e.num_new_files = nf, rng.randint(0, 12)
    keep += [cps, dl, nf]
    return e


def encode(fn, *args):
    n = C.c_uint64()
    buf = C.create_string_buffer(1 << 20)
    rc = fn(*args, buf, 1 << 20, C.byref(n))
    assert rc == 0
    return buf.raw[:n.value]


def test_edit_bytes_equal_reference():
    lib, ref = ffi.engine(), oracle.load_ref()
    rng = random.Random(7)
    for _ in range(300):
        keep = []
        e = random_edit(rng, keep)
        assert encode(lib.luda_encode_edit, C.byref(e)) == encode(ref.ref_encode_edit, C.byref(e))


def test_edit_reports_the_size_it_needs():
    lib = ffi.engine()
    keep = []
    e = random_edit(random.Random(3), keep)
    e.num_new_files = 5
    full = encode(lib.luda_encode_edit, C.byref(e))
    n = C.c_uint64()
    small = C.create_string_buffer(8)
    assert lib.luda_encode_edit(C.byref(e), small, 8, C.byref(n)) == ffi.LUDA_INVALID_ARGUMENT
    assert n.value == len(full)


def result_of(tables):
    """A luda_result for table metadata (data not needed for the edit)."""
    keys = b"".join(t.smallest + t.largest for t in tables)
    kb = C.create_string_buffer(keys, max(1, len(keys)))
    outs = (ffi.luda_output * max(1, len(tables)))()
    off = 0
    for i, t in enumerate(tables):
        outs[i].file_size, outs[i].num_entries = len(t.data), t.num_entries
        outs[i].smallest_offset, outs[i].smallest_size = off, len(t.smallest)
        off += len(t.smallest)
        outs[i].largest_offset, outs[i].largest_size = off, len(t.largest)
        off += len(t.largest)
    r = ffi.luda_result()
    r.num_outputs = len(tables)
    r.outputs = C.cast(outs, C.POINTER(ffi.luda_output))
    r.keys = C.cast(kb, C.POINTER(C.c_uint8))
    return r, (kb, outs)


@pytest.mark.parametrize("seed", range(40))
def test_compaction_edit_equals_reference_install(seed):
    """InstallCompactionResults' edit for a compaction the oracle ran: inputs
    deleted from level and level + 1, outputs added at level + 1 with caller
    file numbers, plus LogAndApply's header fields and a compact pointer."""
    lib, ref = ffi.engine(), oracle.load_ref()
    tj = TinyJob(seed)
    job, keep = tj.make_job(lambda e, o: [t.data for t in oracle.build_tables(e, o, "oracle")])
    for w in (0, 1):
        for i in range(job.num_inputs[w]):
            job.inputs[w][i].number = 100 * (w + 1) + 3 * i + 7
    rc, tables = oracle.compact(job, "oracle")
    assert rc == 0
    res, hold = result_of(tables)
    numbers = (C.c_uint64 * max(1, len(tables)))(*[500 + 2 * i for i in range(len(tables))])
    base = ffi.luda_edit()
    base.has = ffi.LUDA_EDIT_HAS_LOG_NUMBER | ffi.LUDA_EDIT_HAS_NEXT_FILE | ffi.LUDA_EDIT_HAS_LAST_SEQUENCE
    base.log_number, base.next_file_number, base.last_sequence = 12, 900, 123456
    cp = (ffi.luda_edit_pointer * 1)()
    cp[0].level = tj.level
    cp[0].key = _keep(keep, ikey(b"zz", 77, 1))
    base.compact_pointers, base.num_compact_pointers = cp, 1
    got = encode(lib.luda_encode_compaction_edit, C.byref(job), C.byref(res), numbers, C.byref(base))
    # the same edit through VersionEdit's setters in the reference
    want = ffi.luda_edit()
    C.pointer(want)[0] = base
    dels = [(tj.level + w, job.inputs[w][i].number) for w in (0, 1) for i in range(job.num_inputs[w])]
    dl = (ffi.luda_edit_deleted * max(1, len(dels)))()
    for i, (lv, num) in enumerate(dels):
        dl[i].level, dl[i].number = lv, num
    nf = (ffi.luda_edit_file * max(1, len(tables)))()
    for i, t in enumerate(tables):
        nf[i].level, nf[i].number, nf[i].file_size = tj.level + 1, numbers[i], len(t.data)
        nf[i].smallest, nf[i].largest = _keep(keep, t.smallest), _keep(keep, t.largest)
    want.deleted, want.num_deleted = dl, len(dels)
    want.new_files, want.num_new_files = nf, len(tables)
    assert got == encode(ref.ref_encode_edit, C.byref(want))


@pytest.mark.parametrize("offset", [0, 1, 7, 32000, 32760, 32761, 32762, 32767])
def test_log_records_equal_reference(offset):
    """Records of 0 bytes to three blocks, starting anywhere in a block
    (trailer fill when under 7 bytes are left, FIRST/MIDDLE/LAST fragments)."""
    lib, ref = ffi.engine(), oracle.load_ref()
    rng = random.Random(offset)
    for n in [0, 1, 100, 32761, 32762, 70000, 100000]:
        payload = bytes(rng.randrange(256) for _ in range(n))
        outs = []
        for fn in (lib.luda_log_record, ref.ref_log_record):
            bo = C.c_uint64(offset)
            m = C.c_uint64()
            buf = C.create_string_buffer(n + 1000)
            assert fn(C.byref(bo), payload, n, buf, n + 1000, C.byref(m)) == 0
            outs.append((buf.raw[:m.value], bo.value))
        assert outs[0] == outs[1], n
