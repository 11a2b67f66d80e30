"""luda_verify_tables on the GPU: tables the engine wrote pass; a flipped
byte in a data block, in the filter block or in the index block's stored CRC
is reported as corruption of that table, as the oracle's checksummed reader
also reports it."""
import pytest

import oracle
from luda_amd import ffi, jobs

pytestmark = pytest.mark.gpu


def varint(b, i):
    r = s = 0
    while True:
        c = b[i]
        i += 1
        r |= (c & 127) << s
        if c < 128:
            return r, i
        s += 7


def handles(t):
    """(metaindex, index, filter) handles (offset, size) of a table image."""
    i = len(t) - 48
    moff, i = varint(t, i)
    msize, i = varint(t, i)
    ioff, i = varint(t, i)
    isize, i = varint(t, i)
    j = moff + 1  # shared = 0
    ns, j = varint(t, j)
    vl, j = varint(t, j)
    assert t[j:j + ns] == b"filter.leveldb.BuiltinBloomFilter2"
    foff, k = varint(t, j + ns)
    fsize, k = varint(t, k)
    return (moff, msize), (ioff, isize), (foff, fsize)


@pytest.fixture(scope="module")
def tables():
    eng = jobs.Engine(0)
    opts = ffi.default_options(max_file_size=64 << 10)  # several output tables
    wl = jobs.Workload(0, scale=0.03)
    ins = ([], [])
    for g in wl.groups:
        ins[g.which].extend(t.data for t in jobs.build_group_tables(eng.builder(), opts, g))
    job = jobs.make_job(wl, opts, ins)
    rc, out, msg = eng.compact(job.job)
    assert rc == 0, msg
    yield eng, [t.data for t in out[0]], ins[0] + ins[1]
    eng.close()


def test_engine_tables_pass(tables):
    eng, outs, ins = tables
    assert len(outs) >= 2
    for ts in (outs, ins):
        rc, bad, msg = eng.verify(ts)
        assert rc == ffi.LUDA_OK and bad is None, msg


def flipped(t, at):
    b = bytearray(t)
    b[at] ^= 0x5a
    return bytes(b)


@pytest.mark.parametrize("where", ["data", "filter", "index_crc", "metaindex"])
def test_flipped_byte_is_found(tables, where):
    eng, outs, _ = tables
    k = len(outs) - 1 if where == "filter" else 1
    t = outs[k]
    (moff, msize), (ioff, isize), (foff, fsize) = handles(t)
    at = {"data": 50, "filter": foff + 1, "index_crc": ioff + isize + 1, "metaindex": moff + msize + 2}[where]
    bad_t = flipped(t, at)
    print("verify", where, k, at, flush=True)
    rc, bad, msg = eng.verify(outs[:k] + [bad_t] + outs[k + 1:])
    assert rc == ffi.LUDA_CORRUPTION and bad == k, (rc, bad, msg)
    rc_cpu, _ = oracle.table_entries(bad_t, verify_crc=True)
    if where != "filter":
        assert rc_cpu != 0  # the oracle's reader checks data, index and metaindex block CRCs too


def test_device_resident_tables(tables):
    eng, outs, _ = tables
    ptrs = []
    try:
        for t in outs:
            p = eng.device_alloc(len(t))
            eng.h2d(p, t)
            ptrs.append((p, len(t)))
        rc, bad, msg = eng.verify(ptrs, on_device=True)
        assert rc == ffi.LUDA_OK and bad is None, msg
    finally:
        for p, _ in ptrs:
            eng.device_free(p)
