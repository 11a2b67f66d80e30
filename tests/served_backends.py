"""Test helpers for luda_amd.served: a CPU backend on the oracle (host table
bytes), and checks of a finished tree against the write stream."""
import ctypes as C
import time

import oracle
from luda_amd import ffi, served


def small_options(max_file_size=16 << 10):
    o = ffi.default_options()
    o.max_file_size = max_file_size
    return o


class OracleBackend:
    """Flushes and compactions on the oracle.  File handles are addresses of
    host buffers.  With lifo=True the queued forms finish in reverse order of
    submission at each poll, so the drivers see completions out of order."""

    def __init__(self, opts, lifo=True):
        self.opts = opts
        self.flush_opts = ffi.luda_options()
        C.pointer(self.flush_opts)[0] = opts
        self.flush_opts.max_file_size = (1 << 64) - 1
        self.bufs = {}
        self.queue = []
        self.lifo = lifo

    def _files(self, lsm, tables):
        out = []
        owner = served.Owner(None, len(tables))
        for t in tables:
            buf = C.create_string_buffer(t.data, len(t.data))
            addr = C.addressof(buf)
            self.bufs[addr] = buf
            out.append(served.File(lsm.new_number(), len(t.data), t.smallest, t.largest, t.num_entries, addr, owner))
        owner.release = lambda: [self.bufs.pop(f.handle) for f in out]
        if not out:
            owner.release = None
        return out

    def table_bytes(self, f):
        return C.string_at(f.handle, f.size)

    def flush(self, lsm, mem):
        k, ko, v, vo = mem.arrays()
        tables = oracle.builder()(C.byref(self.flush_opts), mem.n, k.ctypes.data, ko.ctypes.data, v.ctypes.data,
                                  vo.ctypes.data)
        return self._files(lsm, tables), time.monotonic() * 1e3

    def compact(self, lsm, spec):
        job, keep = served._job_struct(spec, self.opts)
        job.flags = 0  # host inputs, host outputs
        rc, tables = oracle.compact(job)
        if rc != 0:
            raise RuntimeError(f"oracle compaction failed ({rc})")
        return self._files(lsm, tables), time.monotonic() * 1e3

    def submit_flush(self, lsm, mem, done):
        self.queue.append(lambda: done(*self.flush(lsm, mem), None))

    def submit_compact(self, lsm, spec, done):
        self.queue.append(lambda: done(*self.compact(lsm, spec)))

    def poll(self):
        q, self.queue = self.queue, []
        if self.lifo:
            q.reverse()
        for fn in q:
            fn()
        return len(q)


def expected_contents(memtables):
    """user key -> (sequence, value) of the newest put."""
    want = {}
    for m in memtables:
        for i in range(m.n):
            k = bytes(m.keys[i])
            seq = int.from_bytes(k[-8:], "little") >> 8
            uk = k[:-8]
            if uk not in want or want[uk][0] < seq:
                want[uk] = (seq, bytes(m.vals[i]))
    return want


def tree_contents(lsm, table_bytes):
    """user key -> (sequence, value) of the newest entry the tree holds, with
    every table read back and its block CRCs checked."""
    got = {}
    for level in lsm.levels:
        for f in level:
            rc, entries = oracle.table_entries(table_bytes(f))
            assert rc == 0, f"table {f.number} unreadable ({rc})"
            assert len(entries) == f.entries
            assert entries and entries[0][0] == f.smallest and entries[-1][0] == f.largest
            for k, v in entries:
                seq = int.from_bytes(k[-8:], "little") >> 8
                uk = k[:-8]
                if uk not in got or got[uk][0] < seq:
                    got[uk] = (seq, v)
    return got
