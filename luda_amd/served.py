"""Served compaction: an LSM tree fed by a fillrandom-style write stream, whose
memtable flushes and leveled compactions run on the engine as they arrive.

This is the host side of the served form.  It keeps the tree's shape the way
the reference's VersionSet does and hands every flush and compaction to the
engine through the C-ABI (luda_build_tables / luda_compact one at a time, or
luda_submit_build / luda_submit queued on the worker slots).  Tables stay in
HBM from the flush that writes them to the compaction that deletes them:
every compaction reads its inputs in place (LUDA_JOB_INPUTS_ON_DEVICE) and
leaves its outputs there (LUDA_JOB_KEEP_ON_DEVICE).

Rules followed (reference files):
  * memtable size: MemTable::Add's entry encoding plus one skip-list node,
    against write_buffer_size 4 MiB (db/memtable.cc:76-101,
    include/leveldb/options.h:82);
  * flush: BuildTable of the sorted memtable into one level-0 table
    (db/builder.cc:17-77); flushes install in memtable order;
  * the level a flushed table goes to: 0 unless it overlaps nothing there,
    then down to kMaxMemCompactLevel 2 (Version::PickLevelForMemTableOutput
    db/version_set.cc:491-516, DBImpl::WriteLevel0Table db/db_impl.cc:633-679);
  * the fork's level-0 stop trigger, 40 files, delays flushes
    (db/dbformat.h:26-32, DBImpl::MakeRoomForWrite db/db_impl.cc:1866-1928);
  * scores: level 0 files / 4, level L bytes / (10 MiB * 10^(L-1))
    (VersionSet::Finalize db/version_set.cc:1050-1086, MaxBytesForLevel :42-53);
  * picking: the first file past compact_pointer_[level] (level 0: widened
    to every overlapping level-0 file), boundary inputs, the overlapping files
    of level+1, the level's inputs grown when that takes no more level+1
    files, grandparents of level+2 (PickCompaction / SetupOtherInputs /
    AddBoundaryInputs / GetOverlappingInputs, db/version_set.cc:1270-1464,
    :1380-1403, Version::GetOverlappingInputs :519-555);
  * trivial moves (Compaction::IsTrivialMove :1517-1525, DBImpl::
    BackgroundCompaction db/db_impl.cc:694-711);
  * install: inputs deleted, outputs added at level+1 (db/db_impl.cc:1040-1056).

With an output directory (EngineBackend io_dir) every table is also written
to a file, as DoCompactionWork's outputs are (db/db_impl.cc:1490-1565,
FinishCompactionOutputFile :1000-1014): the plain path writes a job's files
before it takes the next job, the served form hands finished results to a
writer thread whose luda_write_outputs copies them from HBM on the context's
I/O stream, beside the slots' kernels, and a job counts as done (and is
installed) once its files are written; a compaction's inputs are deleted
from the directory when it installs (DeleteObsoleteFiles, db/db_impl.cc:
206-284).

The served form may run compactions of disjoint level pairs at the same time
(the reference runs one at a time, db/db_impl.cc:865-871); a job's inputs,
grandparents and deeper-level ranges are fixed when it is submitted, so its
output is the same as luda_compact's on the same job.
"""
import bisect
import ctypes as C
import os
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import ffi

NUM_LEVELS = 7
L0_COMPACTION_TRIGGER = 4      # db/dbformat.h:26
L0_STOP_WRITES_TRIGGER = 40    # db/dbformat.h:32 (this fork's value)
MAX_MEM_COMPACT_LEVEL = 2      # db/dbformat.h:40
WRITE_BUFFER_SIZE = 4 << 20    # include/leveldb/options.h:82
KEY_LEN = 16
VALUE_LEN = 100
# MemTable::Add: varint32(ikey) + ikey + varint32(value) + value, plus a skip
# list node (key pointer + 4/3 next pointers on average) in the arena.
MEMTABLE_ENTRY_BYTES = 1 + (KEY_LEN + 8) + 1 + VALUE_LEN + 19


def max_bytes_for_level(level, base=10.0 * 1048576.0):
    """MaxBytesForLevel, db/version_set.cc:42-53 (level >= 1); `base` is level
    1's 10 MiB."""
    r = float(base)
    while level > 1:
        r *= 10
        level -= 1
    return r


def ukey(ikey):
    return ikey[:-8]


def ikey_order(ikey):
    """InternalKeyComparator order as a sort key: user key ascending, then
    (sequence, type) descending (db/dbformat.cc:51-67)."""
    return (ikey[:-8], -int.from_bytes(ikey[-8:], "little"))


# ---------------------------------------------------------------------------
# write stream

@dataclass
class Memtable:
    """A sorted memtable: entry i is keys[i] (internal key, 24 bytes) and
    vals[i] (100 bytes), in InternalKeyComparator order."""
    n: int
    keys: np.ndarray      # (n, 24) uint8
    vals: np.ndarray      # (n, 100) uint8
    first_seq: int
    last_seq: int

    @property
    def nbytes(self):
        return int(self.keys.size + self.vals.size)

    def arrays(self):
        """ctypes views for luda_build_tables (kept alive by the caller)."""
        ko = np.arange(self.n + 1, dtype=np.uint64) * np.uint64(self.keys.shape[1])
        vo = np.arange(self.n + 1, dtype=np.uint64) * np.uint64(self.vals.shape[1])
        return self.keys, ko, self.vals, vo


def write_stream(puts, seed, key_space=10 ** 9, entries_per_flush=None):
    """fillrandom: `puts` puts of 16-byte zero-padded decimal keys uniform over
    key_space with 100-byte random values, sequence numbers from 1, cut into
    memtables of write_buffer_size.  Yields Memtable objects."""
    epf = entries_per_flush or WRITE_BUFFER_SIZE // MEMTABLE_ENTRY_BYTES
    rng = np.random.default_rng(seed)
    seq = 1
    left = puts
    while left > 0:
        n = min(epf, left)
        k = rng.integers(0, key_space, n, dtype=np.uint64)
        seqs = np.arange(seq, seq + n, dtype=np.uint64)
        order = np.lexsort((-(seqs.astype(np.int64)), k))
        k = k[order]
        s = seqs[order]
        keys = np.empty((n, KEY_LEN + 8), dtype=np.uint8)
        x = k.copy()
        for j in range(KEY_LEN - 1, -1, -1):
            keys[:, j] = (x % np.uint64(10)).astype(np.uint8) + 48
            x //= np.uint64(10)
        tags = (s << np.uint64(8)) | np.uint64(1)  # kTypeValue
        keys[:, KEY_LEN:] = tags.astype("<u8").view(np.uint8).reshape(n, 8)
        vals = rng.integers(0, 256, (n, VALUE_LEN), dtype=np.uint8)
        yield Memtable(n, keys, vals, seq, seq + n - 1)
        seq += n
        left -= n


# ---------------------------------------------------------------------------
# the tree

_LEVEL_FILE = np.dtype([("sd", "<u8"), ("ss", "<u8"), ("ld", "<u8"), ("ls", "<u8"), ("level", "<i4"),
                        ("reserved", "<u4")])
assert _LEVEL_FILE.itemsize == C.sizeof(ffi.luda_level_file)


def _cbufs(f):
    """C copies of a file's smallest/largest user keys and largest internal
    key, made once per file."""
    cb = getattr(f, "_cb", None)
    if cb is None:
        cb = (C.create_string_buffer(ukey(f.smallest), len(f.smallest) - 8),
              C.create_string_buffer(ukey(f.largest), len(f.largest) - 8),
              C.create_string_buffer(f.largest, len(f.largest)))
        f._cb = cb
    return cb


def _level_row(f, level):
    """f's luda_level_file as a record (f keeps the key buffers alive)."""
    b = _cbufs(f)
    return np.array([(C.addressof(b[0]), len(f.smallest) - 8, C.addressof(b[1]), len(f.largest) - 8, level, 0)],
                    dtype=_LEVEL_FILE)


def _row_bytes(f, level):
    """_level_row's bytes, kept with the file while it stays on one level."""
    rb = getattr(f, "_rb", None)
    if rb is None or rb[0] != level:
        rb = (level, _level_row(f, level).tobytes())
        f._rb = rb
    return rb[1]



@dataclass(eq=False)
class File:
    """FileMetaData (db/version_edit.h:18-27) plus where its bytes are: a
    backend handle (device pointer, or host bytes for a CPU backend) and the
    result that owns them."""
    number: int
    size: int
    smallest: bytes
    largest: bytes
    entries: int
    handle: object
    owner: object = None


@dataclass(eq=False)
class JobSpec:
    """One compaction, fixed when it is picked."""
    level: int
    inputs: tuple            # ([File] of level, [File] of level + 1)
    grandparents: list       # [File] of level + 2 overlapping the inputs
    deeper_files: list       # [(level, [File])]: files of levels >= level + 2 that overlap the job
    smallest_snapshot: int
    arrival: float = 0.0
    bytes_in: int = 0
    deeper_rows: object = None  # the same files as luda_level_file rows (a numpy record array)

    @property
    def deeper(self):
        """[(level, File)] in level order."""
        return [(d, f) for d, fs in self.deeper_files for f in fs]


class Owner:
    """Reference count over the files one result holds; the backend's release
    runs when the last of them is deleted."""

    def __init__(self, release, nfiles):
        self.release = release
        self.refs = nfiles

    def drop(self):
        self.refs -= 1
        if self.refs == 0 and self.release is not None:
            self.release()
            self.release = None


class Lsm:
    """The levels of one DB (Version::files_) and the compaction picker."""

    def __init__(self, max_file_size=2 << 20, level1_bytes=10 << 20):
        self.level1_bytes = level1_bytes
        self.levels = [[] for _ in range(NUM_LEVELS)]
        self.compact_pointer = [None] * NUM_LEVELS
        self.busy = set()          # levels a running compaction reads or writes
        self.on_delete = None      # called with each file a compaction deletes
        self.next_number = 2
        self.last_seq = 0
        self.max_file_size = max_file_size
        self.moves = 0
        # per level, kept in step with levels[]: smallest and largest user
        # keys in level order, and the level's bytes
        self._sk = [[] for _ in range(NUM_LEVELS)]
        self._lk = [[] for _ in range(NUM_LEVELS)]
        self._bytes = [0] * NUM_LEVELS
        # and the levels' luda_level_file rows for the deeper-level ranges of
        # jobs, rebuilt when a changed level is next asked for
        self._rows = [np.zeros(0, dtype=_LEVEL_FILE) for _ in range(NUM_LEVELS)]
        self._dirty = set()

    def level_rows(self, level):
        if level in self._dirty:
            self._rows[level] = np.frombuffer(b"".join(_row_bytes(f, level) for f in self.levels[level]),
                                              dtype=_LEVEL_FILE)
            self._dirty.discard(level)
        return self._rows[level]

    def _remove(self, level, f):
        i = self.levels[level].index(f)
        del self.levels[level][i]
        del self._sk[level][i]
        del self._lk[level][i]
        self._dirty.add(level)
        self._bytes[level] -= f.size

    def new_number(self):
        self.next_number += 1
        return self.next_number - 1

    # -- Version::GetOverlappingInputs, db/version_set.cc:519-555
    def overlapping(self, level, begin, end):
        files = self.levels[level]
        if level > 0:
            i, j = self.overlap_range(level, begin, end)
            return files[i:j]
        out = []
        i = 0
        while i < len(files):
            f = files[i]
            i += 1
            fs, fl = ukey(f.smallest), ukey(f.largest)
            if (begin is not None and fl < begin) or (end is not None and fs > end):
                continue
            out.append(f)
            if level == 0:
                # level-0 files overlap each other: widen and start over
                if begin is not None and fs < begin:
                    begin, out, i = fs, [], 0
                elif end is not None and fl > end:
                    end, out, i = fl, [], 0
        return out

    def overlap_range(self, level, begin, end):
        """Level >= 1 is sorted, and disjoint but for a shared boundary user
        key: the files overlapping [begin, end] are one run [i, j)."""
        sk, lk = self._sk[level], self._lk[level]
        i = 0 if begin is None else bisect.bisect_left(lk, begin)
        j = len(sk) if end is None else bisect.bisect_right(sk, end)
        return i, max(i, j)

    @staticmethod
    def key_range(files):
        lo = min((f.smallest for f in files), key=ikey_order)
        hi = max((f.largest for f in files), key=ikey_order)
        return lo, hi

    def level_bytes(self, level):
        return self._bytes[level]

    def score(self, level):
        if level == 0:
            return len(self.levels[0]) / L0_COMPACTION_TRIGGER
        return self.level_bytes(level) / max_bytes_for_level(level, self.level1_bytes)

    def needs(self):
        """Levels whose score is at least 1 (VersionSet::Finalize)."""
        return [L for L in range(NUM_LEVELS - 1) if self.score(L) >= 1]

    # -- AddBoundaryInputs, db/version_set.cc:1380-1403
    def add_boundary(self, level, picked):
        files = self.levels[level]
        if not picked:
            return picked
        largest = max((f.largest for f in picked), key=ikey_order)
        if level > 0:
            # the smallest file past `largest` that starts on its user key is
            # the next file in level order
            sk = self._sk[level]
            j = bisect.bisect_left(sk, ukey(largest))
            while j < len(files):
                f = files[j]
                j += 1
                if f in picked or ikey_order(f.smallest) <= ikey_order(largest):
                    continue
                if ukey(f.smallest) != ukey(largest):
                    break
                picked.append(f)
                largest = f.largest
            return picked
        while True:
            cand = None
            for f in files:
                if f in picked:
                    continue
                if ikey_order(f.smallest) > ikey_order(largest) and ukey(f.smallest) == ukey(largest):
                    if cand is None or ikey_order(f.smallest) < ikey_order(cand.smallest):
                        cand = f
            if cand is None:
                return picked
            picked.append(cand)
            largest = cand.largest

    def pick(self):
        """The highest-scoring level with a score >= 1 whose level pair is not
        in use; returns a JobSpec, a trivially moved file (installed here), or
        None."""
        # VersionSet::Finalize (db/version_set.cc:1050-1086) keeps the first
        # level with the strictly highest score; PickCompaction runs it only
        # when that score is >= 1 (:1276-1277).
        best, best_score = None, -1.0
        for L in range(NUM_LEVELS - 1):
            if L in self.busy or L + 1 in self.busy:
                continue
            s = self.score(L)
            if s > best_score:
                best, best_score = L, s
        if best is None or best_score < 1.0:
            return None
        L = best
        files = self.levels[L]
        first = None
        ptr = self.compact_pointer[L]
        i = 0
        if ptr is not None and L > 0:
            i = bisect.bisect_left(self._lk[L], ukey(ptr))
        for f in files[i:]:
            if ptr is None or ikey_order(f.largest) > ikey_order(ptr):
                first = f
                break
        if first is None:
            first = files[0]
        inputs0 = [first]
        if L == 0:
            inputs0 = self.overlapping(0, ukey(first.smallest), ukey(first.largest))
        # SetupOtherInputs (db/version_set.cc:1405-1464)
        inputs0 = self.add_boundary(L, inputs0)
        lo, hi = self.key_range(inputs0)
        inputs1 = self.overlapping(L + 1, ukey(lo), ukey(hi))
        alo, ahi = self.key_range(inputs0 + inputs1)
        if inputs1:
            # grow the level's inputs when that takes in no more level+1 files
            # and stays under ExpandedCompactionByteSizeLimit (:37-39)
            exp0 = self.add_boundary(L, self.overlapping(L, ukey(alo), ukey(ahi)))
            size1 = sum(f.size for f in inputs1)
            if len(exp0) > len(inputs0) and size1 + sum(f.size for f in exp0) < 25 * self.max_file_size:
                nlo, nhi = self.key_range(exp0)
                exp1 = self.overlapping(L + 1, ukey(nlo), ukey(nhi))
                if len(exp1) == len(inputs1):
                    lo, hi = nlo, nhi
                    inputs0, inputs1 = exp0, exp1
                    alo, ahi = self.key_range(inputs0 + inputs1)
        gps = self.overlapping(L + 2, ukey(alo), ukey(ahi)) if L + 2 < NUM_LEVELS else []
        self.compact_pointer[L] = hi
        if len(inputs0) == 1 and not inputs1 and sum(f.size for f in gps) <= 10 * self.max_file_size:
            # trivial move (Compaction::IsTrivialMove, db/version_set.cc:1517-1525)
            f = inputs0[0]
            self._remove(L, f)
            self.add_file(L + 1, f)
            self.moves += 1
            return f
        # IsBaseLevelForKey (db/version_set.cc:1535-1554) is only asked about
        # the job's own keys, so deeper files outside the job's key range can
        # never answer it: they are left out
        dfiles, drows = [], []
        for d in range(L + 2, NUM_LEVELS):
            i, j = self.overlap_range(d, ukey(alo), ukey(ahi))
            if j > i:
                dfiles.append((d, self.levels[d][i:j]))
                drows.append(self.level_rows(d)[i:j])
        order0 = sorted(inputs0, key=lambda f: ikey_order(f.smallest))
        order1 = sorted(inputs1, key=lambda f: ikey_order(f.smallest))
        spec = JobSpec(L, (order0, order1), gps, dfiles, self.last_seq,
                       deeper_rows=np.concatenate(drows) if drows else np.zeros(0, dtype=_LEVEL_FILE))
        spec.bytes_in = sum(f.size for f in order0 + order1)
        self.busy.update((L, L + 1))
        return spec

    def flush_level(self, f):
        """Version::PickLevelForMemTableOutput (db/version_set.cc:491-516) for
        a flushed table, with the levels a running compaction uses left alone."""
        lo, hi = ukey(f.smallest), ukey(f.largest)
        level = 0
        if self.overlapping(0, lo, hi) or 0 in self.busy:
            return 0
        while level < MAX_MEM_COMPACT_LEVEL:
            if level + 1 in self.busy or self.overlapping(level + 1, lo, hi):
                break
            if level + 2 < NUM_LEVELS:
                if sum(x.size for x in self.overlapping(level + 2, lo, hi)) > 10 * self.max_file_size:
                    break
            level += 1
        return level

    def add_file(self, level, f):
        files = self.levels[level]
        if level > 0:
            sk = self._sk[level]
            i = bisect.bisect_left(sk, ukey(f.smallest))
            while i < len(files) and ikey_order(files[i].smallest) <= ikey_order(f.smallest):
                i += 1
        else:
            keys = [ikey_order(x.smallest) for x in files]
            i = bisect.bisect_right(keys, ikey_order(f.smallest))
        files.insert(i, f)
        self._sk[level].insert(i, ukey(f.smallest))
        self._lk[level].insert(i, ukey(f.largest))
        self._dirty.add(level)
        self._bytes[level] += f.size

    def install(self, spec, outputs):
        """InstallCompactionResults (db/db_impl.cc:1040-1056)."""
        for which in (0, 1):
            for f in spec.inputs[which]:
                self._remove(spec.level + which, f)
                if self.on_delete is not None:
                    self.on_delete(f)
        for f in outputs:
            self.add_file(spec.level + 1, f)
        for which in (0, 1):
            for f in spec.inputs[which]:
                if f.owner is not None:
                    f.owner.drop()
        self.busy.discard(spec.level)
        self.busy.discard(spec.level + 1)

    def check(self):
        """Levels >= 1 are sorted and their files disjoint (neighbours may
        share only a boundary user key, newer versions first)."""
        for L in range(1, NUM_LEVELS):
            fs = self.levels[L]
            for a, b in zip(fs, fs[1:]):
                if not (ukey(a.largest) <= ukey(b.smallest) and ikey_order(a.largest) < ikey_order(b.smallest)):
                    raise AssertionError(f"level {L} files overlap or are out of order")


# ---------------------------------------------------------------------------
# the engine as the backend: tables in HBM

def _job_struct(spec, opts):
    """luda_job for a JobSpec whose file handles are device pointers."""
    keep = []
    job = ffi.luda_job()
    job.level = spec.level
    job.flags = ffi.LUDA_JOB_INPUTS_ON_DEVICE | ffi.LUDA_JOB_KEEP_ON_DEVICE
    job.smallest_snapshot = spec.smallest_snapshot
    job.options = opts
    for which in (0, 1):
        fs = spec.inputs[which]
        arr = (ffi.luda_file * max(1, len(fs)))()
        for i, f in enumerate(fs):
            arr[i].data = C.c_void_p(f.handle)
            arr[i].size = f.size
            arr[i].number = f.number
        keep.append(arr)
        job.inputs[which] = C.cast(arr, C.POINTER(ffi.luda_file))
        job.num_inputs[which] = len(fs)
    gp = (ffi.luda_grandparent * max(1, len(spec.grandparents)))()
    for i, f in enumerate(spec.grandparents):
        b = _cbufs(f)[2]
        keep.append(f)
        gp[i].largest.data = C.cast(b, C.c_void_p)
        gp[i].largest.size = len(f.largest)
        gp[i].file_size = f.size
    keep.append(gp)
    job.grandparents = C.cast(gp, C.POINTER(ffi.luda_grandparent))
    job.num_grandparents = len(spec.grandparents)
    if spec.deeper_rows is not None:
        rows = spec.deeper_rows
    else:
        rows = np.zeros(len(spec.deeper), dtype=_LEVEL_FILE)
        for i, (lvl, f) in enumerate(spec.deeper):
            rows[i] = _level_row(f, lvl)
    keep.append(spec.deeper_files)  # the key buffers the rows point into
    rows = np.ascontiguousarray(rows) if len(rows) else np.zeros(1, dtype=_LEVEL_FILE)
    keep.append(rows)
    job.deeper = C.cast(C.c_void_p(rows.ctypes.data), C.POINTER(ffi.luda_level_file))
    job.num_deeper = len(spec.deeper)
    return job, keep


class EngineBackend:
    """Flushes and compactions on a luda context; tables stay in HBM, and with
    `resident` their decoded entries too (LUDA_JOB_KEEP_DECODED), so that a
    compaction of tables the engine wrote decodes nothing."""

    def __init__(self, engine, opts, resident=True, io_dir=None, io_sync=False, io_threads=1):
        self.eng = engine
        self.opts = opts
        self.resident = resident
        self.flush_opts = ffi.luda_options()
        C.pointer(self.flush_opts)[0] = opts
        self.flush_opts.max_file_size = (1 << 64) - 1  # BuildTable writes one table
        self.acct = {"jobs": 0, "device_ms": 0.0, "call_ms": 0.0}  # summed over the engine's calls
        self.io_dir = io_dir    # tables also written to files here (None: HBM only)
        self.io_sync = io_sync  # fsync each file, as FinishCompactionOutputFile does
        self.io = {"files": 0, "bytes": 0, "write_ms": 0.0, "deleted": 0}
        # queued jobs' writer threads, each on a context of its own (a
        # context's I/O staging buffers serve one write at a time)
        self.io_threads = io_threads
        self._wq = self._doneq = None
        self._writers = []
        self._io_lock = threading.Lock()

    # -- output files
    def path_of(self, f):
        return os.path.join(self.io_dir, "%06d.ldb" % f.number)

    def _write(self, res, files, eng=None):
        t = time.perf_counter()
        rc, msg = (eng or self.eng).write_outputs(res, [self.path_of(f) for f in files], sync=self.io_sync)
        if rc != ffi.LUDA_OK:
            raise RuntimeError(f"luda_write_outputs failed ({rc}): {msg}")
        with self._io_lock:
            self.io["files"] += len(files)
            self.io["bytes"] += sum(f.size for f in files)
            self.io["write_ms"] += (time.perf_counter() - t) * 1e3

    def delete_file(self, f):
        """Lsm.on_delete: a compaction's input leaves the directory."""
        if self.io_dir is not None:
            os.unlink(self.path_of(f))
            self.io["deleted"] += 1

    def _writer_loop(self, eng):
        while True:
            item = self._wq.get()
            if item is None:
                return
            res, files, finish = item
            try:
                self._write(res, files, eng)
                self._doneq.put((finish, None, _now_ms()))
            except Exception as e:  # re-raised by poll on the driving thread
                self._doneq.put((finish, e, None))

    def _finish_async(self, res, files, finish):
        """Queued jobs: finish(done_ms) runs from poll once the files are
        written (at once without an output directory)."""
        if self.io_dir is None or not files:
            finish(None)
            return
        if not self._writers:
            from . import jobs
            self._wq, self._doneq = queue.Queue(), queue.Queue()
            for i in range(max(1, self.io_threads)):
                eng = self.eng if i == 0 else jobs.Engine(self.eng.device)
                t = threading.Thread(target=self._writer_loop, args=(eng,), daemon=True)
                t.start()
                self._writers.append((t, eng))
        self._wq.put((res, files, finish))

    def close(self):
        for _ in self._writers:
            self._wq.put(None)
        for t, eng in self._writers:
            t.join()
            if eng is not self.eng:
                eng.close()
        self._writers = []

    def _account(self, stats):
        self.acct["jobs"] += 1
        self.acct["device_ms"] += stats["device_ms"]
        self.acct["call_ms"] += stats["total_ms"]

    def _files(self, lsm, res):
        r = res.contents
        if r.num_outputs == 0:
            self.eng.free(res)
            return []
        keys = C.string_at(r.keys, sum(r.outputs[i].smallest_size + r.outputs[i].largest_size
                                       for i in range(r.num_outputs)))
        owner = Owner(lambda: self.eng.free(res), r.num_outputs)
        out = []
        for i in range(r.num_outputs):
            o = r.outputs[i]
            out.append(File(lsm.new_number(), o.file_size,
                            keys[o.smallest_offset:o.smallest_offset + o.smallest_size],
                            keys[o.largest_offset:o.largest_offset + o.largest_size],
                            o.num_entries, r.device_data + o.offset, owner))
        return out

    def table_bytes(self, f):
        return self.eng.d2h(f.handle, f.size)

    # -- one at a time
    def flush(self, lsm, mem):
        k, ko, v, vo = mem.arrays()
        rc, out, msg = self.eng.build_tables(C.byref(self.flush_opts), mem.n, k.ctypes.data, ko.ctypes.data,
                                             v.ctypes.data, vo.ctypes.data, keep_on_device=True,
                                             keep_decoded=self.resident)
        if rc != ffi.LUDA_OK:
            raise RuntimeError(f"luda_build_tables failed ({rc}): {msg}")
        self._account(out[1])
        files = self._files(lsm, out[2])
        if self.io_dir is not None and files:
            self._write(out[2], files)
            return files, _now_ms()
        return files, out[1]["end_clock_ms"]

    def compact(self, lsm, spec):
        job, keep = _job_struct(spec, self.opts)
        rc, out, msg = self.eng.compact(job, keep_on_device=True, keep_decoded=self.resident)
        if rc != ffi.LUDA_OK:
            raise RuntimeError(f"luda_compact failed ({rc}): {msg}")
        self._account(out[1])
        files = self._files(lsm, out[2])
        if self.io_dir is not None and files:
            self._write(out[2], files)
            return files, _now_ms()
        return files, out[1]["end_clock_ms"]

    # -- queued on the worker slots
    def submit_flush(self, lsm, mem, done):
        k, ko, v, vo = mem.arrays()
        hold = (k, ko, v, vo)

        def cb(rc, out, msg):
            if rc != ffi.LUDA_OK:
                raise RuntimeError(f"queued flush failed ({rc}): {msg}")
            self._account(out[1])
            files, end = self._files(lsm, out[2]), out[1]["end_clock_ms"]
            self._finish_async(out[2], files, lambda t: done(files, t or end, hold))
        self.eng.submit_build(self.flush_opts, mem.n, k.ctypes.data, ko.ctypes.data, v.ctypes.data, vo.ctypes.data,
                              cb, keep_on_device=True, keep_decoded=self.resident)

    def submit_compact(self, lsm, spec, done):
        job, keep = _job_struct(spec, self.opts)
        spec._job = (job, keep)

        def cb(rc, out, msg):
            if rc != ffi.LUDA_OK:
                raise RuntimeError(f"queued compaction failed ({rc}): {msg}")
            self._account(out[1])
            files, end = self._files(lsm, out[2]), out[1]["end_clock_ms"]
            self._finish_async(out[2], files, lambda t: done(files, t or end))
        self.eng.submit(job, cb, keep_on_device=True, keep_decoded=self.resident)

    def poll(self):
        n = self.eng.poll()
        while self._doneq is not None and not self._doneq.empty():
            finish, err, done_ms = self._doneq.get()
            if err is not None:
                raise err
            finish(done_ms)
        return n


# ---------------------------------------------------------------------------
# drivers

@dataclass
class RunStats:
    jobs: list = field(default_factory=list)   # (kind, level, arrival_ms, done_ms, bytes_in, bytes_out)
    t0_ms: float = 0.0
    end_ms: float = 0.0

    def summary(self):
        lat = sorted(d - a for _, _, a, d, _, _ in self.jobs)
        span_s = max(1e-9, (self.end_ms - self.t0_ms) / 1e3)
        bin_ = sum(j[4] for j in self.jobs)

        def pct(p):
            if not lat:
                return 0.0
            return lat[min(len(lat) - 1, int(round(p / 100.0 * (len(lat) - 1))))]
        return {
            "jobs": len(self.jobs),
            "flushes": sum(1 for j in self.jobs if j[0] == "flush"),
            "compactions": sum(1 for j in self.jobs if j[0] == "compact"),
            "input_bytes": bin_,
            "output_bytes": sum(j[5] for j in self.jobs),
            "seconds": round(span_s, 4),
            "gbps": round(bin_ / span_s / 1e9, 3),
            "p50_ms": round(pct(50), 3),
            "p99_ms": round(pct(99), 3),
        }


def _now_ms():
    return time.monotonic() * 1e3


def _note_needs(lsm, need_since, now):
    for L in lsm.needs():
        if need_since.get(L) is None:
            need_since[L] = now


def run_plain(backend, memtables, interval_ms, lsm=None, on_job=None):
    """on_job(spec, outputs) runs for every compaction before it is installed
    (while its input files are still held).

    One job at a time on the calling thread, as the reference's single
    background thread does: a flush that has arrived goes first
    (DBImpl::BackgroundCompaction, db/db_impl.cc:683-687), then one
    compaction.  Flush i arrives at i * interval_ms."""
    lsm = lsm or Lsm()
    st = RunStats()
    st.t0_ms = t0 = _now_ms()
    need_since = {}
    nxt = 0
    while True:
        now = _now_ms()
        if nxt < len(memtables) and t0 + nxt * interval_ms <= now and len(lsm.levels[0]) < L0_STOP_WRITES_TRIGGER:
            mem = memtables[nxt]
            files, done = backend.flush(lsm, mem)
            for f in files:
                lsm.add_file(lsm.flush_level(f), f)
            lsm.last_seq = mem.last_seq
            st.jobs.append(("flush", -1, t0 + nxt * interval_ms, done, mem.nbytes, sum(f.size for f in files)))
            nxt += 1
            _note_needs(lsm, need_since, done)
            continue
        spec = lsm.pick()
        if isinstance(spec, File):
            _note_needs(lsm, need_since, _now_ms())
            continue
        if spec is not None:
            spec.arrival = need_since.pop(spec.level, None) or _now_ms()
            files, done = backend.compact(lsm, spec)
            if on_job:
                on_job(spec, files)
            lsm.install(spec, files)
            st.jobs.append(("compact", spec.level, spec.arrival, done, spec.bytes_in, sum(f.size for f in files)))
            _note_needs(lsm, need_since, done)
            continue
        if nxt >= len(memtables):
            break
        wait = t0 + nxt * interval_ms - _now_ms()
        if wait > 0:
            time.sleep(min(wait, 1.0) / 1e3)
    st.end_ms = max([st.t0_ms] + [j[3] for j in st.jobs])
    return lsm, st


def run_served(backend, memtables, interval_ms, lsm=None, on_job=None, max_compactions=3):
    """Flushes are queued the moment they arrive; compactions of disjoint
    level pairs are queued as soon as the tree needs them; the engine admits
    queued jobs as its slots and HBM budget allow, so their work overlaps."""
    lsm = lsm or Lsm()
    st = RunStats()
    st.t0_ms = t0 = _now_ms()
    need_since = {}
    nxt = 0
    flushes_out = 0          # queued, not yet installed
    pending_install = {}     # memtable index -> (files, done_ms) waiting for older flushes
    next_install = 0
    running = [0]

    def flush_done(i, arrival, mem):
        def done(files, done_ms, _hold):
            pending_install[i] = (files, done_ms, arrival, mem)
        return done

    def compact_done(spec):
        def done(files, done_ms):
            running[0] -= 1
            if on_job:
                on_job(spec, files)
            lsm.install(spec, files)
            st.jobs.append(("compact", spec.level, spec.arrival, done_ms, spec.bytes_in, sum(f.size for f in files)))
            _note_needs(lsm, need_since, done_ms)
        return done

    while True:
        now = _now_ms()
        while (nxt < len(memtables) and t0 + nxt * interval_ms <= now
               and len(lsm.levels[0]) + flushes_out < L0_STOP_WRITES_TRIGGER):
            arrival = t0 + nxt * interval_ms
            backend.submit_flush(lsm, memtables[nxt], flush_done(nxt, arrival, memtables[nxt]))
            flushes_out += 1
            nxt += 1
        backend.poll()
        while next_install in pending_install:  # flushes install in memtable order
            files, done_ms, arrival, mem = pending_install.pop(next_install)
            for f in files:
                lsm.add_file(lsm.flush_level(f), f)
            lsm.last_seq = mem.last_seq
            st.jobs.append(("flush", -1, arrival, done_ms, mem.nbytes, sum(f.size for f in files)))
            next_install += 1
            flushes_out -= 1
            _note_needs(lsm, need_since, done_ms)
        while running[0] < max_compactions:
            spec = lsm.pick()
            if spec is None:
                break
            if isinstance(spec, File):
                _note_needs(lsm, need_since, _now_ms())
                continue
            spec.arrival = need_since.pop(spec.level, None) or _now_ms()
            running[0] += 1
            backend.submit_compact(lsm, spec, compact_done(spec))
        if nxt >= len(memtables) and flushes_out == 0 and running[0] == 0 and not lsm.needs():
            break
        time.sleep(0.0001)
    st.end_ms = max([st.t0_ms] + [j[3] for j in st.jobs])
    return lsm, st
