"""Python-side plumbing around the C-ABI: seeded workloads, input tables and
luda_job assembly, and reading luda_result back.

This mirrors how DBImpl::DoCompactionWork's caller would fill a luda_job from
a leveldb::Compaction (inputs_[0..1] in Compaction order, grandparents_,
Version::files_ of the deeper levels, smallest_snapshot); see INTEGRATION.md
for the C++ side.  No table bytes are produced here: a table builder (the
engine's luda_build_tables, or a CPU builder passed in by tests) does that.
"""
import ctypes as C
from dataclasses import dataclass, field

from . import ffi


@dataclass
class Table:
    data: bytes
    smallest: bytes
    largest: bytes
    num_entries: int


def result_tables(res):
    """Host copies of every table in a luda_result (which must hold host data)."""
    r = res.contents
    out = []
    keys = C.string_at(r.keys, max(1, sum(r.outputs[i].smallest_size + r.outputs[i].largest_size
                                          for i in range(r.num_outputs)))) if r.num_outputs else b""
    base = C.addressof(r.data.contents) if r.num_outputs and r.data else None
    for i in range(r.num_outputs):
        o = r.outputs[i]
        data = C.string_at(base + o.offset, o.file_size) if base is not None else b""
        out.append(Table(data=data,
                         smallest=keys[o.smallest_offset:o.smallest_offset + o.smallest_size],
                         largest=keys[o.largest_offset:o.largest_offset + o.largest_size],
                         num_entries=o.num_entries))
    return out


def device_result_tables(eng, res):
    """Like result_tables for a result kept on the device
    (LUDA_JOB_KEEP_ON_DEVICE): table bytes are read back from HBM."""
    r = res.contents
    out = []
    keys = C.string_at(r.keys, max(1, sum(r.outputs[i].smallest_size + r.outputs[i].largest_size
                                          for i in range(r.num_outputs)))) if r.num_outputs else b""
    for i in range(r.num_outputs):
        o = r.outputs[i]
        out.append(Table(data=eng.d2h(r.device_data + o.offset, o.file_size),
                         smallest=keys[o.smallest_offset:o.smallest_offset + o.smallest_size],
                         largest=keys[o.largest_offset:o.largest_offset + o.largest_size],
                         num_entries=o.num_entries))
    return out


class Workload:
    """A seeded job from libluda_workload.so (see luda_amd/workload/workload.h)."""

    def __init__(self, config, seed=None, scale=1.0):
        lib = ffi.workload_lib()
        if seed is None:
            seed = 0x4C554441 + config
        self._lib = lib
        self.ptr = lib.lw_make(config, seed, scale)
        if not self.ptr:
            raise ValueError(f"unknown workload config {config}")
        self.w = self.ptr.contents
        self.config = config

    def close(self):
        if self.ptr:
            self._lib.lw_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def groups(self):
        return [self.w.groups[i] for i in range(self.w.num_groups)]

    def group_entries(self, gi):
        """(internal key, value) pairs of one group, as bytes."""
        g = self.w.groups[gi]
        ko = [g.key_off[i] for i in range(g.n + 1)]
        vo = [g.val_off[i] for i in range(g.n + 1)]
        kb = C.string_at(g.keys, ko[-1]) if ko[-1] else b""
        vb = C.string_at(g.vals, vo[-1]) if vo[-1] else b""
        return [(kb[ko[i]:ko[i + 1]], vb[vo[i]:vo[i + 1]]) for i in range(g.n)]


def mix_specs(n=64, seed=3):
    """The mixed batch of n compaction jobs: (shape, workload seed, scale,
    MiB) with input sizes log-uniform in [4, 256] MiB.  Half are L0->L1 jobs
    with 4-12 level-0 tables (shapes 3 and 4), half Ln->Ln+1 jobs (shapes 1
    and 2); half use the 256 MiB job's key/value distribution (shapes 1 and 3)
    and half the skewed 512 MiB job's (shapes 2 and 4)."""
    import random
    rng = random.Random(seed)
    out = []
    for k in range(n):
        mib = 4.0 * (64.0 ** rng.random())
        shape = [3, 1, 4, 2][k % 4]
        scale = mib / (512.0 if shape == 2 else 256.0)
        out.append((shape, 0x4C554441 + 1000 + k, scale, mib))
    return out


def output_digest(tables):
    """SHA-256 over every output table's SHA-256, size, entry count and
    boundary keys, in output order: one value that pins a whole job."""
    import hashlib
    h = hashlib.sha256()
    for t in tables:
        h.update(hashlib.sha256(bytes(t.data)).digest())
        h.update(b"%d:%d:" % (len(t.data), t.num_entries))
        h.update(bytes(t.smallest) + b"|" + bytes(t.largest) + b";")
    return h.hexdigest()


def group_opts(opts, g):
    o = ffi.luda_options()
    C.pointer(o)[0] = opts
    o.max_file_size = g.max_file_size
    return o


def build_group_tables(builder, opts, g):
    """builder(opts_ptr, n, keys, key_off, vals, val_off) -> list[Table]."""
    return builder(C.byref(group_opts(opts, g)), g.n, C.cast(g.keys, C.c_void_p), C.cast(g.key_off, C.c_void_p),
                   C.cast(g.vals, C.c_void_p), C.cast(g.val_off, C.c_void_p))


def build_device_inputs(eng, wl, opts):
    """A workload's input tables built on the GPU by luda_build_tables and
    left in HBM: per which, [(device pointer, size)], plus the results that
    own them (free each with eng.free)."""
    dev = ([], [])
    keep = []
    for g in wl.groups:
        o = group_opts(opts, g)
        rc, out, msg = eng.build_tables(C.byref(o), g.n, C.cast(g.keys, C.c_void_p), C.cast(g.key_off, C.c_void_p),
                                        C.cast(g.vals, C.c_void_p), C.cast(g.val_off, C.c_void_p),
                                        keep_on_device=True)
        if rc != ffi.LUDA_OK:
            raise RuntimeError(f"luda_build_tables failed ({rc}): {msg}")
        res = out[2]
        keep.append(res)
        r = res.contents
        for i in range(r.num_outputs):
            dev[g.which].append((r.device_data + r.outputs[i].offset, r.outputs[i].file_size))
    return dev, keep


@dataclass
class Job:
    """A luda_job plus the Python objects whose memory it points into."""
    job: ffi.luda_job
    keep: list = field(default_factory=list)
    inputs: tuple = ((), ())

    @property
    def bytes_in(self):
        return sum(sum(s) for s in self.inputs)


def make_job(wl, opts, inputs, flags=0, device_inputs=None):
    """inputs: (list of table bytes for inputs_[0], list for inputs_[1]).
    device_inputs: optional ((ptr, size), ...) pairs per which, replacing the
    bytes with device pointers (sets LUDA_JOB_INPUTS_ON_DEVICE)."""
    w = wl.w
    keep = []
    job = ffi.luda_job()
    job.level = w.level
    job.flags = flags
    job.smallest_snapshot = w.smallest_snapshot
    job.options = opts
    sizes = ([], [])
    for which in (0, 1):
        if device_inputs is not None:
            files = device_inputs[which]
            arr = (ffi.luda_file * max(1, len(files)))()
            for i, (ptr, size) in enumerate(files):
                arr[i].data = ptr
                arr[i].size = size
                arr[i].number = 1000 * which + i + 1
                sizes[which].append(size)
            n = len(files)
        else:
            files = inputs[which]
            arr = (ffi.luda_file * max(1, len(files)))()
            for i, t in enumerate(files):
                buf = C.create_string_buffer(bytes(t), len(t))
                keep.append(buf)
                arr[i].data = C.cast(buf, C.c_void_p)
                arr[i].size = len(t)
                arr[i].number = 1000 * which + i + 1
                sizes[which].append(len(t))
            n = len(files)
        keep.append(arr)
        job.inputs[which] = C.cast(arr, C.POINTER(ffi.luda_file))
        job.num_inputs[which] = n
    if device_inputs is not None:
        job.flags |= ffi.LUDA_JOB_INPUTS_ON_DEVICE
    job.grandparents = w.grandparents
    job.num_grandparents = w.num_grandparents
    job.deeper = w.deeper
    job.num_deeper = w.num_deeper
    keep.append(wl)
    return Job(job=job, keep=keep, inputs=(tuple(sizes[0]), tuple(sizes[1])))


class Engine:
    """A luda_ctx on one HIP device (the product's C-ABI, include/luda.h)."""

    def __init__(self, device=0, budget=0):
        self.lib = ffi.engine()
        self.device = device
        self.ctx = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = self.lib.luda_init(device, budget, C.byref(self.ctx), err, 512)
        if rc != ffi.LUDA_OK:
            raise RuntimeError(f"luda_init failed ({rc}): {err.value.decode()}")

    def close(self):
        if self.ctx:
            self.lib.luda_shutdown(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _finish(self, rc, res, err, keep_on_device=False):
        if rc != ffi.LUDA_OK:
            return rc, None, err.value.decode()
        try:
            stats = res.contents.stats
            tables = [] if keep_on_device else result_tables(res)
            return rc, (tables, _stats_dict(stats), res), ""
        finally:
            if not keep_on_device:
                self.lib.luda_free_result(self.ctx, res)

    def compact(self, job, keep_on_device=False, keep_result=False, keep_decoded=False):
        """(status, (tables, stats, result_ptr) | None, message).  With
        keep_on_device (outputs left in HBM) or keep_result (host copies read,
        result kept) the result stays alive; free it with free().
        keep_decoded: LUDA_JOB_KEEP_DECODED."""
        res = ffi.RESULT_P()
        err = C.create_string_buffer(512)
        if keep_on_device:
            job.flags |= ffi.LUDA_JOB_KEEP_ON_DEVICE
        if keep_decoded:
            job.flags |= ffi.LUDA_JOB_KEEP_DECODED
        rc = self.lib.luda_compact(self.ctx, C.byref(job), C.byref(res), err, 512)
        if keep_result and rc == ffi.LUDA_OK:
            return rc, (result_tables(res), _stats_dict(res.contents.stats), res), ""
        return self._finish(rc, res, err, keep_on_device)

    def build_tables(self, opts_p, n, keys, key_off, vals, val_off, keep_on_device=False, keep_decoded=False):
        res = ffi.RESULT_P()
        err = C.create_string_buffer(512)
        flags = (ffi.LUDA_JOB_KEEP_ON_DEVICE if keep_on_device else 0) | \
            (ffi.LUDA_JOB_KEEP_DECODED if keep_decoded else 0)
        rc = self.lib.luda_build_tables(self.ctx, opts_p, n, keys, key_off, vals, val_off, flags, C.byref(res), err,
                                        512)
        return self._finish(rc, res, err, keep_on_device)

    def free(self, res):
        self.lib.luda_free_result(self.ctx, res)

    def verify(self, tables, on_device=False):
        """luda_verify_tables over table images (bytes, or (device pointer,
        size) pairs with on_device): (status, first bad table or None, message)."""
        arr = (ffi.luda_file * max(1, len(tables)))()
        keep = []
        for i, t in enumerate(tables):
            if on_device:
                arr[i].data, arr[i].size = t[0], t[1]
            else:
                b = C.create_string_buffer(bytes(t), len(t))
                keep.append(b)
                arr[i].data, arr[i].size = C.cast(b, C.c_void_p), len(t)
            arr[i].number = i + 1
        bad = C.c_uint32()
        err = C.create_string_buffer(512)
        rc = self.lib.luda_verify_tables(self.ctx, arr, len(tables), ffi.LUDA_JOB_INPUTS_ON_DEVICE if on_device else 0,
                                         C.byref(bad), err, 512)
        return rc, (None if bad.value == 0xffffffff else bad.value), err.value.decode()

    def set_tuning(self, rules):
        """luda_set_tuning from dicts with min_bytes, max_bytes,
        min_key_bytes, max_key_bytes, write_blocks_groups, decode_groups
        (missing bounds: all jobs; missing shapes: the defaults)."""
        arr = (ffi.luda_tuning_rule * max(1, len(rules)))()
        for i, r in enumerate(rules):
            arr[i].min_bytes = r.get("min_bytes", 0)
            arr[i].max_bytes = r.get("max_bytes", ffi.U64_MAX)
            arr[i].min_key_bytes = r.get("min_key_bytes", 0)
            arr[i].max_key_bytes = r.get("max_key_bytes", 0xFFFFFFFF)
            arr[i].tuning.write_blocks_groups = r.get("write_blocks_groups", 0)
            arr[i].tuning.decode_groups = r.get("decode_groups", 0)
        rc = self.lib.luda_set_tuning(self.ctx, arr, len(rules))
        if rc != ffi.LUDA_OK:
            raise RuntimeError(f"luda_set_tuning failed ({rc})")

    # -- batched / served form (luda_submit, luda_poll, luda_drain) --------
    def set_slots(self, slots):
        rc = self.lib.luda_set_slots(self.ctx, slots)
        if rc != ffi.LUDA_OK:
            raise RuntimeError(f"luda_set_slots failed ({rc})")

    def submit(self, job, on_done, keep_on_device=False, keep_decoded=False):
        """Queues `job`; on_done(status, (tables, stats, result_ptr) | None,
        message) runs from poll()/drain().  The job must stay alive until then."""
        if keep_on_device:
            job.flags |= ffi.LUDA_JOB_KEEP_ON_DEVICE
        if keep_decoded:
            job.flags |= ffi.LUDA_JOB_KEEP_DECODED
        return self._queue(lambda uid: self.lib.luda_submit(self.ctx, C.byref(job), C.cast(self._done_fn, C.c_void_p),
                                                            C.c_void_p(uid)), on_done, job, keep_on_device)

    def submit_build(self, opts, n, keys, key_off, vals, val_off, on_done, keep_on_device=False, keep_decoded=False):
        """Queues a flush (luda_submit_build); the arrays must stay alive until
        on_done runs."""
        flags = (ffi.LUDA_JOB_KEEP_ON_DEVICE if keep_on_device else 0) | \
            (ffi.LUDA_JOB_KEEP_DECODED if keep_decoded else 0)
        hold = (opts, keys, key_off, vals, val_off)
        return self._queue(lambda uid: self.lib.luda_submit_build(
            self.ctx, C.byref(opts), n, keys, key_off, vals, val_off, flags, C.cast(self._done_fn, C.c_void_p),
            C.c_void_p(uid)), on_done, hold, keep_on_device)

    def _queue(self, call, on_done, hold, keep_on_device):
        if not hasattr(self, "_pending"):
            self._pending = {}
            self._next_id = 1

            def trampoline(user, status, res, err):
                # an exception cannot cross the C frame of luda_poll/luda_drain:
                # keep the first one and raise it when that call returns
                try:
                    cb, jb, kod = self._pending.pop(user)
                    if status != ffi.LUDA_OK:
                        cb(status, None, (err or b"").decode())
                        return
                    try:
                        stats = res.contents.stats
                        tables = [] if kod else result_tables(res)
                        cb(status, (tables, _stats_dict(stats), res), "")
                    finally:
                        if not kod:
                            self.lib.luda_free_result(self.ctx, res)
                except BaseException as e:  # noqa: BLE001
                    if self._cb_error is None:
                        self._cb_error = e
            self._cb_error = None
            self._done_fn = ffi.DONE_FN(trampoline)
        uid = self._next_id
        self._next_id += 1
        self._pending[uid] = (on_done, hold, keep_on_device)
        rc = call(uid)
        if rc != ffi.LUDA_OK:
            self._pending.pop(uid)
            raise RuntimeError(f"luda_submit failed ({rc})")
        return uid

    def _reraise(self):
        e = getattr(self, "_cb_error", None)
        if e is not None:
            self._cb_error = None
            raise e

    def poll(self):
        n = self.lib.luda_poll(self.ctx)
        self._reraise()
        return n

    def drain(self):
        n = self.lib.luda_drain(self.ctx)
        self._reraise()
        return n

    def builder(self):
        """A table builder for build_group_tables backed by the engine."""
        def build(opts_p, n, keys, key_off, vals, val_off):
            rc, out, msg = self.build_tables(opts_p, n, keys, key_off, vals, val_off)
            if rc != ffi.LUDA_OK:
                raise RuntimeError(f"luda_build_tables failed ({rc}): {msg}")
            return out[0]
        return build

    def build_entries(self, entries, opts):
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
        return self.build_tables(C.byref(opts), n, C.cast(kb, C.c_void_p), C.cast(ko, C.c_void_p),
                                 C.cast(vb, C.c_void_p), C.cast(vo, C.c_void_p))

    def read_tables(self, tables, verify=False, on_device=False):
        """luda_read_tables: (status, ([entries per table of (key, value)]) |
        None, message); tables are bytes, or (device pointer, size) pairs with
        on_device."""
        arr = (ffi.luda_file * max(1, len(tables)))()
        keep = []
        for i, t in enumerate(tables):
            if on_device:
                arr[i].data, arr[i].size = t
            else:
                b = C.create_string_buffer(bytes(t), len(t))
                keep.append(b)
                arr[i].data, arr[i].size = C.cast(b, C.c_void_p), len(t)
        flags = (ffi.LUDA_READ_VERIFY if verify else 0) | (ffi.LUDA_JOB_INPUTS_ON_DEVICE if on_device else 0)
        out = ffi.ENTRIES_P()
        err = C.create_string_buffer(512)
        rc = self.lib.luda_read_tables(self.ctx, arr, len(tables), flags, C.byref(out), err, 512)
        if rc != ffi.LUDA_OK:
            return rc, None, err.value.decode()
        try:
            e = out.contents
            n = e.n
            ko = [e.key_off[i] for i in range(n + 1)]
            vo = [e.val_off[i] for i in range(n + 1)]
            kb = C.string_at(e.keys, ko[n]) if ko[n] else b""
            vb = C.string_at(e.vals, vo[n]) if vo[n] else b""
            tf = [e.table_first[i] for i in range(e.num_tables + 1)]
            ents = [(kb[ko[i]:ko[i + 1]], vb[vo[i]:vo[i + 1]]) for i in range(n)]
            return rc, [ents[tf[t]:tf[t + 1]] for t in range(e.num_tables)], ""
        finally:
            self.lib.luda_free_entries(out)

    # -- file I/O (luda_read_files, luda_write_outputs) --------------------
    def read_files(self, paths, numbers=None):
        """(status, luda_file_set pointer | None, message); free the set with
        free_files."""
        arr = (C.c_char_p * max(1, len(paths)))(*[p.encode() for p in paths])
        nums = (C.c_uint64 * max(1, len(paths)))(*(numbers or [])) if numbers else None
        out = ffi.FILESET_P()
        err = C.create_string_buffer(512)
        rc = self.lib.luda_read_files(self.ctx, arr, nums, len(paths), C.byref(out), err, 512)
        return rc, (out if rc == ffi.LUDA_OK else None), err.value.decode()

    def free_files(self, fs):
        self.lib.luda_free_files(self.ctx, fs)

    def write_outputs(self, res, paths, sync=True):
        arr = (C.c_char_p * max(1, len(paths)))(*[p.encode() for p in paths])
        err = C.create_string_buffer(512)
        rc = self.lib.luda_write_outputs(self.ctx, res, arr, 1 if sync else 0, err, 512)
        return rc, err.value.decode()

    def device_alloc(self, nbytes):
        p = C.c_void_p()
        if self.lib.luda_device_alloc(self.ctx, nbytes, C.byref(p)) != ffi.LUDA_OK:
            raise RuntimeError("luda_device_alloc failed")
        return p.value

    def device_free(self, p):
        self.lib.luda_device_free(self.ctx, p)

    def h2d(self, dst, data):
        buf = C.create_string_buffer(bytes(data), len(data))
        if self.lib.luda_memcpy_h2d(self.ctx, dst, buf, len(data)) != ffi.LUDA_OK:
            raise RuntimeError("luda_memcpy_h2d failed")

    def d2h(self, src, nbytes):
        buf = C.create_string_buffer(nbytes)
        if self.lib.luda_memcpy_d2h(self.ctx, buf, src, nbytes) != ffi.LUDA_OK:
            raise RuntimeError("luda_memcpy_d2h failed")
        return buf.raw


def _stats_dict(s):
    return {k: (getattr(s, k).decode() if k == "top_kernel" else getattr(s, k)) for k, _ in ffi.luda_stats._fields_}
