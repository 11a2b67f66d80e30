"""Decoded entries resident across chained jobs (LUDA_JOB_KEEP_DECODED):
tables written by luda_build_tables and luda_compact keep their decoded
entries in HBM, and a compaction whose inputs are all such tables gathers
them instead of decoding (no phase A, no decode: blocks_in stays 0).  Its
tables equal the oracle's on host copies of the same inputs and the decoding
path's on unregistered copies, through a chain of jobs."""
import ctypes as C

import pytest

import oracle
from luda_amd import ffi, jobs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = jobs.Engine(0)
    yield e
    e.close()


def build_resident(eng, wl, opts, groups):
    """Tables of the given workload groups built on the GPU, kept decoded:
    [(device pointer, size)] per group, and the results that own them."""
    out, keep = [], []
    for gi in groups:
        g = wl.groups[gi]
        o = jobs.group_opts(opts, g)
        rc, res, msg = eng.build_tables(C.byref(o), g.n, C.cast(g.keys, C.c_void_p), C.cast(g.key_off, C.c_void_p),
                                        C.cast(g.vals, C.c_void_p), C.cast(g.val_off, C.c_void_p),
                                        keep_on_device=True, keep_decoded=True)
        assert rc == 0, msg
        r = res[2].contents
        keep.append(res[2])
        out.append([(r.device_data + r.outputs[i].offset, r.outputs[i].file_size) for i in range(r.num_outputs)])
    return out, keep


def make_device_job(level, opts, snapshot, ins, wl=None):
    job = ffi.luda_job()
    job.level = level
    job.flags = ffi.LUDA_JOB_INPUTS_ON_DEVICE
    job.smallest_snapshot = snapshot
    job.options = opts
    keep = []
    for which in (0, 1):
        arr = (ffi.luda_file * max(1, len(ins[which])))()
        for i, (p, n) in enumerate(ins[which]):
            arr[i].data, arr[i].size, arr[i].number = p, n, 100 * which + i + 1
        keep.append(arr)
        job.inputs[which] = C.cast(arr, C.POINTER(ffi.luda_file))
        job.num_inputs[which] = len(ins[which])
    if wl is not None:
        job.grandparents, job.num_grandparents = wl.w.grandparents, wl.w.num_grandparents
        job.deeper, job.num_deeper = wl.w.deeper, wl.w.num_deeper
        keep.append(wl)
    return job, keep


def host_copy_job(eng, job):
    """The same job on host copies of its inputs (decoded by the engine, or
    run by the oracle)."""
    hold = []
    h = ffi.luda_job()
    C.pointer(h)[0] = job
    h.flags = 0
    for which in (0, 1):
        n = job.num_inputs[which]
        arr = (ffi.luda_file * max(1, n))()
        for i in range(n):
            f = job.inputs[which][i]
            b = C.create_string_buffer(eng.d2h(f.data, f.size), f.size)
            hold.append(b)
            arr[i].data, arr[i].size, arr[i].number = C.cast(b, C.c_void_p), f.size, f.number
        hold.append(arr)
        h.inputs[which] = C.cast(arr, C.POINTER(ffi.luda_file))
    return h, hold


def tables_of(out):
    return [(t.data, t.smallest, t.largest, t.num_entries) for t in out[0]]


@pytest.mark.parametrize("shape,scale", [(3, 0.02), (4, 0.05), (1, 0.02)])
def test_chained_jobs_gather_resident_entries(eng, shape, scale):
    opts = ffi.default_options()
    wl = jobs.Workload(shape, scale=scale)
    ngroups = len(wl.groups)
    which_of = [g.which for g in wl.groups]
    built, keep = build_resident(eng, wl, opts, range(ngroups))
    ins = ([], [])
    for gi in range(ngroups):
        ins[which_of[gi]].extend(built[gi])
    print("resident job", shape, scale, [len(x) for x in ins], flush=True)
    job, hold = make_device_job(wl.w.level, opts, wl.w.smallest_snapshot, ins, wl)
    rc, out1, msg = eng.compact(job, keep_on_device=True, keep_decoded=True)
    assert rc == 0, msg
    assert out1[1]["blocks_in"] == 0, "inputs were decoded, not gathered"
    first = jobs.device_result_tables(eng, out1[2])
    hjob, hhold = host_copy_job(eng, job)
    rc, want = oracle.compact(hjob, "oracle")
    assert rc == 0
    got = [(t.data, t.smallest, t.largest, t.num_entries) for t in first]
    assert got == [(t.data, t.smallest, t.largest, t.num_entries) for t in want]
    rc, dec, msg = eng.compact(hjob)  # the decoding path on host copies
    assert rc == 0 and dec[1]["blocks_in"] > 0
    assert tables_of(dec) == got
    # next in the chain: the outputs (resident) under new level-0 tables
    r = out1[2].contents
    outs = [(r.device_data + r.outputs[i].offset, r.outputs[i].file_size) for i in range(r.num_outputs)]
    wl2 = jobs.Workload(3, seed=777 + shape, scale=scale / 2)
    l0_groups = [i for i, g in enumerate(wl2.groups) if g.which == 0]
    built2, keep2 = build_resident(eng, wl2, opts, l0_groups)
    ins2 = ([p for b in built2 for p in b], outs)
    snap = max(wl.w.smallest_snapshot, wl2.w.smallest_snapshot) + (1 << 40)
    # the second job's level-0 tables must be newer: shift their sequence
    # numbers by building from the workload as is and a snapshot past all
    job2, hold2 = make_device_job(0, opts, snap, ins2)
    rc, out2, msg = eng.compact(job2, keep_on_device=True, keep_decoded=True)
    assert rc == 0, msg
    assert out2[1]["blocks_in"] == 0
    hjob2, hhold2 = host_copy_job(eng, job2)
    rc, want2 = oracle.compact(hjob2, "oracle")
    assert rc == 0
    assert [(t.data, t.smallest, t.largest, t.num_entries) for t in jobs.device_result_tables(eng, out2[2])] == \
        [(t.data, t.smallest, t.largest, t.num_entries) for t in want2]
    # one unregistered input: the job decodes, with the same result
    p, n = ins2[1][0]
    copy = eng.device_alloc(n + 16)
    eng.h2d(copy, eng.d2h(p, n))
    ins3 = (ins2[0], [(copy, n)] + ins2[1][1:])
    job3, hold3 = make_device_job(0, opts, snap, ins3)
    rc, out3, msg = eng.compact(job3)
    assert rc == 0, msg
    assert out3[1]["blocks_in"] > 0
    assert tables_of(out3) == [(t.data, t.smallest, t.largest, t.num_entries) for t in want2]
    eng.device_free(copy)
    for res in keep + keep2 + [out1[2], out2[2]]:
        eng.free(res)
