"""File I/O around a compaction, through the C-ABI: input tables read from
disk into HBM (luda_read_files), compacted from there, outputs written back
to files (luda_write_outputs), byte for byte against the oracle; the file
reads overlapping a compaction that runs on another thread of the same
context."""
import ctypes as C
import os
import threading

import pytest

import oracle
from luda_amd import ffi, jobs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = jobs.Engine(0)
    yield e
    e.close()


def host_job_tables(cfg, scale, seed=None):
    opts = ffi.default_options()
    wl = jobs.Workload(cfg, seed=seed, scale=scale)
    ins = ([], [])
    for g in wl.groups:
        ins[g.which].extend(t.data for t in jobs.build_group_tables(oracle.builder("oracle"), opts, g))
    return wl, opts, ins


def job_from_set(wl, opts, fs, counts, flags=0):
    """A luda_job whose inputs are the file set's device-resident files:
    the first counts[0] are inputs_[0], the rest inputs_[1]."""
    s = fs.contents
    dev = ([], [])
    for i in range(s.n):
        dev[0 if i < counts[0] else 1].append((s.files[i].data, s.files[i].size))
    return jobs.make_job(wl, opts, None, flags=flags, device_inputs=dev)


def write_tables(tmp, tables, prefix):
    paths = []
    for i, t in enumerate(tables):
        p = os.path.join(tmp, f"{prefix}{i:06d}.ldb")
        with open(p, "wb") as f:
            f.write(t)
        paths.append(p)
    return paths


@pytest.mark.parametrize("keep_on_device", [True, False])
def test_files_in_files_out_match_oracle(eng, tmp_path, keep_on_device):
    wl, opts, ins = host_job_tables(0, 0.3)
    hjob = jobs.make_job(wl, opts, ins)
    rc, want = oracle.compact(hjob.job, "oracle")
    assert rc == 0
    paths = write_tables(str(tmp_path), ins[0] + ins[1], "in")
    print("read", len(paths), "files", flush=True)
    rc, fs, msg = eng.read_files(paths, numbers=list(range(11, 11 + len(paths))))
    assert rc == 0, msg
    try:
        s = fs.contents
        assert s.n == len(paths) and s.total_bytes == sum(map(len, ins[0] + ins[1]))
        assert [s.files[i].number for i in range(s.n)] == list(range(11, 11 + len(paths)))
        assert eng.d2h(s.files[3].data, s.files[3].size) == (ins[0] + ins[1])[3]
        job = job_from_set(wl, opts, fs, (len(ins[0]), len(ins[1])))
        if keep_on_device:
            rc, out, msg = eng.compact(job.job, keep_on_device=True)
        else:
            rc, out, msg = eng.compact(job.job, keep_result=True)
        assert rc == 0, msg
        res = out[2]
        opaths = [os.path.join(str(tmp_path), f"out{i:06d}.ldb") for i in range(res.contents.num_outputs)]
        rc, msg = eng.write_outputs(res, opaths, sync=True)
        assert rc == 0, msg
        got = []
        for p in opaths:
            with open(p, "rb") as f:
                got.append(f.read())
        assert got == [t.data for t in want]
        eng.free(res)
    finally:
        eng.free_files(fs)


def test_missing_file_is_io_error(eng, tmp_path):
    rc, fs, msg = eng.read_files([os.path.join(str(tmp_path), "absent.ldb")])
    assert rc == ffi.LUDA_IO_ERROR and fs is None and "absent" in msg


def test_reads_overlap_a_running_compaction(eng, tmp_path):
    """Another job's files are read into HBM on the I/O stream while a
    compaction runs on the compute stream of the same context; both results
    are the ones each gives alone."""
    wl1, opts, ins1 = host_job_tables(1, 0.05)
    wl2, _, ins2 = host_job_tables(0, 0.5, seed=99)
    p1 = write_tables(str(tmp_path), ins1[0] + ins1[1], "a")
    p2 = write_tables(str(tmp_path), ins2[0] + ins2[1], "b")
    rc, fs1, msg = eng.read_files(p1)
    assert rc == 0, msg
    job1 = job_from_set(wl1, opts, fs1, (len(ins1[0]), len(ins1[1])))
    rc, alone, msg = eng.compact(job1.job)
    assert rc == 0, msg
    box = {}

    def reader():
        box["read"] = eng.read_files(p2)
    th = threading.Thread(target=reader)
    th.start()
    rc, during, msg = eng.compact(job1.job)
    th.join()
    assert rc == 0, msg
    assert [t.data for t in during[0]] == [t.data for t in alone[0]]
    rc2, fs2, msg2 = box["read"]
    assert rc2 == 0, msg2
    job2 = job_from_set(wl2, opts, fs2, (len(ins2[0]), len(ins2[1])))
    rc, out2, msg = eng.compact(job2.job)
    assert rc == 0, msg
    hjob2 = jobs.make_job(wl2, opts, ins2)
    rc, want2 = oracle.compact(hjob2.job, "oracle")
    assert [t.data for t in out2[0]] == [t.data for t in want2]
    eng.free_files(fs1)
    eng.free_files(fs2)
