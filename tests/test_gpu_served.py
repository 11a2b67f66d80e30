"""The served form on the GPU: flushes and compactions of a write stream
queued on the engine's worker slots (luda_submit_build / luda_submit), with
tables resident in HBM from flush to deletion.  Every sampled job's output is
the same as the same job run alone through luda_compact and as the oracle's
on host copies of its inputs, and the finished tree holds exactly the newest
put of every key."""
import ctypes as C
import hashlib

import pytest

import oracle
from luda_amd import ffi, jobs, served
from served_backends import expected_contents, small_options, tree_contents

pytestmark = pytest.mark.gpu


def _sha(b):
    return hashlib.sha256(b).hexdigest()


@pytest.mark.parametrize("slots", [3, 1])
def test_served_stream_matches_alone(slots):
    print(f"served stream, {slots} slots", flush=True)
    mems = list(served.write_stream(60000, 11, key_space=40000, entries_per_flush=2000))
    opts = small_options(32 << 10)
    beng = jobs.Engine(0)
    beng.set_slots(slots)
    peng = jobs.Engine(0)
    try:
        be = served.EngineBackend(beng, opts)
        lsm = served.Lsm(max_file_size=opts.max_file_size, level1_bytes=96 << 10)
        samples = []

        def on_job(spec, files):
            if len(samples) < 24:
                for which in (0, 1):
                    for f in spec.inputs[which]:
                        f.owner.refs += 1  # held until the job is re-run alone
                samples.append((spec, [_sha(be.table_bytes(f)) for f in files]))
        lsm, st = served.run_served(be, mems, 0.0, lsm=lsm, on_job=on_job)
        lsm.check()
        summ = st.summary()
        assert summ["flushes"] == len(mems) and summ["compactions"] >= 10
        assert samples
        pbe = served.EngineBackend(peng, opts)
        scratch = served.Lsm()
        for spec, want in samples:
            files, _ = pbe.compact(scratch, spec)
            got = [_sha(pbe.table_bytes(f)) for f in files]
            assert got == want, f"level-{spec.level} job differs when run alone"
            # the oracle on host copies of the same inputs
            hjob, keep = served._job_struct(spec, opts)
            hjob.flags = 0
            host = []
            for which in (0, 1):
                for i, f in enumerate(spec.inputs[which]):
                    b = C.create_string_buffer(be.table_bytes(f), f.size)
                    host.append(b)
                    hjob.inputs[which][i].data = C.cast(b, C.c_void_p)
            rc, tables = oracle.compact(hjob)
            assert rc == 0 and [_sha(t.data) for t in tables] == want, "differs from the oracle"
            if files:
                files[0].owner.refs = 1
                files[0].owner.drop()
            for which in (0, 1):
                for f in spec.inputs[which]:
                    f.owner.drop()
        assert tree_contents(lsm, be.table_bytes) == expected_contents(mems)
    finally:
        peng.close()
        beng.close()


@pytest.mark.parametrize("mode", ["served", "plain"])
def test_stream_with_output_files(tmp_path, mode):
    """Every table also written to a file (the served form on its writer
    thread beside the slots' kernels, the plain path before its next job):
    the directory then holds exactly the finished tree's files, byte for byte
    the tables in HBM, every compaction having deleted its inputs' files, and
    the tree holds the newest put of every key."""
    print(f"stream with output files, {mode}", flush=True)
    mems = list(served.write_stream(40000, 12, key_space=30000, entries_per_flush=2000))
    opts = small_options(32 << 10)
    eng = jobs.Engine(0)
    if mode == "served":
        eng.set_slots(3)
    try:
        be = served.EngineBackend(eng, opts, io_dir=str(tmp_path), io_sync=True)
        lsm = served.Lsm(max_file_size=opts.max_file_size, level1_bytes=96 << 10)
        lsm.on_delete = be.delete_file
        run = served.run_served if mode == "served" else served.run_plain
        lsm, st = run(be, mems, 0.0, lsm=lsm)
        be.close()
        lsm.check()
        summ = st.summary()
        assert summ["compactions"] >= 5
        tree = {"%06d.ldb" % f.number: f for level in lsm.levels for f in level}
        assert set(p.name for p in tmp_path.iterdir()) == set(tree)
        for name, f in tree.items():
            assert (tmp_path / name).read_bytes() == be.table_bytes(f), name
        assert be.io["files"] == be.io["deleted"] + len(tree)
        assert tree_contents(lsm, be.table_bytes) == expected_contents(mems)
    finally:
        eng.close()
