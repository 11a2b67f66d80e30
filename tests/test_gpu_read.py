"""The GPU table reader (luda_read_tables) against the oracle's reader (the
restatement of Table::Open and a full Block::Iter iteration, with block CRCs
checked): every entry of every table, in order, on corner-case tables, job
inputs and outputs, wide keys; and LUDA_READ_VERIFY finding a flipped byte."""
import json
import os

import pytest

import oracle
from luda_amd import ffi, jobs
from tests.tiny_jobs import wide_key_job

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def eng():
    e = jobs.Engine(0)
    yield e
    e.close()


def oracle_entries(t):
    rc, ents = oracle.table_entries(t)
    assert rc == 0
    return ents


def test_golden_corner_tables(eng):
    with open(os.path.join(HERE, "golden", "tables.json")) as f:
        meta = json.load(f)
    tables = []
    for name in sorted(meta):
        with open(os.path.join(HERE, "golden", "tables", name + ".ldb"), "rb") as f:
            tables.append(f.read())
    rc, got, msg = eng.read_tables(tables, verify=True)
    assert rc == 0, msg
    assert got == [oracle_entries(t) for t in tables]
    # the entries the tables were built from
    for name, ents in zip(sorted(meta), got):
        assert ents == [(bytes.fromhex(k), bytes.fromhex(v)) for k, v in meta[name]["entries"]], name


@pytest.mark.parametrize("cfg,scale", [(0, 0.2), (2, 0.01), (3, 0.01)])
def test_job_inputs_and_outputs(eng, cfg, scale):
    opts = ffi.default_options()
    wl = jobs.Workload(cfg, scale=scale)
    ins = ([], [])
    for g in wl.groups:
        ins[g.which].extend(t.data for t in jobs.build_group_tables(oracle.builder("oracle"), opts, g))
    tables = ins[0] + ins[1]
    rc, got, msg = eng.read_tables(tables, verify=True)
    assert rc == 0, msg
    assert got == [oracle_entries(t) for t in tables]
    job = jobs.make_job(wl, opts, ins)
    rc, out, msg = eng.compact(job.job)
    assert rc == 0, msg
    outs = [t.data for t in out[0]]
    rc, got, msg = eng.read_tables(outs)
    assert rc == 0, msg
    assert got == [oracle_entries(t) for t in outs]
    assert sum(len(g) for g in got) == out[1]["entries_kept"]


def test_wide_keys_and_device_resident(eng):
    tj = wide_key_job(2)
    job, keep = tj.make_job(lambda e, o: [t.data for t in oracle.build_tables(e, o, "oracle")])
    tables = tj.tables[0] + tj.tables[1]
    ptrs = []
    for t in tables:
        p = eng.device_alloc(len(t) + 16)
        eng.h2d(p, t)
        ptrs.append((p, len(t)))
    try:
        rc, got, msg = eng.read_tables(ptrs, verify=True, on_device=True)
        assert rc == 0, msg
        assert got == [oracle_entries(t) for t in tables]
    finally:
        for p, _ in ptrs:
            eng.device_free(p)


def test_verify_flag_rejects_a_flipped_byte(eng):
    opts = ffi.default_options()
    wl = jobs.Workload(0, scale=0.05)
    g = wl.groups[-1]
    tables = [t.data for t in jobs.build_group_tables(oracle.builder("oracle"), opts, g)]
    bad = bytearray(tables[0])
    bad[200] ^= 1
    rc, got, msg = eng.read_tables([tables[0], bytes(bad)], verify=True)
    assert rc == ffi.LUDA_CORRUPTION
    rc, got, msg = eng.read_tables([tables[0]], verify=True)
    assert rc == 0 and got == [oracle_entries(tables[0])]
