"""Pins the oracle to data the reference's own code produced: whole-job
SHA-256 fixtures and corner-case tables written by tests/golden/make_golden.py
from oracle/_ref (the reference compiled from its sources)."""
import hashlib
import json
import os

import pytest

import oracle
from luda_amd import ffi, jobs

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")

with open(os.path.join(GOLD, "jobs.json")) as f:
    JOBS = json.load(f)
with open(os.path.join(GOLD, "tables.json")) as f:
    TABLES = json.load(f)

OPTION_SETS = {
    "default": dict(),
    "small_blocks": dict(block_size=256, block_restart_interval=4, max_file_size=1 << 20),
    "no_filter": dict(bloom_bits_per_key=0),
}


def sha(b):
    return hashlib.sha256(b).hexdigest()


@pytest.mark.parametrize("rec", JOBS, ids=lambda r: f"cfg{r['config']}-{r['seed']}-{r['scale']}-{r['options']}")
def test_oracle_reproduces_reference_jobs(rec):
    opts = ffi.default_options(**OPTION_SETS[rec["options"]])
    wl = jobs.Workload(rec["config"], seed=rec["seed"], scale=rec["scale"])
    ins = ([], [])
    for g in wl.groups:
        ins[g.which].extend(t.data for t in jobs.build_group_tables(oracle.builder("oracle"), opts, g))
    assert [[sha(t) for t in ins[0]], [sha(t) for t in ins[1]]] == rec["inputs"]
    job = jobs.make_job(wl, opts, ins)
    rc, outs = oracle.compact(job.job, "oracle")
    assert rc == 0
    got = [{"sha256": sha(t.data), "file_size": len(t.data), "num_entries": t.num_entries,
            "smallest": t.smallest.hex(), "largest": t.largest.hex()} for t in outs]
    assert got == rec["outputs"]
    # every output is readable with CRCs checked, strictly sorted, and its
    # smallest/largest match its first/last entry
    for t in outs:
        rc, ents = oracle.table_entries(t.data)
        assert rc == 0 and len(ents) == t.num_entries
        assert ents[0][0] == t.smallest and ents[-1][0] == t.largest
        o = oracle.load()
        for a, b in zip(ents, ents[1:]):
            assert o.or_ikey_compare(a[0], len(a[0]), b[0], len(b[0])) < 0


@pytest.mark.parametrize("name", sorted(TABLES))
def test_oracle_reproduces_reference_tables(name):
    meta = TABLES[name]
    opts = ffi.luda_options(**meta["options"])
    entries = [(bytes.fromhex(k), bytes.fromhex(v)) for k, v in meta["entries"]]
    tables = oracle.build_tables(entries, opts, "oracle")
    assert len(tables) == 1
    with open(os.path.join(GOLD, "tables", name + ".ldb"), "rb") as f:
        ref = f.read()
    assert sha(ref) == meta["sha256"]
    assert tables[0].data == ref
    rc, ents = oracle.table_entries(ref)
    assert rc == 0 and ents == entries


def test_reader_rejects_bad_crc():
    with open(os.path.join(GOLD, "tables", "decimal.ldb"), "rb") as f:
        data = bytearray(f.read())
    data[100] ^= 0x40
    rc, _ = oracle.table_entries(bytes(data), verify_crc=True)
    assert rc == ffi.LUDA_CORRUPTION
    rc, _ = oracle.table_entries(bytes(data), verify_crc=False)
    assert rc == 0


with open(os.path.join(GOLD, "full_jobs.json")) as f:
    FULL = json.load(f)
with open(os.path.join(GOLD, "mix_jobs.json")) as f:
    MIX = json.load(f)


def oracle_outputs(cfg, seed, scale):
    opts = ffi.default_options()
    wl = jobs.Workload(cfg, seed=seed, scale=scale)
    ins = ([], [])
    for g in wl.groups:
        ins[g.which].extend(t.data for t in jobs.build_group_tables(oracle.builder("oracle"), opts, g))
    job = jobs.make_job(wl, opts, ins)
    rc, outs = oracle.compact(job.job)
    assert rc == 0
    return job.bytes_in, outs


def test_oracle_full_size_l0_job_matches_reference():
    """Config 0 at its full size (20 MiB), every output table."""
    rec = [r for r in FULL if r["config"] == 0][0]
    nbytes, outs = oracle_outputs(0, rec["seed"], 1.0)
    assert nbytes == rec["input_bytes"]
    assert [sha(t.data) for t in outs] == [o["sha256"] for o in rec["outputs"]]
    assert jobs.output_digest(outs) == rec["digest"]


@pytest.mark.parametrize("k", [0, 2, 3, 6])
def test_oracle_mix_jobs_match_reference(k):
    """Mixed-batch jobs of every shape, including L0->L1 jobs with 4-12
    level-0 tables (only jobs small enough for the CPU suite)."""
    small = [(i, r) for i, r in enumerate(MIX["records"]) if r["input_bytes"] <= (24 << 20)]
    i, rec = small[k]
    nbytes, outs = oracle_outputs(rec["shape"], rec["seed"], rec["scale"])
    assert nbytes == rec["input_bytes"]
    assert jobs.output_digest(outs) == rec["digest"]
