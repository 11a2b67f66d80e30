"""Parity of the HIP engine with the oracle (and, through the golden
fixtures, with the reference's own code): byte-identical tables and identical
table metadata.  Every case calls the engine through the C-ABI."""
import ctypes as C
import hashlib
import json
import os
import random
import struct

import pytest

import oracle
from luda_amd import ffi, jobs
from tests.tiny_jobs import (TinyJob, ikey, interval_job, malformed_job, many_children_job, many_grandparents_job,
                             wide_key_job)

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
OPTION_SETS = {
    "default": dict(),
    "small_blocks": dict(block_size=256, block_restart_interval=4, max_file_size=1 << 20),
    "no_filter": dict(bloom_bits_per_key=0),
}


@pytest.fixture(scope="module")
def eng():
    e = jobs.Engine(0)
    yield e
    e.close()


def sha(b):
    return hashlib.sha256(b).hexdigest()


def summary(ts):
    return [(sha(t.data), len(t.data), t.smallest, t.largest, t.num_entries) for t in ts]


def test_build_golden_corner_tables(eng):
    with open(os.path.join(GOLD, "tables.json")) as f:
        meta = json.load(f)
    for name in sorted(meta):
        m = meta[name]
        print("table", name, flush=True)
        opts = ffi.luda_options(**m["options"])
        entries = [(bytes.fromhex(k), bytes.fromhex(v)) for k, v in m["entries"]]
        rc, out, msg = eng.build_entries(entries, opts)
        assert rc == 0, (name, msg)
        tables = out[0]
        assert len(tables) == 1, name
        with open(os.path.join(GOLD, "tables", name + ".ldb"), "rb") as f:
            assert tables[0].data == f.read(), name


def test_build_30_entries_one_block(eng):
    """The first case that ran on the GPU: 30 entries in one 684-byte block,
    assembled by wave 0 at LDS offset 0, CRC pad of 83 bytes."""
    ents = [(b"key%03d" % i + struct.pack("<Q", (5 << 8) | 1), b"x" * 10) for i in range(30)]
    opts = ffi.default_options(max_file_size=ffi.U64_MAX)
    rc, out, msg = eng.build_entries(ents, opts)
    assert rc == 0, msg
    assert summary(out[0]) == summary(oracle.build_tables(ents, opts, "oracle"))


def random_entries(rng, n, klen_range, vlen_choices, versions=1):
    keys = set()
    while len(keys) < n:
        L = rng.randint(*klen_range)
        keys.add(bytes(rng.randrange(256) for _ in range(L)))
    ents = []
    seq = 1
    for u in sorted(keys):
        for v in range(versions):
            ents.append((u, 10 ** 6 - seq, rng.choice([0, 1, 1, 1])))
            seq += 1
    ents.sort(key=lambda e: (e[0], -e[1]))
    return [(ikey(u, s, t), b"" if t == 0 else bytes(rng.randrange(256) for _ in range(rng.choice(vlen_choices))))
            for u, s, t in ents]


@pytest.mark.parametrize("case", range(12))
def test_build_tables_match_oracle(eng, case):
    rng = random.Random(1000 + case)
    opts = ffi.default_options(block_size=rng.choice([64, 256, 1024, 4096]),
                               block_restart_interval=rng.choice([1, 3, 16]),
                               bloom_bits_per_key=rng.choice([0, 10, 7]),
                               max_file_size=rng.choice([2000, 20000, ffi.U64_MAX]))
    ents = random_entries(rng, rng.choice([1, 5, 300, 2000]), (0, rng.choice([4, 30, 80])), [0, 3, 100, 700],
                          versions=rng.choice([1, 3]))
    print("case", case, len(ents), opts.block_size, opts.max_file_size, flush=True)
    rc, out, msg = eng.build_entries(ents, opts)
    assert rc == 0, msg
    want = oracle.build_tables(ents, opts, "oracle")
    assert summary(out[0]) == summary(want)


@pytest.mark.parametrize("block_size,interval,nkeys,vmax", [(8192, 16, 40, 0), (8192, 1000, 60, 0), (7000, 5, 30, 3),
                                                            (8192, 1, 30, 1), (1, 16, 3, 30)])
def test_block_cuts_many_small_entries(eng, block_size, interval, nkeys, vmax):
    """Blocks of hundreds of 10-byte entries (versions of a few user keys):
    the block cut's staged span is exceeded and finished from global
    memory; restart intervals from 1 to 1000."""
    rng = random.Random(block_size * 7 + interval)
    ents = [(ikey(b"k%03d" % u, 100000 - v, 1), bytes(rng.randrange(256) for _ in range(rng.randint(0, vmax))))
            for u in range(nkeys) for v in range(100)]
    opts = ffi.default_options(block_size=block_size, block_restart_interval=interval, max_file_size=20000)
    print("cuts", block_size, interval, len(ents), flush=True)
    rc, out, msg = eng.build_entries(ents, opts)
    assert rc == 0, msg
    assert summary(out[0]) == summary(oracle.build_tables(ents, opts, "oracle"))


def engine_tables(eng, ents, opts):
    rc, out, msg = eng.build_entries(ents, opts)
    assert rc == 0, msg
    return [t.data for t in out[0]]


@pytest.mark.parametrize("block", range(4))
def test_tiny_jobs_match_oracle(eng, block):
    for seed in range(block * 100, (block + 1) * 100):
        tj = TinyJob(seed)
        job, keep = tj.make_job(lambda e, o: [t.data for t in oracle.build_tables(e, o, "oracle")])
        rc1, o1 = oracle.compact(job, "oracle")
        rc2, out, msg = eng.compact(job)
        assert rc1 == 0 and rc2 == 0, (seed, msg)
        assert summary(out[0]) == summary(o1), seed


def test_fresh_context_first_job_wider_keys_than_guessed():
    """Each job on a context of its own: the key planes are first sized from
    the keys at restart points only, so a longer key further in overflows
    them and the job must be retried wider, not rejected (kernels after the
    overflow see truncated keys and may report spurious faults)."""
    for seed in range(500, 524):
        tj = TinyJob(seed)
        job, keep = tj.make_job(lambda e, o: [t.data for t in oracle.build_tables(e, o, "oracle")])
        rc1, want = oracle.compact(job, "oracle")
        e = jobs.Engine(0)
        rc2, out, msg = e.compact(job)
        e.close()
        assert rc1 == 0 and rc2 == 0, (seed, msg)
        assert summary(out[0]) == summary(want), seed


with open(os.path.join(GOLD, "jobs.json")) as f:
    GOLD_JOBS = json.load(f)


@pytest.mark.parametrize("rec", GOLD_JOBS, ids=lambda r: f"cfg{r['config']}-{r['seed']}-{r['scale']}-{r['options']}")
def test_golden_jobs_on_engine(eng, rec):
    """Inputs built by the engine's table writer, compacted by the engine,
    against the SHA-256 the reference's own code produced."""
    print("job", rec["config"], rec["seed"], rec["scale"], rec["options"], flush=True)
    opts = ffi.default_options(**OPTION_SETS[rec["options"]])
    wl = jobs.Workload(rec["config"], seed=rec["seed"], scale=rec["scale"])
    ins = ([], [])
    for g in wl.groups:
        ins[g.which].extend(t.data for t in jobs.build_group_tables(eng.builder(), opts, g))
    assert [[sha(t) for t in ins[0]], [sha(t) for t in ins[1]]] == rec["inputs"]
    job = jobs.make_job(wl, opts, ins)
    rc, out, msg = eng.compact(job.job)
    assert rc == 0, msg
    got = [{"sha256": sha(t.data), "file_size": len(t.data), "num_entries": t.num_entries,
            "smallest": t.smallest.hex(), "largest": t.largest.hex()} for t in out[0]]
    assert got == rec["outputs"]


@pytest.mark.parametrize("ngp", [5, 2047, 2049, 5000])
def test_many_grandparents_match_oracle(eng, ngp):
    print("grandparents", ngp, flush=True)
    tj = many_grandparents_job(ngp)
    job, keep = tj.make_job(lambda e, o: [t.data for t in oracle.build_tables(e, o, "oracle")])
    rc1, want = oracle.compact(job, "oracle")
    rc2, out, msg = eng.compact(job)
    assert rc1 == 0 and rc2 == 0, msg
    assert len(want) > 10
    assert summary(out[0]) == summary(want)


def oracle_build(e, o):
    return [t.data for t in oracle.build_tables(e, o, "oracle")]


def check_against_oracle(eng, tj, label):
    job, keep = tj.make_job(oracle_build)
    rc1, want = oracle.compact(job, "oracle")
    rc2, out, msg = eng.compact(job)
    assert rc1 == 0 and rc2 == 0, (label, msg)
    assert summary(out[0]) == summary(want), label
    return out


@pytest.mark.parametrize("nl0", [6, 7, 9, 12, 13])
def test_many_children_match_oracle(eng, nl0):
    """Level-0 jobs of 6-13 merge children (3-4 merge rounds)."""
    for seed in range(4):
        print("children", nl0, seed, flush=True)
        out = check_against_oracle(eng, many_children_job(nl0, 100 * nl0 + seed), (nl0, seed))
        assert out[1]["entries_in"] > 0


def test_decoder_ahead_of_the_input_counts():
    """The decoder is launched behind the input scan before the host has read
    its counts, sized like the context's last jobs: used when the counts fit
    (stats decode_ahead 1), launched again after them when they do not (a
    first job, longer keys, a corrupt input, more entries than guessed).
    Every job's tables equal the oracle's."""
    e = jobs.Engine(0)
    try:
        def run(tj, label):
            out = check_against_oracle(e, tj, label)
            return out[1]["decode_ahead"]

        print("decode ahead: repeated job", flush=True)
        assert run(many_children_job(4, 3), "first") == 0
        assert run(many_children_job(4, 3), "again") == 1
        print("decode ahead: longer keys", flush=True)
        assert run(wide_key_job(1, lo=40, hi=64, nl0=3), "longer keys") == 0
        assert run(wide_key_job(1, lo=40, hi=64, nl0=3), "longer keys again") == 1
        print("decode ahead: corrupt input", flush=True)
        tj, (job, keep) = small_job(4)
        t0 = bytearray(tj.tables[0][0])
        bad = bytes(t0[:-1] + bytes([t0[-1] ^ 1]))
        b = C.create_string_buffer(bad, len(bad))
        keep.append(b)
        job.inputs[0][0].data = C.cast(b, C.c_void_p)
        job.inputs[0][0].size = len(bad)
        rc, _, msg = e.compact(job)
        assert rc == ffi.LUDA_CORRUPTION, (rc, msg)
        run(many_children_job(4, 3), "after corrupt")
        print("decode ahead: tiny jobs", flush=True)
        seen = [run(TinyJob(seed), seed) for seed in range(700, 760)]
        assert 0 in seen and 1 in seen, seen
    finally:
        e.close()


@pytest.mark.parametrize("seed", range(4))
def test_wide_keys_match_oracle(eng, seed):
    """User keys of 65-256 bytes: key planes of 9-32 words (the grouped
    decoder and the untiled merge), six merge children."""
    print("wide keys", seed, flush=True)
    out = check_against_oracle(eng, wide_key_job(seed), seed)
    assert out[1]["key_width"] > 64


@pytest.mark.parametrize("seed", range(3))
def test_keys_of_257_to_2048_bytes_match_oracle(eng, seed):
    """User keys of 257-2048 bytes: planes of up to 256 words, the grouped
    decoder with more than 64 KiB of LDS."""
    print("very wide keys", seed, flush=True)
    out = check_against_oracle(eng, wide_key_job(seed, lo=257, hi=2048, nl0=3), seed)
    assert out[1]["key_width"] > 256


def test_key_over_2048_bytes_is_left_to_the_cpu(eng):
    tj = wide_key_job(9, lo=1900, hi=2048)
    long_key = ikey(b"\xfe" * 2049, 10 ** 6, 1)  # sorts after every other key
    tj.files[0][0].append((long_key, b"v"))
    job, keep = tj.make_job(oracle_build)
    assert max(len(k) - 8 for f in tj.files[0] + tj.files[1] for k, _ in f) == 2049
    assert oracle.compact(job, "oracle")[0] == 0
    rc, out, msg = eng.compact(job)
    assert rc == ffi.LUDA_NOT_SUPPORTED, msg


@pytest.mark.parametrize("block", range(2))
def test_malformed_tags_match_oracle(eng, block):
    """Entries whose tag type is above kTypeValue: kept, and the drop rule's
    state is reset (db/db_impl.cc:1113-1119)."""
    for seed in range(block * 60, (block + 1) * 60):
        check_against_oracle(eng, malformed_job(seed), seed)


@pytest.mark.parametrize("interval", [17, 31, 100, 127, 128, 200, 255])
def test_restart_intervals_match_oracle(eng, interval):
    """Restart groups of up to 255 entries in inputs and outputs."""
    print("interval", interval, flush=True)
    check_against_oracle(eng, interval_job(interval, interval), interval)


@pytest.mark.parametrize("interval", [256, 1000])
def test_count_spilling_into_restart_offset_is_corruption(eng, interval):
    """A restart group of 256+ entries spills its count into the offset bits
    (table/block_builder.cc:62-68,89-94); the fork's reader then fails with
    Corruption (oracle/_ref agrees), and so does the engine."""
    tj = interval_job(interval, interval)
    job, keep = tj.make_job(oracle_build)
    rc1, _ = oracle.compact(job, "oracle")
    rc2, out, msg = eng.compact(job)
    assert rc1 == ffi.LUDA_CORRUPTION and rc2 == ffi.LUDA_CORRUPTION, msg


def small_job(seed=3):
    tj = many_children_job(4, seed)
    return tj, tj.make_job(oracle_build)


def test_error_statuses(eng):
    """Each status the caller maps onto leveldb::Status, with the context
    usable afterwards."""
    tj, (job, keep) = small_job()
    # compression other than kNoCompression -> NOT_SUPPORTED
    job.options.compression = 1
    rc, _, msg = eng.compact(job)
    assert rc == ffi.LUDA_NOT_SUPPORTED, msg
    job.options.compression = 0
    # block_restart_interval 0 -> INVALID_ARGUMENT; level out of range too
    job.options.block_restart_interval = 0
    assert eng.compact(job)[0] == ffi.LUDA_INVALID_ARGUMENT
    job.options.block_restart_interval = 16
    job.level = 7
    assert eng.compact(job)[0] == ffi.LUDA_INVALID_ARGUMENT
    job.level = 0
    # still works
    rc, out, msg = eng.compact(job)
    assert rc == 0, msg
    rc1, want = oracle.compact(job, "oracle")
    assert summary(out[0]) == summary(want)


def test_corrupt_inputs_are_corruption(eng):
    """A broken footer magic, an index handle past the end and a truncated
    table are CORRUPTION (Table::Open, table/table.cc:38-82; the caller then
    runs its CPU path)."""
    tj, (job, keep) = small_job(4)
    t0 = bytearray(tj.tables[0][0])
    cases = {
        "magic": t0[:-1] + bytes([t0[-1] ^ 1]),
        "truncated": t0[:-20],
        "index_handle": t0[:-48] + bytes([0xff] * 8) + t0[-40:],
    }
    for name, bad in cases.items():
        job2, keep2 = tj.make_job(oracle_build)
        b = C.create_string_buffer(bytes(bad), len(bad))
        keep2.append(b)
        job2.inputs[0][0].data = C.cast(b, C.c_void_p)
        job2.inputs[0][0].size = len(bad)
        rc, _, msg = eng.compact(job2)
        assert rc == ffi.LUDA_CORRUPTION, (name, rc, msg)
        rc1, _ = oracle.compact(job2, "oracle")
        assert rc1 == ffi.LUDA_CORRUPTION, name


def test_job_over_budget_is_not_supported():
    """A job whose working memory exceeds the context's HBM budget is left to
    the CPU path."""
    tj, (job, keep) = small_job(5)
    e = jobs.Engine(0, budget=1 << 16)
    try:
        rc, _, msg = e.compact(job)
        assert rc == ffi.LUDA_NOT_SUPPORTED and "budget" in msg, msg
    finally:
        e.close()
