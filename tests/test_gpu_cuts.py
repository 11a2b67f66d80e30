"""The tiled output block cuts (k_cuts.hip) on the GPU: the same jobs with the
tiled cuts and with the orbit kernels they replace (LUDA_BLOCK_CUTS=orbit)
give identical tables, equal to the oracle's; kernel timing
(LUDA_KERNEL_TIMING=cut_tiles) shows which path a job took, including the
fallback of a job whose blocks hold more entries than a tile takes."""
import contextlib
import os
import random

import pytest

import oracle
from luda_amd import ffi, jobs
from tests.tiny_jobs import TinyJob, ikey, many_grandparents_job

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = jobs.Engine(0)
    yield e
    e.close()


@contextlib.contextmanager
def env(**kv):
    old = {k: os.environ.get(k) for k in kv}
    os.environ.update({k: v for k, v in kv.items()})
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def tables_of(out):
    return [(t.data, t.smallest, t.largest, t.num_entries) for t in out[0]]


def build_both(eng, ents, opts):
    """(tiled tables, orbit tables, tiled launches of cut_tiles)."""
    with env(LUDA_KERNEL_TIMING="cut_tiles"):
        rc, out, msg = eng.build_entries(ents, opts)
    assert rc == 0, msg
    with env(LUDA_BLOCK_CUTS="orbit"):
        rc2, out2, msg2 = eng.build_entries(ents, opts)
    assert rc2 == 0, msg2
    st = out[1]
    launches = st["top_kernel_launches"] if st["top_kernel"] == "cut_tiles" else 0
    return tables_of(out), tables_of(out2), launches


def oracle_tables(ents, opts):
    return [(t.data, t.smallest, t.largest, t.num_entries) for t in oracle.build_tables(ents, opts, "oracle")]


@pytest.mark.parametrize("n,vmin,vmax,interval,block_size,max_file",
                         [(30000, 100, 100, 16, 4096, 2 << 20),   # fixed-size entries, many tiles
                          (20000, 0, 400, 16, 4096, 1 << 20),     # varied, several tables
                          (12000, 50, 150, 1, 4096, 300000),      # every entry a restart point
                          (9000, 10, 40, 64, 1024, 100000),       # long restart runs, small blocks
                          (4097, 100, 100, 16, 4096, 2 << 20),    # one entry past a tile
                          (8192, 5, 20, 3, 1024, 50000)])         # ~20 entries per block
def test_tiled_cuts_equal_orbit_and_oracle(eng, n, vmin, vmax, interval, block_size, max_file):
    rng = random.Random(n * 31 + interval)
    keys = sorted({b"%016d" % rng.randrange(10 ** 12) for _ in range(n)})
    ents = [(ikey(k, 1000 + i, 1), bytes(rng.randrange(256) for _ in range(rng.randint(vmin, vmax))))
            for i, k in enumerate(keys)]
    opts = ffi.default_options(block_size=block_size, block_restart_interval=interval, max_file_size=max_file)
    print("cuts", n, vmin, vmax, interval, block_size, max_file, flush=True)
    tiled, orbit, launches = build_both(eng, ents, opts)
    assert launches == 1  # the tiled cuts ran, once
    assert tiled == orbit
    assert tiled == oracle_tables(ents, opts)


def test_blocks_of_many_entries_fall_back_to_the_orbit(eng):
    """Entries of ~12 bytes in 8 KiB blocks: more entry points per tile than
    the tiles take, so the job runs again on the orbit kernels."""
    rng = random.Random(5)
    ents = [(ikey(b"%06d" % i, 7, 1), bytes(rng.randrange(256) for _ in range(rng.randint(0, 3))))
            for i in range(12000)]
    opts = ffi.default_options(block_size=8192, max_file_size=1 << 20)
    with env(LUDA_KERNEL_TIMING="cut_tiles"):
        rc, out, msg = eng.build_entries(ents, opts)
    assert rc == 0, msg
    assert out[1]["attempts"] >= 2
    assert tables_of(out) == oracle_tables(ents, opts)


@pytest.mark.parametrize("seed", range(6))
def test_compactions_with_table_cuts_equal_orbit(eng, seed):
    """Compactions whose ShouldStopBefore cuts start tables (forced block
    starts in the kept stream), tiled against orbit and against the oracle."""
    tj = many_grandparents_job(5 + 700 * seed) if seed % 2 else TinyJob(900 + seed)
    job, keep = tj.make_job(lambda e, o: [t.data for t in oracle.build_tables(e, o, "oracle")])
    rc1, want = oracle.compact(job, "oracle")
    rc2, out, msg = eng.compact(job)
    assert rc1 == 0 and rc2 == 0, msg
    with env(LUDA_BLOCK_CUTS="orbit"):
        rc3, out3, msg3 = eng.compact(job)
    assert rc3 == 0, msg3
    got = [(t.data, t.smallest, t.largest, t.num_entries) for t in out[0]]
    assert got == [(t.data, t.smallest, t.largest, t.num_entries) for t in out3[0]]
    assert got == [(t.data, t.smallest, t.largest, t.num_entries) for t in want]


@pytest.mark.parametrize("max_file", [1 << 20, 1 << 30])
def test_tables_longer_than_the_staged_span(eng, max_file):
    """Tables of more than 8192 blocks (~45-byte entries in 64-byte blocks:
    ~10 K blocks to a 1 MiB table, or one table of ~25 K blocks when
    max_file_size exceeds the output) end past k_table_next's staged span, so
    the chain follows them over the tiles' totals."""
    rng = random.Random(11)
    ents = [(ikey(b"%016d" % i, 5 + i, 1), rng.randbytes(20)) for i in range(50000)]
    opts = ffi.default_options(block_size=64, max_file_size=max_file)
    rc, out, msg = eng.build_entries(ents, opts)
    assert rc == 0, msg
    assert tables_of(out) == oracle_tables(ents, opts)


def test_long_tables_cut_by_should_stop_before(eng):
    """ShouldStopBefore cuts (grandparents of 30 MB, over the 10 MB overlap
    limit) every 20000 keys (~10 K blocks), in a stream whose size limit
    alone makes tables of ~10 K blocks: forced and size ends past the staged
    span."""
    rng = random.Random(12)
    users = [b"g%07d" % i for i in range(100000)]
    a, b = [], []
    for i, u in enumerate(users):
        (a if rng.random() < 0.3 else b).append((ikey(u, i + 1, 1), rng.randbytes(20)))
    tj = TinyJob(0)
    tj.level = 1
    tj.opts = ffi.default_options(block_size=64, max_file_size=1 << 20)
    tj.files = ([a], [b])
    tj.snapshot = len(users)
    tj.grandparents = [(ikey(users[i] + b"~", 7, 1), 30 << 20) for i in range(7000, 100000, 20000)]
    tj.deeper = []
    job, keep = tj.make_job(lambda e, o: [t.data for t in oracle.build_tables(e, o, "oracle")])
    rc1, want = oracle.compact(job, "oracle")
    rc2, out, msg = eng.compact(job)
    assert rc1 == 0 and rc2 == 0, msg
    assert len(want) >= 5
    got = [(t.data, t.smallest, t.largest, t.num_entries) for t in out[0]]
    assert got == [(t.data, t.smallest, t.largest, t.num_entries) for t in want]
