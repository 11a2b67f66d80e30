"""Brute-force differential test of the oracle against the reference's own
code (oracle/_ref), on seeded tiny jobs that reach every branch of the
compaction loop.  Skipped where the reference could not be built."""
import pytest

import oracle
from tests.tiny_jobs import (TinyJob, interval_job, malformed_job, many_children_job, many_grandparents_job,
                             wide_key_job)
from luda_amd import ffi

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")


def build_with(which):
    return lambda ents, opts: [t.data for t in oracle.build_tables(ents, opts, which)]


@pytest.mark.parametrize("block", range(8))
def test_tiny_jobs_oracle_equals_reference(block):
    for seed in range(block * 250, (block + 1) * 250):
        tj = TinyJob(seed)
        job, keep = tj.make_job(build_with("ref"))
        # inputs built by the two builders are identical
        for which in (0, 1):
            for ents in tj.files[which]:
                one = tj.opts
                a = oracle.build_tables(ents, one, "oracle")
                b = oracle.build_tables(ents, one, "ref")
                assert [t.data for t in a] == [t.data for t in b], seed
        rc1, o1 = oracle.compact(job, "oracle")
        rc2, o2 = oracle.compact(job, "ref")
        assert rc1 == rc2 == 0, seed
        assert [(t.data, t.smallest, t.largest, t.num_entries) for t in o1] == \
               [(t.data, t.smallest, t.largest, t.num_entries) for t in o2], seed


@pytest.mark.parametrize("ngp", [5, 2049, 5000])
def test_many_grandparents_oracle_equals_reference(ngp):
    """ShouldStopBefore over thousands of small grandparents (cuts every few)."""
    tj = many_grandparents_job(ngp)
    job, keep = tj.make_job(build_with("ref"))
    rc1, o1 = oracle.compact(job, "oracle")
    rc2, o2 = oracle.compact(job, "ref")
    assert rc1 == rc2 == 0
    assert len(o1) > 10
    assert [(t.data, t.smallest, t.largest, t.num_entries) for t in o1] == \
           [(t.data, t.smallest, t.largest, t.num_entries) for t in o2]


def both(tj):
    job, keep = tj.make_job(build_with("ref"))
    rc1, o1 = oracle.compact(job, "oracle")
    rc2, o2 = oracle.compact(job, "ref")
    return (rc1, [(t.data, t.smallest, t.largest, t.num_entries) for t in o1 or []]), \
           (rc2, [(t.data, t.smallest, t.largest, t.num_entries) for t in o2 or []])


@pytest.mark.parametrize("nl0", [6, 7, 9, 12, 13])
def test_many_children_oracle_equals_reference(nl0):
    for seed in range(4):
        a, b = both(many_children_job(nl0, 100 * nl0 + seed))
        assert a[0] == 0 and a == b, (nl0, seed)


@pytest.mark.parametrize("seed", range(4))
def test_wide_keys_oracle_equals_reference(seed):
    a, b = both(wide_key_job(seed))
    assert a[0] == 0 and a == b


def test_malformed_tags_oracle_equals_reference():
    for seed in range(300):
        a, b = both(malformed_job(seed))
        assert a[0] == 0 and a == b, seed


@pytest.mark.parametrize("interval", [17, 31, 100, 127, 128, 200, 255, 256, 1000])
def test_restart_intervals_oracle_equals_reference(interval):
    """Up to 255, compactions agree byte for byte; from 256 on the first
    restart group's count spills into the offset bits and both readers report
    Corruption."""
    a, b = both(interval_job(interval, interval))
    assert a == b
    assert a[0] == (0 if interval <= 255 else ffi.LUDA_CORRUPTION)


@pytest.mark.parametrize("interval", [17, 64, 127, 128, 255, 256, 300, 1000])
def test_restart_word_overflow_bytes_equal_reference(interval):
    """The fork's (offset << 8) | count restart words (table/block_builder.cc:
    62-68, 89-94), counts past 255 included, written the same way by the
    oracle and by the reference's BlockBuilder."""
    from tests.tiny_jobs import ikey
    ents = [(ikey(b"r%06d" % i, 10 + i, 1), b"v" * (i % 3)) for i in range(3000)]
    opts = ffi.default_options(block_size=8192, block_restart_interval=interval, max_file_size=ffi.U64_MAX)
    a = oracle.build_tables(ents, opts, "oracle")
    b = oracle.build_tables(ents, opts, "ref")
    assert [t.data for t in a] == [t.data for t in b]
