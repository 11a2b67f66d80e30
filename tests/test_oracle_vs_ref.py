"""Brute-force differential test of the oracle against the reference's own
code (oracle/_ref), on seeded tiny jobs that reach every branch of the
compaction loop.  Skipped where the reference could not be built."""
import pytest

import oracle
from tests.tiny_jobs import TinyJob, many_grandparents_job

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
