"""The batched / served form (luda_submit, luda_poll, luda_drain): jobs of
different shapes and sizes queued together on several worker slots give the
same tables as the same jobs run alone through luda_compact, and the same as
the oracle for the small ones; a device-memory budget that admits one job
at a time changes nothing."""
import hashlib
import json
import os

import pytest

import oracle
from luda_amd import ffi, jobs
from tests.tiny_jobs import TinyJob

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def summary(ts):
    return [(hashlib.sha256(t.data).hexdigest(), len(t.data), t.smallest, t.largest, t.num_entries) for t in ts]


def make_mix(builder_eng):
    """Tiny jobs (oracle-checked) plus workload jobs of configs 0, 1, 2, 5."""
    mix = []
    for seed in range(24):
        tj = TinyJob(500 + seed)
        job, keep = tj.make_job(lambda e, o: [t.data for t in oracle.build_tables(e, o, "oracle")])
        rc, want = oracle.compact(job, "oracle")
        assert rc == 0
        mix.append(("tiny", job, keep, summary(want)))
    opts = ffi.default_options()
    for cfg, scale in [(0, 0.05), (1, 0.01), (2, 0.004), (5, 0.2), (1, 0.03)]:
        wl = jobs.Workload(cfg, scale=scale)
        ins = ([], [])
        for g in wl.groups:
            ins[g.which].extend(t.data for t in jobs.build_group_tables(builder_eng.builder(), opts, g))
        j = jobs.make_job(wl, opts, ins)
        mix.append((f"cfg{cfg}", j.job, (wl, j), None))
    return mix


@pytest.fixture(scope="module")
def mix():
    e = jobs.Engine(0)
    m = make_mix(e)
    # each job alone first
    for i, (name, job, keep, want) in enumerate(m):
        rc, out, msg = e.compact(job)
        assert rc == 0, (name, msg)
        got = summary(out[0])
        if want is not None:
            assert got == want, name
        m[i] = (name, job, keep, got)
    e.close()
    return m


def run_batched(mix, slots, budget):
    e = jobs.Engine(0, budget=budget)
    e.set_slots(slots)
    got = {}
    for i, (name, job, keep, want) in enumerate(mix):
        e.submit(job, lambda rc, out, msg, i=i: got.__setitem__(i, (rc, summary(out[0]) if out else None, msg)))
    e.poll()
    n = e.drain()
    e.close()
    return got, n


# 100 MiB: every job fits alone, but the admission estimate (64 MiB or more
# per job) lets only one run at a time.  12 slots: many streams' scans side
# by side, whose look-backs take their tiles by ticket (an 8-slot mix once
# failed jobs when tiles followed workgroup order)
@pytest.mark.parametrize("slots,budget", [(3, 0), (2, 0), (3, 100 << 20), (12, 0)])
def test_batched_matches_alone(mix, slots, budget):
    print("batched", slots, budget, len(mix), flush=True)
    got, n = run_batched(mix, slots, budget)
    assert len(got) == len(mix)
    for i, (name, job, keep, want) in enumerate(mix):
        rc, s, msg = got[i]
        assert rc == 0, (name, msg)
        assert s == want, name
