"""The benchmark configurations at their stated sizes against the
reference's own code: configs 0 (20 MiB L0->L1), 1 (256 MiB L1->L2) and 2
(512 MiB bottommost) table by table, and the jobs of the 64-job mixed batch
by output digest (tests/golden/full_jobs.json and mix_jobs.json, written by
tests/golden/make_golden.py from oracle/_ref).  Inputs are built on the GPU
by luda_build_tables and stay in HBM; their bytes are checked against the
reference's inputs first."""
import hashlib
import json
import os

import pytest

from luda_amd import ffi, jobs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "full_jobs.json")) as f:
    FULL = json.load(f)
with open(os.path.join(HERE, "golden", "mix_jobs.json")) as f:
    MIX = json.load(f)


@pytest.fixture(scope="module")
def eng():
    e = jobs.Engine(0)
    yield e
    e.close()


def compact_resident(eng, wl, opts):
    dev, keep = jobs.build_device_inputs(eng, wl, opts)
    try:
        digest_in = hashlib.sha256(b"".join(hashlib.sha256(eng.d2h(p, n)).digest()
                                            for which in dev for p, n in which)).hexdigest()
        job = jobs.make_job(wl, opts, None, flags=ffi.LUDA_JOB_KEEP_ON_DEVICE, device_inputs=dev)
        rc, out, msg = eng.compact(job.job, keep_on_device=True)
        assert rc == ffi.LUDA_OK, msg
        try:
            tables = jobs.device_result_tables(eng, out[2])
        finally:
            eng.free(out[2])
        return digest_in, tables, out[1]
    finally:
        for r in keep:
            eng.free(r)


@pytest.mark.parametrize("rec", FULL, ids=lambda r: f"config{r['config']}")
def test_full_size_config_matches_reference(eng, rec):
    print("config", rec["config"], "full size", flush=True)
    wl = jobs.Workload(rec["config"], seed=rec["seed"], scale=rec["scale"])
    digest_in, tables, stats = compact_resident(eng, wl, ffi.default_options())
    wl.close()
    assert stats["bytes_read"] == rec["input_bytes"]
    assert digest_in == rec["inputs_digest"]
    got = [{"sha256": hashlib.sha256(t.data).hexdigest(), "file_size": len(t.data), "num_entries": t.num_entries,
            "smallest": t.smallest.hex(), "largest": t.largest.hex()} for t in tables]
    assert len(got) == len(rec["outputs"])
    for i, (g, w) in enumerate(zip(got, rec["outputs"])):
        assert g == w, f"output {i}"
    assert jobs.output_digest(tables) == rec["digest"]


SMALL_MIX = [(k, r) for k, r in enumerate(MIX["records"]) if r["input_bytes"] <= (48 << 20)]


@pytest.mark.parametrize("k,rec", SMALL_MIX, ids=lambda x: str(x) if isinstance(x, int) else f"shape{x['shape']}")
def test_mix_job_matches_reference(eng, k, rec):
    """The mixed batch's jobs of up to 48 MiB (L0->L1 jobs with 4-12 level-0
    tables among them); bench.py --mix checks all 64."""
    print("mix job", k, rec["shape"], rec["input_bytes"], flush=True)
    wl = jobs.Workload(rec["shape"], seed=rec["seed"], scale=rec["scale"])
    _, tables, stats = compact_resident(eng, wl, ffi.default_options())
    wl.close()
    assert stats["bytes_read"] == rec["input_bytes"]
    assert len(tables) == rec["outputs"]
    assert jobs.output_digest(tables) == rec["digest"]
