"""Writes the golden fixtures under tests/golden/ from the reference's own
CPU code (oracle/_ref/libluda_ref.so, built from the reference sources by
oracle/Makefile).  Run in the build container, where the reference tree
exists:

    python tests/golden/make_golden.py            (small jobs and corner tables)
    python tests/golden/make_golden.py --full     (configs 0-2 at full size)
    python tests/golden/make_golden.py --mix      (the 64-job mixed batch)

Outputs (data only; no reference source is copied):
  jobs.json       SHA-256 of every input and output table, output metadata,
                  for small seeded instances of each workload shape
  tables/*.ldb    tiny tables covering block-format corner cases, with the
                  (key, value) entries they were built from in tables.json
  full_jobs.json  configs 0, 1 and 2 at full size: SHA-256 and metadata of
                  every output table
  mix_jobs.json   one digest over each mixed-batch job's outputs
"""
import hashlib
import json
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import ctypes as C  # noqa: E402

import oracle  # noqa: E402
from luda_amd import ffi, jobs  # noqa: E402

# (workload shape, seed, scale): small enough that the whole suite of CPU
# checks runs in seconds.
JOB_CASES = [
    (0, 0x4C554441, 0.05),
    (0, 7, 0.2),
    (1, 0x4C554442, 0.01),
    (1, 11, 0.03),
    (2, 0x4C554443, 0.004),
    (2, 13, 0.01),
    (5, 0x4C554446, 0.2),
    (5, 17, 0.5),
]

OPTION_SETS = {
    "default": dict(),
    "small_blocks": dict(block_size=256, block_restart_interval=4, max_file_size=1 << 20),
    "no_filter": dict(bloom_bits_per_key=0),
}


def sha(b):
    return hashlib.sha256(bytes(b)).hexdigest()


def job_record(cfg, seed, scale, optname):
    opts = ffi.default_options(**OPTION_SETS[optname])
    wl = jobs.Workload(cfg, seed=seed, scale=scale)
    ins = ([], [])
    for g in wl.groups:
        ins[g.which].extend(t.data for t in jobs.build_group_tables(oracle.builder("ref"), opts, g))
    job = jobs.make_job(wl, opts, ins)
    rc, outs = oracle.compact(job.job, "ref")
    assert rc == 0
    return {
        "config": cfg, "seed": seed, "scale": scale, "options": optname,
        "inputs": [[sha(t) for t in ins[0]], [sha(t) for t in ins[1]]],
        "outputs": [{"sha256": sha(t.data), "file_size": len(t.data), "num_entries": t.num_entries,
                     "smallest": t.smallest.hex(), "largest": t.largest.hex()} for t in outs],
    }


def ikey(user, seq, t):
    return user + struct.pack("<Q", (seq << 8) | t)


def corner_tables():
    """Entry lists that exercise block-format corners."""
    cases = {}
    # restart groups of 1 and of 16, a 2-byte value-length varint, deletions
    cases["groups_and_varints"] = ([(ikey(b"k%05d" % i, 1000 - i, 1 if i % 7 else 0), b"v" * (i % 300))
                                    for i in range(40)], dict(block_restart_interval=16))
    cases["restart_every_entry"] = ([(ikey(b"key%03d" % i, 5, 1), b"x" * 10) for i in range(30)],
                                    dict(block_restart_interval=1))
    # 0xff user keys: the successor of the last key cannot shorten
    cases["ff_keys"] = ([(ikey(b"\xff" * n, 9, 1), b"y" * 50) for n in range(1, 20)], dict())
    # many versions of one user key: separators cannot shorten across blocks
    cases["versions"] = ([(ikey(b"same", 10000 - i, 1), b"z" * 200) for i in range(60)], dict())
    # shortened separators between decimal keys; several 2 KB filter windows
    cases["decimal"] = ([(ikey(b"%016d" % (i * 7919), i + 1, 1), bytes((i * j) & 0xff for j in range(100)))
                         for i in range(300)], dict())
    # a block that ends exactly at block_size: 1 entry of 4096 - 4 - 4 - 3 - 8 - 1 value bytes
    k = ikey(b"a", 1, 1)
    vlen = 4096 - 4 - 4 - 3 - len(k)
    cases["exact_block"] = ([(k, b"e" * (vlen - 1)), (ikey(b"b", 1, 1), b"f" * 10)],
                            dict(block_restart_interval=16))
    # block_size 256: many small blocks, so some 2 KB filter windows are empty
    # of block starts and others hold several blocks
    cases["small_blocks"] = ([(ikey(b"s%04d" % i, 3, 1), b"q" * (i % 40)) for i in range(400)],
                             dict(block_size=256))
    # no filter policy: empty metaindex block
    cases["no_filter"] = ([(ikey(b"nf%03d" % i, 2, 1), b"w" * 30) for i in range(50)],
                          dict(bloom_bits_per_key=0))
    return cases


def build_ref_tables(entries, opts):
    keys = b"".join(k for k, _ in entries)
    vals = b"".join(v for _, v in entries)
    ko = (C.c_uint64 * (len(entries) + 1))()
    vo = (C.c_uint64 * (len(entries) + 1))()
    a = b = 0
    for i, (k, v) in enumerate(entries):
        ko[i], vo[i] = a, b
        a += len(k)
        b += len(v)
    ko[len(entries)], vo[len(entries)] = a, b
    kb = C.create_string_buffer(keys, len(keys) + 1)
    vb = C.create_string_buffer(vals, len(vals) + 1)
    return oracle.builder("ref")(C.byref(opts), len(entries), C.cast(kb, C.c_void_p), C.cast(ko, C.c_void_p),
                                 C.cast(vb, C.c_void_p), C.cast(vo, C.c_void_p))


def full_record(cfg):
    """Config cfg at its full size with its default seed: every output's
    SHA-256 and metadata, and the digest over them."""
    opts = ffi.default_options()
    wl = jobs.Workload(cfg)
    ins = ([], [])
    for g in wl.groups:
        ins[g.which].extend(t.data for t in jobs.build_group_tables(oracle.builder("ref"), opts, g))
    job = jobs.make_job(wl, opts, ins)
    rc, outs = oracle.compact(job.job, "ref")
    assert rc == 0
    return {
        "config": cfg, "seed": 0x4C554441 + cfg, "scale": 1.0, "input_bytes": job.bytes_in,
        "inputs_digest": sha(b"".join(hashlib.sha256(t).digest() for t in ins[0] + ins[1])),
        "digest": jobs.output_digest(outs),
        "outputs": [{"sha256": sha(t.data), "file_size": len(t.data), "num_entries": t.num_entries,
                     "smallest": t.smallest.hex(), "largest": t.largest.hex()} for t in outs],
    }


def mix_records(n=64, seed=3):
    """Every job of the mixed batch (luda_amd.jobs.mix_specs) compacted by the
    reference: one output digest per job."""
    opts = ffi.default_options()
    recs = []
    for k, (shape, wseed, scale, mib) in enumerate(jobs.mix_specs(n, seed)):
        wl = jobs.Workload(shape, seed=wseed, scale=scale)
        ins = ([], [])
        for g in wl.groups:
            ins[g.which].extend(t.data for t in jobs.build_group_tables(oracle.builder("ref"), opts, g))
        job = jobs.make_job(wl, opts, ins)
        rc, outs = oracle.compact(job.job, "ref")
        assert rc == 0
        recs.append({"shape": shape, "seed": wseed, "scale": scale, "input_bytes": job.bytes_in,
                     "files": [len(ins[0]), len(ins[1])], "outputs": len(outs), "digest": jobs.output_digest(outs)})
        print("mix job", k, shape, round(mib, 1), "MiB", len(outs), "outputs", flush=True)
        wl.close()
    return {"jobs": n, "seed": seed, "records": recs}


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libluda_ref.so is not built (needs the reference tree)")
    if "--full" in sys.argv:
        recs = [full_record(c) for c in (0, 1, 2)]
        with open(os.path.join(HERE, "full_jobs.json"), "w") as f:
            json.dump(recs, f, indent=0)
        return
    if "--mix" in sys.argv:
        with open(os.path.join(HERE, "mix_jobs.json"), "w") as f:
            json.dump(mix_records(), f, indent=0)
        return
    recs = []
    for cfg, seed, scale in JOB_CASES:
        for optname in OPTION_SETS:
            if optname != "default" and cfg not in (0, 5):
                continue
            recs.append(job_record(cfg, seed, scale, optname))
            print("job", cfg, seed, scale, optname, len(recs[-1]["outputs"]), "outputs")
    with open(os.path.join(HERE, "jobs.json"), "w") as f:
        json.dump(recs, f, indent=1)
    tdir = os.path.join(HERE, "tables")
    os.makedirs(tdir, exist_ok=True)
    meta = {}
    for name, (entries, kw) in corner_tables().items():
        opts = ffi.default_options(max_file_size=ffi.U64_MAX, **kw)
        tables = build_ref_tables(entries, opts)
        assert len(tables) == 1
        data = tables[0].data
        assert len(data) <= 64 * 1024
        with open(os.path.join(tdir, name + ".ldb"), "wb") as f:
            f.write(data)
        meta[name] = {"options": {k: getattr(opts, k) for k, _ in ffi.luda_options._fields_},
                      "entries": [[k.hex(), v.hex()] for k, v in entries], "sha256": sha(data)}
        print("table", name, len(data), "bytes")
    with open(os.path.join(HERE, "tables.json"), "w") as f:
        json.dump(meta, f, indent=0)


if __name__ == "__main__":
    main()
