#!/usr/bin/env python3
"""Compaction throughput of the MI355X engine on one GPU.

Workload (default): the L1->L2 compaction of one 25 MiB L1 table and ~115
overlapping 2 MiB L2 tables, 256 MiB in all (~2.25 M entries of 16-byte
decimal keys uniform over 1e8 and 100-byte values; 10% of L1 keys overwrite
L2 keys, 3% of L1 entries are tombstones that stay because an L3 file covers
the key space).  Inputs are built by the engine's own table writer from a
seeded generator, so they are valid tables of the LevelDB fork, and stay
resident in HBM: the timed region is luda_compact() from inputs in HBM to
outputs in HBM, one compaction per step.

Printed: one JSON line on stdout (progress goes to stderr).  `value` is input
GB/s of whole compactions; `roofline` is the dominant kernel's algorithmic
bytes per launch over its average launch time, timed with HIP events on the
engine's stream during the timed steps; `cpu_baseline` is the CPU compaction
(the reference's own code when oracle/_ref was built, else the C
restatement) timed on this host on the same job; `parity` says whether the
engine's output tables hash the same as the CPU compaction's.
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import shutil
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The batched and served forms run one HIP stream per worker slot.  HIP maps
# a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 unless set),
# and streams that share a queue run their kernels in series, so these forms
# ask for at least 16; the variable is read when the HIP runtime starts.
if any(a.startswith(("--mix", "--serve")) for a in sys.argv[1:]):
    if int(os.environ.get("GPU_MAX_HW_QUEUES") or 0) < 16:
        os.environ["GPU_MAX_HW_QUEUES"] = "16"

from luda_amd import ffi, jobs  # noqa: E402

METRIC = "compaction input GB/s (byte-identical SSTables) vs CPU; p50/p99 job latency"
HBM_PEAK_GBPS = 8000.0
# HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled
# by a plain copy of known size) for the default workload; see the .md beside it.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_l1l2_256MiB.json")


def measured_traffic(kernel, config, scale):
    """HBM bytes per launch of `kernel` from the committed counter run, for
    the workload that run measured (config 1 at full size), else None."""
    if config != 1 or scale != 1.0 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        k = json.load(f).get("kernels", {}).get(kernel)
    return k["bytes"] if k else None

WORKLOADS = {
    0: "L0->L1 compaction: 4 overlapping L0 tables + 6 L1 tables, ~20 MiB, 16-byte keys mod 1e6, "
       "100-byte values, 20% overwrites, 2% deletes",
    1: "L1->L2 compaction: 1 L1 table + ~115 L2 tables, 256 MiB, 16-byte keys uniform over 1e8, "
       "100-byte values, 10% overwrites, 3% tombstones kept (not bottommost)",
    2: "bottommost L1->L2 compaction, 512 MiB, Zipf(0.99) keys of 8-64 bytes over 1e7 ids, 16-1024-byte "
       "values, 5% deletes, snapshots at 25/50/75% of sequence numbers",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sha(b):
    return hashlib.sha256(b).hexdigest()


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * p / 100.0
    f = int(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def build_device_inputs(eng, wl, opts):
    return jobs.build_device_inputs(eng, wl, opts)


def copy_probe(nbytes):
    """A plain copy of nbytes between two HBM buffers (the runtime's blit
    kernel), three times: the reference kernel for HBM byte counters."""
    import torch
    x = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def host_info():
    """The CPU the baseline ran on: logical CPUs the OS reports, CPUs this
    process may use, and the model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def mix_jobs(eng, opts, n, seed):
    """The mixed batch (luda_amd.jobs.mix_specs): n compaction jobs of
    log-uniform input sizes in [4, 256] MiB, half L0->L1 with 4-12 level-0
    tables, half Ln->Ln+1, half with the 256 MiB job's key/value distribution
    and half with the skewed 512 MiB job's; inputs built on the GPU and
    resident in HBM."""
    out = []
    for shape, wseed, scale, _mib in jobs.mix_specs(n, seed):
        wl = jobs.Workload(shape, seed=wseed, scale=scale)
        dev, keep = build_device_inputs(eng, wl, opts)
        job = jobs.make_job(wl, opts, None, flags=ffi.LUDA_JOB_KEEP_ON_DEVICE, device_inputs=dev)
        out.append((shape, sum(s for which in dev for _, s in which), job, keep, wl))
    return out


def run_mix(args):
    """Batched form against the plain one-job-at-a-time path on the same jobs:
    aggregate input GB/s and job latency (from the moment all jobs are
    queued), plus whether every batched result equals the plain one."""
    opts = ffi.default_options()
    eng = jobs.Engine(0)
    t = time.time()
    mix = mix_jobs(eng, opts, args.mix, args.mix_seed)
    total = sum(b for _, b, _, _, _ in mix)
    log(f"{len(mix)} jobs, {total / 2**30:.2f} GiB of inputs built in {time.time() - t:.1f}s")

    def summary(res):
        return jobs.device_result_tables(eng, res)
    tuning = None
    slots = args.slots
    if args.tune:
        # the autotuner measures on a different mix (another seed), then the
        # batched form below runs with what it chose
        from luda_amd import autotune
        t = time.time()
        tmix = mix_jobs(eng, opts, args.tune, args.mix_seed + 1)
        ttotal = sum(b for _, b, _, _, _ in tmix)
        tuning = autotune.autotune_batched([j.job for _, _, j, _, _ in tmix], ttotal, log=log)
        tuning["sample"] = f"{len(tmix)} jobs of the mix with seed {args.mix_seed + 1}, {ttotal / 2**30:.2f} GiB"
        slots = tuning["slots"]
        log(f"autotune: {tuning['slots']} slots, rules {tuning['rules']} ({tuning['gbps']} GB/s on the sample) "
            f"in {time.time() - t:.1f}s")
        for _, _, _, keep, _ in tmix:
            for res in keep:
                eng.free(res)
    # plain: one luda_compact after another
    for shape, b, job, keep, wl in mix[:2]:  # warm-up
        rc, o, msg = eng.compact(job.job, keep_on_device=True)
        eng.free(o[2])
    plain_hash, plain_lat = [], []
    t0 = time.perf_counter()
    for shape, b, job, keep, wl in mix:
        rc, o, msg = eng.compact(job.job, keep_on_device=True)
        if rc != ffi.LUDA_OK:
            raise RuntimeError(f"luda_compact failed ({rc}): {msg}")
        plain_lat.append((time.perf_counter() - t0) * 1e3)
        plain_hash.append(o[2])
    t_plain = time.perf_counter() - t0
    plain_hash = [jobs.output_digest(summary(res)) for res in plain_hash_free(eng, plain_hash, summary)]
    golden = mix_golden(args.mix, args.mix_seed)
    # batched: everything queued at once on the worker slots
    beng = jobs.Engine(0)
    beng.set_slots(slots)
    if tuning:
        beng.set_tuning(tuning["rules"])
    # warm-up: the largest jobs, one per slot, so that every slot's working
    # memory has its final size before the timed batch
    warm_err = []

    def warm_done(rc, o, msg):
        if rc != ffi.LUDA_OK:
            warm_err.append(f"{rc}: {msg}")
        else:
            beng.free(o[2])

    for shape, b, job, keep, wl in sorted(mix, key=lambda x: -x[1])[:slots]:
        beng.submit(job.job, warm_done, keep_on_device=True)
    beng.drain()
    if warm_err:
        raise RuntimeError(f"batched warm-up job failed ({warm_err[0]})")
    got = {}
    t0 = time.perf_counter()
    for i, (shape, b, job, keep, wl) in enumerate(mix):
        beng.submit(job.job, lambda rc, o, msg, i=i: got.__setitem__(i, (rc, o, (time.perf_counter() - t0) * 1e3)),
                    keep_on_device=True)
    beng.drain()
    t_batch = time.perf_counter() - t0
    same = True
    batch_lat = []
    for i in range(len(mix)):
        rc, o, ms = got[i]
        batch_lat.append(ms)
        same &= rc == ffi.LUDA_OK and jobs.output_digest(summary(o[2])) == plain_hash[i]
        beng.free(o[2])
    cpu = None
    if not args.no_cpu:
        cpu = mix_cpu_baseline(eng, opts, mix, plain_hash)
    line = {
        "metric": "batched compaction input GB/s vs the plain path; p50/p99 job latency",
        "jobs": len(mix), "input_bytes": total, "slots": slots, "tuning": tuning,
        "plain_gbps": round(total / t_plain / 1e9, 3), "batched_gbps": round(total / t_batch / 1e9, 3),
        "speedup": round(t_plain / t_batch, 3),
        "plain_p50_ms": round(percentile(plain_lat, 50), 3), "plain_p99_ms": round(percentile(plain_lat, 99), 3),
        "batched_p50_ms": round(percentile(batch_lat, 50), 3), "batched_p99_ms": round(percentile(batch_lat, 99), 3),
        "same_tables": same,
        "same_as_reference": (plain_hash == golden) if golden else None,
        "cpu_baseline": cpu,
        "workload": "log-uniform 4-256 MiB jobs: half L0->L1 with 4-12 level-0 tables, half Ln->Ln+1; half "
                    "16-byte decimal keys with 100-byte values, half Zipf 8-64-byte keys with 16-1024-byte "
                    "values; inputs in HBM",
    }
    print(json.dumps(line), flush=True)
    beng.close()
    for shape, b, job, keep, wl in mix:
        for res in keep:
            eng.free(res)
    eng.close()


def mix_cpu_baseline(eng, opts, mix, gpu_hashes):
    """The CPU compaction (the reference's own code when oracle/_ref was
    built, else the C restatement) on the mix's jobs, from host copies of
    their inputs: one thread over the jobs in order until 1 GiB of input
    (their output hashes checked against the GPU's), then every job at once,
    one job per thread over the host's cores."""
    import oracle  # CPU baseline leg only
    from concurrent.futures import ThreadPoolExecutor
    kind = "reference" if oracle.ref_available() else "port"
    which = "ref" if kind == "reference" else "oracle"
    hjobs = []
    for shape, b, job, keep, wl in mix:
        host = ([], [])
        for w in (0, 1):
            for i in range(job.job.num_inputs[w]):
                f = job.job.inputs[w][i]
                host[w].append(eng.d2h(f.data, f.size))
        hjobs.append((b, jobs.make_job(wl, opts, host)))

    def run(k):
        rc, tabs = oracle.compact(hjobs[k][1].job, which)
        if rc != 0:
            raise RuntimeError(f"CPU compaction of job {k} failed ({rc})")
        return jobs.output_digest(tabs)
    seq_bytes, seq_same, k = 0, True, 0
    t0 = time.perf_counter()
    while k < len(hjobs) and seq_bytes < (1 << 30):
        seq_same &= run(k) == gpu_hashes[k]
        seq_bytes += hjobs[k][0]
        k += 1
    t_seq = time.perf_counter() - t0
    # every CPU this process may use, up to 16: the host CPUs one GPU's job
    # gets on the measurement machines
    host = host_info()
    threads = max(1, min(16, host["usable_cpus"] or 1))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(run, range(len(hjobs))))
    t_par = time.perf_counter() - t0
    total = sum(b for b, _ in hjobs)
    log(f"CPU ({kind}): {k} jobs one after another {seq_bytes / t_seq / 1e9:.3f} GB/s, "
        f"all {len(hjobs)} on {threads} threads {total / t_par / 1e9:.3f} GB/s")
    return {"kind": kind, "unit": "GB/s", "host": host,
            "sequential": {"value": round(seq_bytes / t_seq / 1e9, 4), "cores": 1, "jobs": k,
                           "input_bytes": seq_bytes, "same_tables_as_gpu": seq_same},
            "parallel": {"value": round(total / t_par / 1e9, 4), "cores": threads, "jobs": len(hjobs),
                         "input_bytes": total},
            "sample": "sequential: the first jobs of the mix up to 1 GiB of input, one thread; parallel: every "
                      "job, one job per thread; inputs and outputs in RAM"}


def release_tree(lsm):
    for level in lsm.levels:
        for f in level:
            f.owner.drop()
        level.clear()


def run_serve(args):
    """The served form on a fillrandom write stream: memtable flushes and the
    leveled compactions they cause, tables resident in HBM from flush to
    deletion.  A calibration run pushes the whole stream through the plain
    path (one job at a time, every flush available at once); its time per
    flush is the plain path's sustainable rate.  Then, for each rate multiple,
    flushes arrive at that multiple of the sustainable rate and the same
    stream runs through the plain path and through the served form
    (luda_submit_build / luda_submit on the worker slots).  Reported per run:
    aggregate input GB/s (flush inputs are memtable bytes, compaction inputs
    table bytes) and p50/p99 job latency (a flush from its arrival, a
    compaction from the moment the tree first needed it)."""
    from luda_amd import served
    t = time.time()
    mems = list(served.write_stream(args.serve, args.mix_seed))
    log(f"{args.serve} puts in {len(mems)} memtables generated in {time.time() - t:.1f}s")
    opts = ffi.default_options()
    peng = jobs.Engine(0)
    seng = jobs.Engine(0)
    seng.set_slots(args.slots)
    # the plain path is the drop-in as it is: every job decodes its inputs;
    # the served form keeps decoded entries resident between chained jobs
    pbe = served.EngineBackend(peng, opts, resident=False, io_sync=args.io_sync)
    sbe = served.EngineBackend(seng, opts, resident=True, io_sync=args.io_sync)
    io_root = tempfile.mkdtemp(prefix="luda_serve_") if args.io else None

    def run_on(be, run, interval, tag, **kw):
        """One run of the stream; with --io into a fresh directory, whose
        files must then be exactly the final tree's."""
        lsm = served.Lsm()
        if io_root:
            be.io_dir = os.path.join(io_root, tag)
            os.makedirs(be.io_dir)
            be.io = dict.fromkeys(be.io, 0)
            lsm.on_delete = be.delete_file
        be.acct = dict.fromkeys(be.acct, 0)
        lsm, st = run(be, mems, interval, lsm=lsm, **kw)
        summ = st.summary()
        summ["engine_calls"] = {k: round(v, 3) for k, v in be.acct.items()}
        if io_root:
            tree = {"%06d.ldb" % f.number: f for level in lsm.levels for f in level}
            on_disk = set(os.listdir(be.io_dir))
            same = on_disk == set(tree) and all(os.path.getsize(os.path.join(be.io_dir, n)) == tree[n].size
                                                for n in tree)
            summ["io"] = dict(be.io, write_ms=round(be.io["write_ms"], 3), tree_files_on_disk=same)
            if not same:
                raise RuntimeError(f"{tag}: the directory does not hold the tree's files")
        return lsm, st, summ

    def end_run(be, lsm):
        release_tree(lsm)
        if io_root:
            shutil.rmtree(be.io_dir)
            be.io_dir = None

    # warm-up: the whole stream through both paths, so that every working
    # buffer and output buffer is allocated before anything is timed
    for be, run in ((pbe, served.run_plain), (sbe, served.run_served)):
        lsm, wst, ws = run_on(be, run, 0.0, "warm_" + run.__name__)
        end_run(be, lsm)
        log(f"warm-up {run.__name__}: {ws['jobs']} jobs in {ws['seconds']} s")
    lsm, st, cal = run_on(pbe, served.run_plain, 0.0, "calibration")
    end_run(pbe, lsm)
    per_flush_ms = (st.end_ms - st.t0_ms) / len(mems)
    log(f"plain path, whole stream at once: {cal}")
    out = {"metric": "served compaction: aggregate input GB/s and p50/p99 job latency, served form vs the plain path",
           "puts": args.serve, "memtables": len(mems), "slots": args.slots,
           "workload": "fillrandom: 16-byte keys uniform over 1e9, 100-byte random values, 4 MiB memtables "
                       "(MemTable::Add bytes), LevelDB picking/scoring rules, tables resident in HBM"
                       + ("; every table also written to a file" + (" and synced" if args.io_sync else "")
                          if args.io else ""),
           "calibration": cal, "sustainable_flush_interval_ms": round(per_flush_ms, 4), "rates": []}
    checked = [0, 0]
    for k in [float(x) for x in args.rates.split(",")]:
        interval = per_flush_ms / k
        row = {"rate": k, "flush_interval_ms": round(interval, 4)}
        lsm, st, row["plain"] = run_on(pbe, served.run_plain, interval, f"plain_{k}")
        end_run(pbe, lsm)
        samples = []

        def on_job(spec, files):
            # every 8th compaction is kept (inputs and outputs held) and
            # checked after the timed run
            if len(samples) < 16 and len(samples_seen) % 8 == 0:
                for f in spec.inputs[0] + spec.inputs[1] + files:
                    f.owner.refs += 1
                samples.append((spec, files))
            samples_seen.append(1)
        samples_seen = []
        lsm, st, row["served"] = run_on(sbe, served.run_served, interval, f"served_{k}", on_job=on_job)
        lsm.check()
        end_run(sbe, lsm)
        scratch = served.Lsm()
        for spec, kept in samples:
            want = [sha(sbe.table_bytes(f)) for f in kept]
            files, _ = pbe.compact(scratch, spec)
            checked[0] += 1
            checked[1] += [sha(pbe.table_bytes(f)) for f in files] == want
            if files:
                files[0].owner.refs = 1
                files[0].owner.drop()
            for f in spec.inputs[0] + spec.inputs[1] + kept:
                f.owner.drop()
        row["served_vs_plain_gbps"] = round(row["served"]["gbps"] / max(1e-9, row["plain"]["gbps"]), 3)
        log(f"rate {k}x: {row}")
        out["rates"].append(row)
    out["jobs_checked_alone"] = checked[0]
    out["same_alone"] = checked[0] == checked[1]
    print(json.dumps(out), flush=True)
    sbe.close()
    pbe.close()
    if io_root:
        shutil.rmtree(io_root, ignore_errors=True)
    seng.close()
    peng.close()


def mix_golden(n, seed):
    """Per-job output digests the reference's own code produced for this mix
    (tests/golden/mix_jobs.json), or None when it holds no record for it."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "mix_jobs.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get("jobs") != n or rec.get("seed") != seed:
        return None
    return [j["digest"] for j in rec["records"]]


def plain_hash_free(eng, results, summary):
    for res in results:
        yield res
        eng.free(res)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1, choices=sorted(WORKLOADS))
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mix", type=int, default=0, metavar="JOBS",
                    help="instead: JOBS jobs of 4-256 MiB through the batched form (luda_submit) against the "
                         "plain path, one JSON line of aggregate GB/s and job latency")
    ap.add_argument("--mix-seed", type=int, default=3)
    ap.add_argument("--slots", type=int, default=6)
    ap.add_argument("--tune", type=int, default=0, metavar="JOBS",
                    help="with --mix: first autotune the batched form (slots, launch shapes per job class) on a "
                         "mix of JOBS jobs with the next seed, then measure with its choice")
    ap.add_argument("--serve", type=int, default=0, metavar="PUTS",
                    help="instead: a fillrandom stream of PUTS puts through the served form against the plain "
                         "path at --rates multiples of the plain path's sustainable rate, one JSON line")
    ap.add_argument("--rates", default="1,2,4")
    ap.add_argument("--io", action="store_true",
                    help="with --serve: every table is also written to a file (a fresh directory under $TMPDIR per "
                         "run), the served form on a writer thread beside its kernels, the plain path before its "
                         "next job; compactions delete their inputs' files")
    ap.add_argument("--io-sync", action="store_true", help="with --io: fsync every file")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (the known-size kernel that rocprofv3's "
                         "FETCH_SIZE/WRITE_SIZE are scaled against) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.mix:
        run_mix(args)
        return
    if args.serve:
        run_serve(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: a compaction job runs whole on one device and this build has no "
              "multi-GPU path; run it without --gpus or with --gpus 1", file=sys.stderr)
        sys.exit(2)

    t = time.time()
    wl = jobs.Workload(args.config, scale=args.scale)
    opts = ffi.default_options()
    log(f"workload {args.config} scale {args.scale}: {wl.w.total_entries} entries generated in {time.time() - t:.1f}s")
    eng = jobs.Engine(0)
    t = time.time()
    dev, keep = build_device_inputs(eng, wl, opts)
    bytes_in = sum(s for which in dev for _, s in which)
    log(f"inputs built on the GPU: {len(dev[0])}+{len(dev[1])} tables, {bytes_in / 2**20:.1f} MiB "
        f"in {time.time() - t:.1f}s")
    job = jobs.make_job(wl, opts, None, flags=ffi.LUDA_JOB_KEEP_ON_DEVICE, device_inputs=dev)

    def run_once():
        rc, out, msg = eng.compact(job.job, keep_on_device=True)
        if rc != ffi.LUDA_OK:
            raise RuntimeError(f"luda_compact failed ({rc}): {msg}")
        return out

    os.environ.pop("LUDA_KERNEL_TIMING", None)
    for i in range(args.warmup):
        out = run_once()
        eng.free(out[2])
    # which kernel dominates: one untimed pass with every kernel timed
    os.environ["LUDA_KERNEL_TIMING"] = "1"
    out = run_once()
    top = out[1]["top_kernel"]
    eng.free(out[2])
    log(f"dominant kernel: {top}")
    # timed steps, with HIP events around that kernel only
    os.environ["LUDA_KERNEL_TIMING"] = top
    lat, dev_ms, kms, kbytes, klaunch = [], [], [], [], []
    st = None
    t0 = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        out = run_once()
        lat.append((time.perf_counter() - s0) * 1e3)
        st = out[1]
        dev_ms.append(st["device_ms"])
        kms.append(st["top_kernel_ms"])
        kbytes.append(st["top_kernel_bytes"])
        klaunch.append(st["top_kernel_launches"])
        if i < args.steps - 1:
            eng.free(out[2])
    t1 = time.perf_counter()
    os.environ.pop("LUDA_KERNEL_TIMING", None)
    last = out[2]
    elapsed = t1 - t0
    value = bytes_in * args.steps / elapsed / 1e9
    out_bytes = st["bytes_written"]
    dev_mean = statistics.mean(dev_ms)
    per_launch_ms = sum(kms) / max(1, sum(klaunch))
    per_launch_bytes = kbytes[-1] / max(1, klaunch[-1])
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else None
    log(f"{args.steps} steps in {elapsed:.3f}s: {value:.2f} GB/s input, device {dev_mean:.2f} ms/job, "
        f"{top} {per_launch_ms:.3f} ms/launch")

    # Output tables of the last timed step: verified in HBM (luda_verify_tables,
    # every block's CRC), then read back for the parity check.
    r = last.contents
    outs = [(r.device_data + r.outputs[i].offset, r.outputs[i].file_size) for i in range(r.num_outputs)]
    vt = []
    for _ in range(3):
        v0 = time.perf_counter()
        vrc, vbad, vmsg = eng.verify(outs, on_device=True)
        vt.append(time.perf_counter() - v0)
        if vrc != ffi.LUDA_OK:
            raise RuntimeError(f"luda_verify_tables rejected output table {vbad}: {vmsg}")
    out_bytes = sum(sz for _, sz in outs)
    verify = {"tables": len(outs), "ms": round(min(vt) * 1e3, 3), "gbps": round(out_bytes / min(vt) / 1e9, 3),
              "passed": True}
    log(f"outputs verified on the GPU: {len(outs)} tables in {verify['ms']} ms ({verify['gbps']} GB/s)")
    gpu_tables = [eng.d2h(p, sz) for p, sz in outs]
    eng.free(last)

    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle  # CPU baseline leg only
        kind = "reference" if oracle.ref_available() else "port"
        host_inputs = ([eng.d2h(p, s) for p, s in dev[0]], [eng.d2h(p, s) for p, s in dev[1]])
        hjob = jobs.make_job(wl, opts, host_inputs)
        times = []
        cpu_tables = None
        for i in range(max(1, args.cpu_runs)):
            c0 = time.perf_counter()
            rc, tabs = oracle.compact(hjob.job, "ref" if kind == "reference" else "oracle")
            times.append(time.perf_counter() - c0)
            if rc != 0:
                raise RuntimeError(f"CPU compaction failed ({rc})")
            cpu_tables = tabs
        tmed = statistics.median(times)
        cpu = {"value": round(bytes_in / tmed / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": kind,
               "sample": f"the whole job ({bytes_in / 2**20:.0f} MiB of input tables, median of {len(times)} runs, "
                         f"one thread, inputs and outputs in RAM)",
               "ms_per_job": round(tmed * 1e3, 1), "host": host_info()}
        parity = [sha(t.data) for t in cpu_tables] == [sha(t) for t in gpu_tables]
        log(f"CPU ({kind}): {cpu['value']} GB/s; outputs identical: {parity}")

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.config], "scale": args.scale, "input_bytes": bytes_in,
                   "output_bytes": out_bytes, "entries": st["entries_in"], "kept": st["entries_kept"],
                   "input_tables": [len(dev[0]), len(dev[1])], "output_tables": len(gpu_tables),
                   "block_size": 4096, "restart_interval": 16, "bloom_bits_per_key": 10,
                   "max_file_size": 2 << 20},
        "device_ms_per_job": round(dev_mean, 3),
        "device_input_gbps": round(bytes_in / (dev_mean * 1e-3) / 1e9, 3),
        "output_gbps": round(out_bytes * args.steps / elapsed / 1e9, 3),
        "job_hbm_roofline_frac": round((bytes_in + out_bytes) / (dev_mean * 1e-3) / 1e9 / HBM_PEAK_GBPS, 5),
        "p50_ms": round(percentile(lat, 50), 3),
        "p99_ms": round(percentile(lat, 99), 3),
        "roofline": {"bound": "hbm", "kernel": top,
                     "achieved": round(achieved, 2) if achieved else None, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 5) if achieved else None,
                     "bytes_per_launch": per_launch_bytes, "ms_per_launch": round(per_launch_ms, 4),
                     "traffic": measured_traffic(top, args.config, args.scale)},
        "cpu_baseline": cpu,
        "verify_outputs": verify,
        "parity": parity,
    }
    print(json.dumps(line), flush=True)
    for res in keep:
        eng.free(res)
    eng.close()


if __name__ == "__main__":
    main()
