"""Autotuner for the batched form: picks the worker-slot count and the launch
shapes of the two kernels that hold the most of a CU (the block decoder and
the data block writer), per class of job, by measuring candidates on sample
jobs through the C-ABI (luda_set_slots, luda_set_tuning, luda_submit).

Why these knobs: a large job's block writer takes every CU's LDS (one
16-wave workgroup per CU) for its whole run, and its decoder thousands of
wavefronts, so while they run the other jobs of a batch wait for CUs rather
than overlap.  Fewer workgroups for large jobs leave CUs to the small jobs'
latency-bound kernels.  Which split pays depends on the mix, so it is
measured, not assumed.

Rules are keyed on input bytes and on key width (luda_tuning_rule); the
result is a JSON-able dict that bench.py applies with Engine.set_tuning.
"""
import itertools
import time

from . import ffi, jobs

# job classes the rules distinguish: input bytes
SIZE_CLASSES = [(0, 32 << 20), (32 << 20, 128 << 20), (128 << 20, ffi.U64_MAX)]


def rules_for(shape):
    """shape: {size class index: (write_blocks_groups, decode_groups)} ->
    luda_tuning_rule dicts (any key width)."""
    out = []
    for ci, (wb, dg) in shape.items():
        lo, hi = SIZE_CLASSES[ci]
        out.append({"min_bytes": lo, "max_bytes": hi, "write_blocks_groups": wb, "decode_groups": dg})
    return out


def job_bytes(job):
    return sum(job.inputs[w][i].size for w in (0, 1) for i in range(job.num_inputs[w]))


def measure_batched(jobs_list, total_bytes, slots, rules, repeats=2):
    """Aggregate input GB/s of the jobs all queued at once on a fresh context
    with `slots` worker slots and the given rules (best of `repeats`)."""
    eng = jobs.Engine(0)
    try:
        eng.set_slots(slots)
        eng.set_tuning(rules)
        # warm-up: the largest jobs, one per slot, so that every slot's
        # working memory has its final size
        order = sorted(range(len(jobs_list)), key=lambda i: -job_bytes(jobs_list[i]))
        failed = []

        def warm(rc, o, msg):
            if rc != ffi.LUDA_OK:
                failed.append((rc, msg))
            else:
                eng.free(o[2])

        for i in order[:slots]:
            eng.submit(jobs_list[i], warm, keep_on_device=True)
        eng.drain()
        if failed:
            raise RuntimeError(f"warm-up job failed ({failed[0][0]}): {failed[0][1]}")
        best = 0.0
        for _ in range(repeats):
            res = []
            t0 = time.perf_counter()
            for j in jobs_list:
                eng.submit(j, lambda rc, o, msg: res.append((rc, o, msg)), keep_on_device=True)
            eng.drain()
            dt = time.perf_counter() - t0
            bad = [(rc, msg) for rc, o, msg in res if rc != ffi.LUDA_OK]
            for rc, o, msg in res:
                if rc == ffi.LUDA_OK:
                    eng.free(o[2])
            if bad:
                raise RuntimeError(f"batched job failed ({bad[0][0]}): {bad[0][1]}")
            best = max(best, total_bytes / dt / 1e9)
        return best
    finally:
        eng.close()


def autotune_batched(jobs_list, total_bytes, slots_grid=(4, 6, 8), wb_grid=(0, 128, 64), dg_grid=(0, 1024),
                     log=print):
    """Searches slots x (writer and decoder shapes for the largest class),
    then the middle class's writer shape with the best of those fixed.
    Returns {"slots", "rules", "gbps", "tried": [...]}."""
    tried = []
    best = None
    for slots, wb, dg in itertools.product(slots_grid, wb_grid, dg_grid):
        rules = rules_for({2: (wb, dg)})
        g = measure_batched(jobs_list, total_bytes, slots, rules)
        tried.append({"slots": slots, "large": [wb, dg], "gbps": round(g, 3)})
        log(f"autotune: slots {slots} large-job writer {wb or 256} decoder {dg or 'default'}: {g:.1f} GB/s")
        if best is None or g > best[0]:
            best = (g, slots, {2: (wb, dg)})
    g0, slots, shape = best
    for wb in wb_grid[1:]:
        cand = dict(shape)
        cand[1] = (wb, 0)
        g = measure_batched(jobs_list, total_bytes, slots, rules_for(cand))
        tried.append({"slots": slots, "large": list(shape[2]), "middle": [wb, 0], "gbps": round(g, 3)})
        log(f"autotune: + middle-job writer {wb}: {g:.1f} GB/s")
        if g > best[0]:
            best = (g, slots, cand)
    g, slots, shape = best
    return {"slots": slots, "rules": rules_for(shape), "gbps": round(g, 3), "tried": tried}
