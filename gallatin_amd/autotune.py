"""Autotuner for batched request streams.

Sweeps the allocator geometry, the workgroup shape and the form of the
batched path over one stream and reports pairs/s for each point, beside the
plain per-thread path at the same geometry:

    segment bytes  {4, 16, 64} MiB
    smallest slice {16, 128} B, the stated grid, and 256 B beside it (a
                   coarser first class rounds small requests up further:
                   internal fragmentation for fewer classes)
    workgroup      {256, 512, 1024} threads
    batched form   workgroup batches (batched.hpp) or per-workgroup LDS
                   slice caches (slice_cache.hpp)

The biggest slice is min(segment / 4096, 4096) bytes: a segment must hold
one block of the biggest class.  Each point replays the stream as one
captured hipGraph (all batches plus the final cleanup) and checks the
counters: every malloc issued, every successful malloc freed exactly once,
every tag intact.
"""
import itertools

import numpy as np

from . import (CTR_BAD, CTR_COUNT, CTR_FREE_TICKS, CTR_FREES_SKIPPED, CTR_MALLOC_TICKS,
               CTR_MALLOCS, CTR_MISSES, CTR_PAIRS, CTR_TIMEOUTS, CTR_WAIT_POLLS, CTR_WAIT_TICKS,
               CTR_WAVE_TICKS, BatchPlan, run_batches)
from . import Gallatin, GallatinError

SEGMENTS = (4 << 20, 16 << 20, 64 << 20)
FORMS = {"workgroup": 1, "cache": 2}
SMALLEST = (16, 128, 256)
GRID_SMALLEST = (16, 128)  # the stated sweep; other smallest sizes are reported beside it
THREADS = (256, 512, 1024)


def biggest_for(segment_bytes):
    return min(segment_bytes // 4096, 4096)


def measure(alloc, stream, batched, threads, reps=3, warmup=1, split_reps=3):
    """Time `reps` graph replays of `stream`; returns a result dict.  Its
    `events` are the allocator's event counters (gal_get_stats) that moved
    during the timed replays, summed over them; `wait_polls` counts the
    polls of frees whose entry was not yet published, per replay.
    `wave_time` splits the stream kernel's wave time into malloc rounds,
    waits of frees for their entries, the frees themselves and the rest
    (queue, decode), as shares, from `split_reps` more replays of a timed
    plan (the clock reads slow a replay, so the rate comes from the
    untimed one); 0 skips them."""
    import torch
    ctr = torch.zeros(CTR_COUNT, dtype=torch.int64, device="cuda")
    plan = BatchPlan(alloc, stream, ctr, threads=threads, batched=batched)
    try:
        for _ in range(warmup):
            plan.launch()
        torch.cuda.synchronize()
        ctr.zero_()
        torch.cuda.synchronize()
        s0 = alloc.stats()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        evs[0].record()
        for k in range(reps):
            plan.launch()
            evs[k + 1].record()
        torch.cuda.synchronize()
        s1 = alloc.stats()
    finally:
        plan.destroy()
    c = ctr.cpu().tolist()
    split = None
    if split_reps:
        tctr = torch.zeros(CTR_COUNT, dtype=torch.int64, device="cuda")
        tplan = BatchPlan(alloc, stream, tctr, threads=threads, batched=batched, timed=True)
        try:
            for _ in range(split_reps):
                tplan.launch()
            torch.cuda.synchronize()
        finally:
            tplan.destroy()
        split = _wave_time(tctr.cpu().tolist())
    total_ms = evs[0].elapsed_time(evs[-1])
    steps = sorted(evs[k].elapsed_time(evs[k + 1]) for k in range(reps))
    ok = (c[CTR_MALLOCS] == reps * stream.n_mallocs and c[CTR_BAD] == 0
          and c[CTR_TIMEOUTS] == 0
          and c[CTR_PAIRS] == c[CTR_MALLOCS] - c[CTR_MISSES]
          and (c[CTR_MISSES] > 0 or c[CTR_FREES_SKIPPED] == 0))
    return {
        "pairs_per_s": c[CTR_PAIRS] / (total_ms / 1e3),
        "ms_per_step": total_ms / reps,
        "p50_step_ms": steps[len(steps) // 2],
        "p99_step_ms": steps[-1],
        "misses": c[CTR_MISSES],          # over all reps
        "violations": c[CTR_BAD] + c[CTR_TIMEOUTS] + (0 if ok else 1),
        "pairs_per_step": c[CTR_PAIRS] / reps,
        "events": {k: s1[k] - s0[k] for k in s1 if s1[k] != s0[k]},
        "wave_time": split,
        "wait_polls": c[CTR_WAIT_POLLS] / reps,
    }


def measure_launches(alloc, stream, batched, threads, reps=3, warmup=1):
    """Time `reps` runs of `stream` with one launch per batch and one for
    the cleanup (gal_run_batches), the way a caller that launches a kernel
    per batch drives the allocator.  Same result keys as measure(), without
    the event counters and the wave split."""
    import torch
    ctr = torch.zeros(CTR_COUNT, dtype=torch.int64, device="cuda")
    for _ in range(warmup):
        run_batches(alloc, stream, ctr, threads=threads, batched=batched)
    torch.cuda.synchronize()
    ctr.zero_()
    torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    evs[0].record()
    for k in range(reps):
        run_batches(alloc, stream, ctr, threads=threads, batched=batched)
        evs[k + 1].record()
    torch.cuda.synchronize()
    c = ctr.cpu().tolist()
    total_ms = evs[0].elapsed_time(evs[-1])
    steps = sorted(evs[k].elapsed_time(evs[k + 1]) for k in range(reps))
    ok = (c[CTR_MALLOCS] == reps * stream.n_mallocs and c[CTR_BAD] == 0
          and c[CTR_TIMEOUTS] == 0
          and c[CTR_PAIRS] == c[CTR_MALLOCS] - c[CTR_MISSES]
          and (c[CTR_MISSES] > 0 or c[CTR_FREES_SKIPPED] == 0))
    return {
        "pairs_per_s": c[CTR_PAIRS] / (total_ms / 1e3),
        "ms_per_step": total_ms / reps,
        "p50_step_ms": steps[len(steps) // 2],
        "p99_step_ms": steps[-1],
        "misses": c[CTR_MISSES],
        "violations": c[CTR_BAD] + c[CTR_TIMEOUTS] + (0 if ok else 1),
    }


def _wave_time(c):
    life = c[CTR_WAVE_TICKS]
    if not life:
        return None
    parts = {"malloc": c[CTR_MALLOC_TICKS], "wait": c[CTR_WAIT_TICKS], "free": c[CTR_FREE_TICKS]}
    out = {k: v / life for k, v in parts.items()}
    out["other"] = 1.0 - sum(out.values())
    return out


def tune(stream, pool_bytes, seed=42, reps=3, warmup=1, segments=SEGMENTS,
         smallest=SMALLEST, threads=THREADS, forms=tuple(FORMS), slots_min=32, flags=0,
         log=None):
    """Sweep every (segment, smallest, workgroup, form) point.  Returns
    (best, table): best is the fastest batched point with no violations
    inside the stated grid (smallest slice in GRID_SMALLEST), or outside it
    when the sweep has no such point."""
    table = []
    for seg, mn in itertools.product(segments, smallest):
        big = biggest_for(seg)
        if mn > big:
            continue
        alloc = Gallatin(seg, mn, big).generate_on_device(
            pool_bytes, seed, print_info=False, slots_min=slots_min, flags=flags)
        try:
            for form, wg in itertools.product(forms, threads):
                r = measure(alloc, stream, FORMS[form], wg, reps, warmup)
                r.update(path="batched", form=form, segment_bytes=seg, smallest=mn,
                         biggest=big, threads=wg, in_grid=mn in GRID_SMALLEST)
                table.append(r)
                if log:
                    log(r)
            r = measure(alloc, stream, False, 256, reps, warmup)
            r.update(path="plain", form="per-thread", segment_bytes=seg, smallest=mn,
                     biggest=big, threads=256, in_grid=mn in GRID_SMALLEST)
            table.append(r)
            if log:
                log(r)
            if alloc.info()["errors"]:
                raise GallatinError(f"allocator errors at segment {seg}, smallest {mn}")
        finally:
            alloc.free_on_device()
    good = [r for r in table if r["path"] == "batched" and r["violations"] == 0]
    in_grid = [r for r in good if r["in_grid"]] or good
    best = max(in_grid, key=lambda r: r["pairs_per_s"]) if in_grid else None
    return best, table


def main(argv=None):
    """python -m gallatin_amd.autotune [--seed S] [--pool-gib G] [--reps R]
    [--segments-mib M ...] [--smallest B ...] [--threads T ...] [--forms F ...]
    [--slots-min K]:
    generate the Zipf stream, sweep the given points (the full grid by
    default), print one JSON line per point and the best point last."""
    import json

    import torch

    from . import BatchStream
    from .batch_streams import zipf_stream

    a = parse_args(argv)
    torch.cuda.set_device(0)
    ops, off, nslots = zipf_stream(a.seed, nbatches=a.nbatches)
    stream = BatchStream(ops, off, nslots)
    best, _ = tune(stream, a.pool_gib << 30, seed=a.seed, reps=a.reps, flags=a.flags,
                   segments=tuple(m << 20 for m in a.segments_mib), smallest=tuple(a.smallest),
                   threads=tuple(a.threads), forms=tuple(a.forms), slots_min=a.slots_min,
                   log=lambda r: print(json.dumps(r), flush=True))
    print(json.dumps({"best": best}), flush=True)


def parse_args(argv=None):
    """The command line of main(); the defaults are the full grid."""
    import argparse

    p = argparse.ArgumentParser()
    p.add_argument("--seed", type=int, default=1234)
    p.add_argument("--pool-gib", type=int, default=64)
    p.add_argument("--reps", type=int, default=3)
    p.add_argument("--nbatches", type=int, default=256)
    p.add_argument("--flags", type=int, default=0)
    p.add_argument("--segments-mib", type=int, nargs="+", default=[s >> 20 for s in SEGMENTS])
    p.add_argument("--smallest", type=int, nargs="+", default=list(SMALLEST))
    p.add_argument("--threads", type=int, nargs="+", default=list(THREADS))
    p.add_argument("--forms", nargs="+", choices=tuple(FORMS), default=list(FORMS))
    p.add_argument("--slots-min", type=int, default=32,
                   help="pinned blocks per class at least (gal_config.slots_min)")
    return p.parse_args(argv)


if __name__ == "__main__":
    main()
