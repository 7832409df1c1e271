#!/usr/bin/env python3
"""Benchmark: device malloc+free pairs/s on the mixed-size churn, one MI355X.

Workload (one step = one pass of the hot path over one batch):
  mixed-size churn, 8 GiB pool, Gallatin<16 MiB, 16, 4096>, 2^24 threads x 10
  rounds; sizes log-uniform in [1, 4096) B; each round a thread replaces its
  live allocation with probability 1/2; every allocation is filled with a
  uint64 (thread, round) pattern and verified before it is freed.

Both paths run the same request streams: the batched path (workgroup-batched
malloc/free in one persistent launch, the headline `value`) and the plain
per-thread path (`plain` beside it).  Pairs are counted on the device.  The
roofline is the headline kernel's: payload bytes written and read against HBM
peak, and its atomic issue rate against the measured atomic ceiling.
`cpu_baseline` is the oracle's sequential replay of the same request streams
on one host core; multithreaded glibc malloc/free of the same streams is
reported beside it.

Beside the headline: the plain small method (1 GiB, malloc(16)), the block
and multi-segment batches (64 GiB), the batched Zipf stream with its
autotuner sweep and the serving loop, each with its own workload string.

  python bench.py [--steps K] [--warmup W] [--path batched|plain|both]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("malloc+free pairs/sec (mixed 16B–4KiB churn, 1 GPU) batched vs plain; "
          "p99 step")
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--path", choices=("batched", "plain", "both"), default="both")
    p.add_argument("--threads", type=int, default=1 << 24)
    p.add_argument("--rounds", type=int, default=10)
    p.add_argument("--pool-gib", type=int, default=8)
    p.add_argument("--seed", type=int, default=1234)
    p.add_argument("--slots0", type=int, default=128)
    p.add_argument("--slots-min", type=int, default=32)
    p.add_argument("--partitions", type=int, default=0)
    p.add_argument("--flags", type=int, default=0)
    p.add_argument("--block-threads", type=int, default=256)
    p.add_argument("--workgroups", type=int, default=0)
    p.add_argument("--batch-threads", type=int, default=1024)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-ceilings", action="store_true")
    p.add_argument("--cpu-sample-threads", type=int, default=1 << 22)
    p.add_argument("--no-zipf", action="store_true",
                   help="skip the batched Zipf-stream sweep")
    p.add_argument("--zipf-pool-gib", type=int, default=64)
    p.add_argument("--zipf-reps", type=int, default=3)
    p.add_argument("--no-big", action="store_true",
                   help="skip the block and multi-segment batches")
    p.add_argument("--big-steps", type=int, default=20)
    p.add_argument("--no-serving", action="store_true", help="skip the serving loop")
    p.add_argument("--no-word-churn", action="store_true",
                   help="skip the churn with one word per allocation")
    p.add_argument("--serving-pool-gib", type=int, default=200)
    # the timed window covers a whole admission cycle: the pool fills, the
    # gate closes, long-lived sessions drain, the gate reopens
    p.add_argument("--serving-steps", type=int, default=1000)
    p.add_argument("--serving-warmup", type=int, default=300)
    p.add_argument("--serving-reserve", type=float, default=0.10,
                   help="admission reserve: fraction of the pool's segments kept free")
    p.add_argument("--cpu-serving-steps", type=int, default=30)
    p.add_argument("--serving-gate", type=int, default=0, choices=(0, 1),
                   help="admission gate: 0 free segments of the segment tree, 1 free memory "
                        "in segments (free segments plus free blocks of formatted segments)")
    p.add_argument("--detail-json", default="",
                   help="also write the full result (every section, the autotune table) here")
    return p.parse_args()


def host_cores():
    """Host cores for the CPU baselines: every online core this process may
    run on (its affinity mask), limited by the cgroup CPU quota when one is
    set (the box gives a process a share of the machine, and more threads
    than the quota only time-slice).  Returns (used, online)."""
    online = os.sysconf("SC_NPROCESSORS_ONLN")
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = online
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(quota) // int(period)))
    except (OSError, ValueError):
        pass
    return max(1, n), online


def host_available_bytes():
    """MemAvailable of the host, or None."""
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable:"):
                    return int(line.split()[1]) * 1024
    except (OSError, ValueError):
        pass
    return None


def measured_traffic(kernel, args):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (profiles/hbm_traffic.json), when it was taken with these settings."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    mine = {"threads": args.threads, "rounds": args.rounds, "pool_gib": args.pool_gib,
            "slots_min": args.slots_min, "batch_threads": args.batch_threads}
    if any(mine.get(k) != v for k, v in d.get("match", {}).items()):
        return None
    k = d.get("kernels", {}).get(kernel)
    return k["bytes"] if k else None


def atomic_ceiling(G, torch):
    """Agent-scope atomicAdd(u32) with returned value: distinct consecutive
    addresses, one address per 128-byte line, one address per wave, one
    address for the whole device."""
    words = torch.zeros(1 << 26, dtype=torch.int32, device="cuda")
    out = {}
    for name, span, lanes, stride, nops in (
            ("distinct", 1 << 26, 1, 1, 1 << 28),
            ("per_line", 1 << 26, 1, 32, 1 << 27),
            ("per_wave", 1 << 26, 64, 32, 1 << 26),
            ("one_address", 1, 1, 1, 1 << 22)):
        G.bench_atomic(words, span, nops, lanes, stride)  # warm
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        G.bench_atomic(words, span, nops, lanes, stride)
        e1.record()
        torch.cuda.synchronize()
        out[name] = nops / (e0.elapsed_time(e1) / 1e3)
    del words
    torch.cuda.empty_cache()
    return out


def hbm_ceiling(G, torch):
    """Read+write rates (GB/s, both directions counted) of 2 GiB copies:
    contiguous 64 KiB tiles per workgroup, and scattered pieces of 256 B and
    2 KiB (pieces read and written at permuted positions, each wave walking
    consecutive 16-byte vectors as the churn's fill and verify do). The
    churn's payload bytes are spread evenly over allocation sizes 1..4096 B,
    so the 2 KiB scattered copy is the rate its access pattern can reach."""
    src = torch.empty(1 << 31, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    out = {}

    def rate(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return 5 * 2 * src.numel() / (e0.elapsed_time(e1) / 1e3) / 1e9

    out["copy_gbs"] = rate(lambda: G.bench_copy(src, dst, src.numel()))
    for piece in (256, 2048):
        out[f"scatter_{piece}_gbs"] = rate(lambda: G.bench_scatter(src, dst, src.numel(), piece))
    del src, dst
    torch.cuda.empty_cache()
    return out


def tree_probe(args, G, torch):
    """Wave-cooperative successor searches on a fresh 64 GiB instance's
    trees (3 levels, 4096 segments): alternately the segment tree (dense)
    and class tree 0 (sparse: its boot segments only), from pseudo-random
    starts, read only. 4096 waves x 2000 searches."""
    a = G.Gallatin(16 << 20, 16, 4096).generate_on_device(64 << 30, args.seed, print_info=False)
    waves, searches = 4096, 2000
    G.bench_tree(a, 0, 100, waves)  # warm
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    G.bench_tree(a, 0, searches, waves)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    a.free_on_device()
    return {"workload": "64 GiB instance, 4096 waves x 2000 find_ge, segment tree and "
                        "class tree 0 alternately",
            "searches_per_s": waves * searches / (ms / 1e3),
            "us_per_search_per_wave": ms * 1e3 / searches}


def plain_small(args, G, torch, flags=0, partitions=0):
    """Plain method, small: 1 GiB pool, Gallatin<16 MiB, 16, 4096>, one
    launch of 2^20 threads each doing malloc(16), write its uint32 id, read
    it back, free; 10 launches per step. With GAL_FLAG_NO_PAYLOAD_FENCE the
    id is written and read with device-coherent accesses and the allocator
    skips its payload fences; with GAL_FLAG_LAZY_BOOT and one partition per
    XCD a home free releases only to its own XCD's L2 and a malloc needs no
    acquire."""
    a = G.Gallatin(16 << 20, 16, 4096).generate_on_device(1 << 30, args.seed, print_info=False,
                                                          flags=flags, partitions=partitions)
    ctr = torch.zeros(G.CTR_COUNT, dtype=torch.int64, device="cuda")
    n, launches, steps = 1 << 20, 10, 5
    for _ in range(launches):  # warm-up step
        a.run_plain(n, 16, counters=ctr)
    torch.cuda.synchronize()
    ctr.zero_()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps * launches):
        a.run_plain(n, 16, counters=ctr)
    e1.record()
    torch.cuda.synchronize()
    c = ctr.cpu().tolist()
    info = a.info()
    a_parts = a.partitions
    a.free_on_device()
    secs = e0.elapsed_time(e1) / 1e3
    return {"workload": "1 GiB pool, 2^20 threads x malloc(16)/write id/read/free, "
                        "10 launches per step", "xcd_partitions": a_parts,
            "value": c[G.CTR_PAIRS] / secs, "unit": "pairs/s",
            "ms_per_step": secs * 1e3 / steps, "misses": c[G.CTR_MISSES],
            "violations": c[G.CTR_BAD] + info["errors"]}


def block_multiseg(args, G, torch):
    """Block and multi-segment path: 64 GiB pool (4096 segments, three-level
    trees).  One step is one batch of 4096 requests uniform in [64 KiB,
    64 MiB] and 64 uniform in [256 MiB, 2 GiB]: a launch that mallocs and
    tags each one, then a launch that checks and frees them in a
    pseudo-random order.  Demand is far above the pool, so many requests
    miss.  After the timed steps a drain must leave every segment free."""
    a = G.Gallatin(16 << 20, 16, 4096).generate_on_device(64 << 30, args.seed,
                                                          print_info=False)
    ctr = torch.zeros(G.CTR_COUNT, dtype=torch.int64, device="cuda")
    ptrs = torch.zeros(4096 + 64, dtype=torch.int64, device="cuda")
    warm, steps = 3, args.big_steps
    for s in range(warm):
        a.run_big(args.seed, s, ptrs=ptrs, counters=ctr)
    torch.cuda.synchronize()
    ctr.zero_()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    ev[0].record()
    for i, s in enumerate(range(warm, warm + steps)):
        a.run_big(args.seed, s, ptrs=ptrs, counters=ctr)
        ev[i + 1].record()
    torch.cuda.synchronize()
    c = ctr.cpu().tolist()
    per = sorted(ev[i].elapsed_time(ev[i + 1]) for i in range(steps))
    secs = ev[0].elapsed_time(ev[-1]) / 1e3
    a.drain()
    info = a.info()
    a.free_on_device()
    return {"workload": "64 GiB pool, per step 4096 requests uniform in [64 KiB, 64 MiB] + 64 "
                        "in [256 MiB, 2 GiB]: malloc and tag, then check and free in a "
                        "pseudo-random order (two launches)",
            "value": c[G.CTR_PAIRS] / secs, "unit": "pairs/s", "steps": steps,
            "ms_per_step": secs * 1e3 / steps, "p50_step_ms": per[steps // 2],
            "p99_step_ms": per[min(steps - 1, (99 * steps) // 100)],
            "requests_per_step": 4160, "pairs_per_step": c[G.CTR_PAIRS] / steps,
            "misses_per_step": c[G.CTR_MISSES] / steps,
            "violations": c[G.CTR_BAD] + info["errors"],
            "fresh_after_drain": info["segments_free"] == info["num_segments"]
                                 and info["segments_formatted"] == 0
                                 and info["segments_multi"] == 0}


def zipf_sweep(args, G, torch):
    """Batched persistent form: 256 batches with Zipf(1.2) request counts on
    [1, 2^20], 60 % malloc of lognormal(6, 2) bytes clipped to [1, 16 MiB] /
    40 % free of a uniformly chosen live pointer, replayed as one hipGraph;
    the autotuner sweeps segment size, smallest slice and workgroup size."""
    from gallatin_amd import autotune
    from gallatin_amd.batch_streams import zipf_stream
    ops, off, nslots = zipf_stream(args.seed)
    stream = G.BatchStream(ops, off, nslots)
    # the stream's only payload is an 8-byte tag written and read with
    # device-coherent atomics, so the instances skip the allocator's payload
    # fences (GAL_FLAG_NO_PAYLOAD_FENCE)
    best, table = autotune.tune(stream, args.zipf_pool_gib << 30, seed=args.seed,
                                reps=args.zipf_reps, flags=G.GAL_FLAG_NO_PAYLOAD_FENCE)
    launches = None
    if best is not None:
        # the same stream with one launch per batch (plain per-thread and the
        # best batched form), at the best point's geometry
        alloc = G.Gallatin(best["segment_bytes"], best["smallest"], best["biggest"]) \
            .generate_on_device(args.zipf_pool_gib << 30, args.seed, print_info=False,
                                flags=G.GAL_FLAG_NO_PAYLOAD_FENCE)
        try:
            launches = {
                "plain": autotune.measure_launches(alloc, stream, False, 256, args.zipf_reps),
                "batched": autotune.measure_launches(alloc, stream,
                                                     autotune.FORMS[best["form"]],
                                                     best["threads"], args.zipf_reps),
            }
        finally:
            alloc.free_on_device()
    del stream
    torch.cuda.empty_cache()
    if best is None:
        return {"error": "no batched point without violations"}
    plain = [r for r in table if r["path"] == "plain"
             and r["segment_bytes"] == best["segment_bytes"] and r["smallest"] == best["smallest"]]
    plain_grid = [r for r in table if r["path"] == "plain" and r["in_grid"]]
    plain_best = max(plain_grid, key=lambda r: r["pairs_per_s"])
    outside = [r for r in table if r["path"] == "batched" and not r["in_grid"]
               and r["violations"] == 0]
    best_outside = max(outside, key=lambda r: r["pairs_per_s"]) if outside else None
    keys = ("form", "segment_bytes", "smallest", "biggest", "threads")
    out = {
        "workload": "256 batches, Zipf(1.2) counts on [1, 2^20], 60% malloc lognormal(6,2) "
                    "clipped [1, 16 MiB] / 40% free of a random live pointer, u64 tag "
                    "write+verify, one persistent launch per step (hipGraph replay)",
        "flags": "GAL_FLAG_NO_PAYLOAD_FENCE (tag accesses are device-coherent atomics)",
        "pool_bytes": args.zipf_pool_gib << 30,
        "ops": int(off[-1]), "mallocs": int(nslots),
        "grid": "segment {4,16,64} MiB x smallest {16,128} B x workgroup {256,512,1024}; "
                "smallest 256 B reported beside it",
        "value": best["pairs_per_s"], "unit": "pairs/s",
        "ms_per_step": best["ms_per_step"], "p50_step_ms": best["p50_step_ms"],
        "p99_step_ms": best["p99_step_ms"],
        "best": {k: best[k] for k in keys},
        "misses": best["misses"], "violations": best["violations"],
        # the stream kernel's wave time: malloc rounds, frees waiting for the
        # entry an earlier chunk's malloc publishes, frees, the rest
        "wave_time": best.get("wave_time"),
        "plain_same_geometry": plain[0]["pairs_per_s"] if plain else None,
        "plain_best": plain_best["pairs_per_s"],
        "batched_over_plain": best["pairs_per_s"] / plain_best["pairs_per_s"],
        # one kernel launch per batch and one for the cleanup, the way a
        # caller that launches a kernel per batch runs the stream
        "per_batch_launch": None if launches is None else {
            form: {k: r[k] for k in ("pairs_per_s", "ms_per_step", "misses", "violations")}
            for form, r in launches.items()},
        "persistent_batched_over_per_batch_launch_plain":
            None if launches is None else best["pairs_per_s"] / launches["plain"]["pairs_per_s"],
        "outside_grid": None if best_outside is None else dict(
            {k: best_outside[k] for k in keys}, value=best_outside["pairs_per_s"],
            ms_per_step=best_outside["ms_per_step"]),
        "table": [{k: (round(v, 1) if isinstance(v, float) else v) for k, v in r.items()}
                  for r in table],
    }
    return out


def serving_run(args, G, torch):
    """Serving form: 200 GiB global allocator, Poisson(2000) arrivals per
    step, geometric(0.01) lifetimes, lognormal(3, 1) tokens x 256 B appended
    per session per step to a chunked vector (chunks of 128 B doubling to
    64 KiB, then more 64 KiB chunks), admission gated on a reserve of free
    segments, 2^20 hash-table operations per step (70 % insert / 30 %
    remove, keys uniform in [1, 2^28])."""
    from gallatin_amd import serving
    G.init_global_allocator(args.serving_pool_gib << 30, args.seed, print_info=False)
    try:
        r = serving.run(torch, steps=args.serving_steps, warmup=args.serving_warmup,
                        seed=args.seed, reserve_fraction=args.serving_reserve, log=log,
                        gate=args.serving_gate)
        g = G.global_allocator()
        r["allocator_errors"] = g.info()["errors"]
    finally:
        G.free_global_allocator()
    r["workload"] = (f"{args.serving_pool_gib} GiB pool, Poisson(2000) arrivals/step, "
                     "geometric(0.01) lifetime, lognormal(3,1) tokens x 256 B per session-step "
                     "into chunks of 128 B doubling to 64 KiB then 64 KiB each, "
                     f"{args.serving_reserve:.0%} admission reserve of "
                     f"{'free memory in segments' if args.serving_gate else 'free segments'}, "
                     "2^20 table "
                     "ops/step (70% insert / 30% remove, keys in [1, 2^28]), "
                     f"{args.serving_warmup} warm-up steps")
    r["pool_bytes"] = args.serving_pool_gib << 30
    return r


def cpu_baselines(args):
    import oracle
    res = {}
    a = oracle.Allocator(args.pool_gib << 30)
    n = args.cpu_sample_threads
    t0 = time.perf_counter()
    pairs, misses, bad = a.run_churn_sequential(args.seed, 0, n, args.rounds)
    dt = time.perf_counter() - t0
    a.close()
    log(f"oracle replay: {pairs / dt:.4g} pairs/s")
    res["cpu_baseline"] = {
        "value": pairs / dt, "unit": "pairs/s", "cores": 1, "kind": "port",
        "sample": f"oracle sequential replay of stream threads 0..{n - 1} "
                  f"({pairs} pairs, {args.rounds} rounds, fill+verify) in {dt:.1f} s",
    }
    from gallatin_amd import cpu_baseline as C
    cores, online = host_cores()
    r = C.churn(args.seed, 0, args.threads, args.rounds, cores)
    res["cpu_glibc"] = {"value": r["pairs"] / r["seconds"], "unit": "pairs/s", "cores": cores,
                        "online_cores": online, "kind": "glibc malloc/free",
                        "sample": f"all {args.threads} stream threads, {r['pairs']} pairs "
                                  f"in {r['seconds']:.2f} s, pattern errors {r['bad']}"}
    r = C.plain(1 << 20, 10, cores)
    res["cpu_glibc_plain_small"] = {
        "value": r["pairs"] / r["seconds"], "unit": "pairs/s", "cores": cores,
        "kind": "glibc malloc/free",
        "sample": f"the plain small stream, 2^20 ids x 10 passes of malloc(16)/write id/read/"
                  f"free, {r['pairs']} pairs in {r['seconds']:.2f} s, id errors {r['bad']}"}
    if not args.no_zipf:
        from gallatin_amd.batch_streams import zipf_stream
        ops, off, nslots = zipf_stream(args.seed)
        r = C.batches(ops, off, nslots, cores)
        res["cpu_glibc_batched_zipf"] = {
            "value": r["pairs"] / r["seconds"], "unit": "pairs/s", "cores": cores,
            "kind": "glibc malloc/free",
            "sample": f"the same Zipf stream ({len(ops)} ops in {len(off) - 1} batches, a "
                      f"barrier between batches), {r['pairs']} pairs in {r['seconds']:.2f} s, "
                      f"tag errors {r['bad']}"}
    if not args.no_serving:
        import numpy as np
        from gallatin_amd import serving as S
        pool = args.serving_pool_gib << 30
        # the host holds what the device pool holds; keep well inside the
        # host's memory
        avail = host_available_bytes()
        if avail and pool > 0.75 * avail:
            pool = int(0.75 * avail)
        warm, steps = args.serving_warmup, args.cpu_serving_steps
        r, ms = C.serving(args.seed, warm, steps, 1 << 19, 2000.0, 0.01, 1 << 20,
                          S.tokens_table(), pool, args.serving_reserve, cores)
        srt = np.sort(ms)
        res["cpu_glibc_serving"] = {
            "value": float(srt[len(srt) // 2]), "unit": "ms per step (p50)",
            "p99_step_ms": float(srt[min(len(srt) - 1, int(round(0.99 * (len(srt) - 1))))]),
            "cores": cores, "live_sessions": r["live"], "gb_live": r["held_bytes"] / 1e9,
            "refused_sessions": r["refused"], "admitted_sessions": r["admitted"],
            "kind": "glibc malloc/free + std::unordered_map sharded per thread",
            "sample": f"{warm} untimed + {steps} timed steps of the serving step (Poisson(2000) "
                      f"arrivals, geometric(0.01) lifetimes, lognormal(3,1) x 32 words, 2^20 "
                      f"table ops, admission while chunk bytes < {1 - args.serving_reserve:.2f} "
                      f"x {pool / 2**30:.0f} GiB); at the end {r['live']} live sessions holding "
                      f"{r['held_bytes'] / 1e9:.1f} GB and {r['table_keys']} keys, word errors "
                      f"{r['bad']}"}
    if not args.no_big:
        r = C.big(args.seed, 3, args.big_steps, 4096, 64, cores)
        res["cpu_glibc_block_multiseg"] = {
            "value": r["pairs"] / r["seconds"], "unit": "pairs/s", "cores": cores,
            "kind": "glibc malloc/free",
            "sample": f"the same {args.big_steps} block and multi-segment steps, {r['pairs']} "
                      f"pairs (glibc serves every request) in {r['seconds']:.2f} s, tag errors "
                      f"{r['bad']}"}
    return res


def run_path(args, G, torch, path):
    """Warm up, then time exactly args.steps launches of one path."""
    alloc = G.Gallatin(16 << 20, 16, 4096).generate_on_device(
        args.pool_gib << 30, args.seed, print_info=False, flags=args.flags,
        slots0=args.slots0, slots_min=args.slots_min, partitions=args.partitions)
    ctr = torch.zeros(G.CTR_COUNT, dtype=torch.int64, device="cuda")

    def step(seed):
        if path == "batched":
            alloc.run_churn_batched(seed, args.threads, args.rounds, counters=ctr,
                                    workgroups=args.workgroups, threads=args.batch_threads)
        else:
            alloc.run_churn(seed, args.threads, args.rounds, counters=ctr,
                            block_threads=args.block_threads)

    for w in range(args.warmup):
        step(args.seed + w)
    torch.cuda.synchronize()
    ctr.zero_()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    evs[0].record()
    for k in range(args.steps):
        step(args.seed + 100 + k)
        evs[k + 1].record()
    torch.cuda.synchronize()
    total_ms = evs[0].elapsed_time(evs[-1])
    step_ms = sorted(evs[k].elapsed_time(evs[k + 1]) for k in range(args.steps))
    c = ctr.cpu().tolist()
    info = alloc.info()
    alloc.free_on_device()
    pairs = c[G.CTR_PAIRS]
    return {
        "value": pairs / (total_ms / 1e3),
        "ms_per_step": total_ms / args.steps,
        "p50_step_ms": step_ms[len(step_ms) // 2],
        "p99_step_ms": step_ms[min(len(step_ms) - 1, int(round(0.99 * (len(step_ms) - 1))))],
        "pairs_per_step": pairs / args.steps,
        "bytes_per_step": c[G.CTR_BYTES] / args.steps,
        "misses": c[G.CTR_MISSES],
        "violations": c[G.CTR_BAD] + info["errors"],
        "xcd_partitions": info["partitions"],
        "kernel": f"k_churn_batched<{args.batch_threads}>" if path == "batched" else "k_churn",
    }


def churn_one_word(args, G, torch):
    """The headline's request stream with one 64-bit word of each allocation
    written and checked, the pattern of the reference's churn test, through
    the workgroup-batched path and the per-thread path: the allocator rather
    than HBM sets the rate."""
    res = {"workload": f"mixed-size churn: {args.pool_gib} GiB pool, {args.threads} threads "
                       f"x {args.rounds} rounds, log-uniform 1..4095 B, 50% replace per round, "
                       "one u64 written and checked per allocation"}
    for name, batched, threads in (("batched", True, args.batch_threads),
                                   ("plain", False, args.block_threads)):
        alloc = G.Gallatin(16 << 20, 16, 4096).generate_on_device(
            args.pool_gib << 30, args.seed, print_info=False, flags=args.flags,
            slots0=args.slots0, slots_min=args.slots_min, partitions=args.partitions)
        ctr = torch.zeros(G.CTR_COUNT, dtype=torch.int64, device="cuda")
        for w in range(args.warmup):
            alloc.run_churn_word(args.seed + w, args.threads, args.rounds, counters=ctr,
                                 batched=batched, threads=threads)
        torch.cuda.synchronize()
        ctr.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for k in range(args.steps):
            alloc.run_churn_word(args.seed + 100 + k, args.threads, args.rounds, counters=ctr,
                                 batched=batched, threads=threads)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        c = ctr.cpu().tolist()
        info = alloc.info()
        alloc.free_on_device()
        res[name] = {"value": c[G.CTR_PAIRS] / (ms / 1e3), "ms_per_step": ms / args.steps,
                     "misses": c[G.CTR_MISSES], "violations": c[G.CTR_BAD] + info["errors"],
                     "kernel": (f"k_churn_batched_word<{threads}>" if batched
                                else "k_churn_word")}
    res["value"] = res["batched"]["value"]
    res["batched_over_plain"] = res["batched"]["value"] / res["plain"]["value"]
    return res


def section_roofline(pairs_per_s, bytes_per_pair, atomic_peak):
    """Atomic and HBM roofline fractions of a section: two algorithmic
    atomics per pair (one claim, one release) against the measured
    one-per-line atomic ceiling, and its payload bytes against HBM peak."""
    gbs = pairs_per_s * bytes_per_pair / 1e9
    return {"atomics_per_s": 2.0 * pairs_per_s, "atomic_peak": atomic_peak,
            "atomic_frac": 2.0 * pairs_per_s / atomic_peak,
            "payload_bytes_per_pair": bytes_per_pair, "payload_gbs": gbs,
            "hbm_peak_gbs": HBM_PEAK_GBS, "hbm_frac": gbs / HBM_PEAK_GBS}


def add_section_rooflines(out, atomic_peak):
    """Roofline fractions beside the absolute rates of every section."""
    ps = out.get("plain_small")
    if ps:
        for sec in (ps, ps.get("coherent_payload"), ps.get("partitioned_lazy_boot")):
            if sec:
                sec["roofline"] = section_roofline(sec["value"], 8, atomic_peak)
    for key, per_pair in (("block_multiseg", 16), ("batched_zipf", 16), ("churn_one_word", 16)):
        sec = out.get(key)
        if sec and "value" in sec:
            sec["roofline"] = section_roofline(sec["value"], per_pair, atomic_peak)
            if key == "batched_zipf":
                sec["roofline_plain"] = section_roofline(sec["plain_best"], per_pair,
                                                         atomic_peak)
    srv = out.get("serving")
    if srv and "words_per_s" in srv:
        # appended words (8 B each, written once) and the table's 16-byte
        # entries, one per operation; each allocator operation (a chunk
        # malloc or free) is one claim or one release
        gbs = (srv["words_per_s"] * 8 + srv["table_ops_per_s"] * 16) / 1e9
        srv["roofline"] = {"allocator_atomics_per_s": srv["allocator_ops_per_s"],
                           "atomic_peak": atomic_peak,
                           "atomic_frac": srv["allocator_ops_per_s"] / atomic_peak,
                           "payload_gbs": gbs, "hbm_peak_gbs": HBM_PEAK_GBS,
                           "hbm_frac": gbs / HBM_PEAK_GBS}


_T0 = time.perf_counter()


def log(msg):
    """A progress line on stderr (the result lines go to stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def _r(v, digits=4):
    """Round floats to `digits` significant digits (ints and others pass)."""
    if isinstance(v, float):
        return float(f"{v:.{digits}g}")
    if isinstance(v, dict):
        return {k: _r(x, digits) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_r(x, digits) for x in v]
    return v


def _pick(d, *keys):
    return {k: d[k] for k in keys if d and k in d}


def compact(out):
    """The headline record: the driver reads the last stdout line, so it
    carries the metric, the headline roofline and CPU baseline, the plain
    path beside the batched one, and one small summary per workload."""
    keep = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
            "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "p50_step_ms",
            "p99_step_ms", "misses", "violations", "batched_over_plain")
    c = {k: out[k] for k in keep if k in out}
    cfg = out.get("config", {})
    c["config"] = _pick(cfg, "workload", "path", "allocator", "pool_bytes", "threads",
                        "rounds", "xcd_partitions")
    c["roofline"] = _pick(out.get("roofline", {}), "bound", "achieved", "peak", "unit",
                          "frac", "traffic", "kernel", "algorithmic_bytes_per_launch")
    if "atomic_roofline" in out:
        c["atomic_roofline"] = _pick(out["atomic_roofline"], "achieved", "peak", "unit",
                                     "frac", "per_pair")
    if "hbm_ceilings" in out:
        c["hbm_ceilings"] = _pick(out["hbm_ceilings"], "copy_gbs", "scatter_2048_gbs",
                                  "traffic_frac_of_copy")
    if "plain" in out:
        c["plain"] = _pick(out["plain"], "value", "ms_per_step", "p99_step_ms", "misses",
                           "violations", "kernel")
    if "cpu_baseline" in out:
        c["cpu_baseline"] = dict(out["cpu_baseline"])
    if "cpu_glibc" in out:
        c["cpu_glibc"] = _pick(out["cpu_glibc"], "value", "unit", "cores", "online_cores",
                               "kind")
    w = {}
    ps = out.get("plain_small")
    if ps:
        w["plain_small"] = dict(
            _pick(ps, "value", "ms_per_step", "misses", "violations"),
            atomic_frac=ps.get("roofline", {}).get("atomic_frac"),
            coherent_payload=(ps.get("coherent_payload") or {}).get("value"),
            partitioned_lazy_boot=(ps.get("partitioned_lazy_boot") or {}).get("value"),
            cpu_glibc=(out.get("cpu_glibc_plain_small") or {}).get("value"))
    cw = out.get("churn_one_word")
    if cw:
        w["churn_one_word"] = dict(
            value=cw["value"], plain=cw["plain"]["value"],
            batched_over_plain=cw["batched_over_plain"],
            violations=cw["batched"]["violations"] + cw["plain"]["violations"],
            misses=cw["batched"]["misses"] + cw["plain"]["misses"],
            atomic_frac=cw.get("roofline", {}).get("atomic_frac"))
    bm = out.get("block_multiseg")
    if bm:
        w["block_multiseg"] = dict(
            _pick(bm, "value", "ms_per_step", "p50_step_ms", "p99_step_ms", "pairs_per_step",
                  "misses_per_step", "violations", "fresh_after_drain"),
            cpu_glibc=(out.get("cpu_glibc_block_multiseg") or {}).get("value"))
    bz = out.get("batched_zipf")
    if bz:
        w["batched_zipf"] = dict(
            _pick(bz, "value", "ms_per_step", "p50_step_ms", "p99_step_ms", "best", "misses",
                  "violations", "wave_time", "plain_best", "batched_over_plain",
                  "outside_grid", "error"),
            atomic_frac=bz.get("roofline", {}).get("atomic_frac"),
            per_batch_launch_plain=((bz.get("per_batch_launch") or {}).get("plain")
                                    or {}).get("pairs_per_s"),
            over_per_batch_launch_plain=bz.get("persistent_batched_over_per_batch_launch_plain"),
            cpu_glibc=(out.get("cpu_glibc_batched_zipf") or {}).get("value"))
    sv = out.get("serving")
    if sv:
        cpu = out.get("cpu_glibc_serving") or {}
        w["serving"] = dict(
            _pick(sv, "p50_step_ms", "p99_step_ms", "mean_step_ms", "p50_head_ms", "p99_head_ms",
                  "steps", "warmup", "table_ops_per_s",
                  "words_per_s", "allocator_ops_per_s", "live_sessions", "session_words_bad",
                  "table_keys", "refused_sessions", "words_refused", "chunk_failures",
                  "allocator_errors", "gb_live"),
            cpu_glibc_p50_step_ms=cpu.get("value"), cpu_glibc_p99_step_ms=cpu.get("p99_step_ms"),
            cpu_glibc_live_sessions=cpu.get("live_sessions"))
    c["workloads"] = w
    return _r(c)


def emit(out, detail_path=""):
    """Print every section as a JSON line of its own, then the compact
    headline record as the last line (under 4 KB)."""
    if detail_path:
        with open(detail_path, "w") as f:
            json.dump(out, f)
    for key in ("plain_small", "churn_one_word", "block_multiseg", "serving", "hbm_ceilings",
                "tree_search", "atomic_roofline"):
        if key in out:
            print(json.dumps({"section": key, **out[key]}), flush=True)
    bz = out.get("batched_zipf")
    if bz:
        for row in bz.get("table", []):
            print(json.dumps({"section": "batched_zipf.autotune", **row}), flush=True)
        print(json.dumps({"section": "batched_zipf",
                          **{k: v for k, v in bz.items() if k != "table"}}), flush=True)
    for key, v in out.items():
        if key.startswith("cpu_"):
            print(json.dumps({"section": key, **v}), flush=True)
    print(json.dumps(compact(out)), flush=True)


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py measures one GPU: the allocator is a per-GPU instance and this "
              "build runs on a single MI355X; --gpus above 1 is not supported.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import gallatin_amd as G

    torch.cuda.set_device(0)
    paths = ["batched", "plain"] if args.path == "both" else [args.path]
    res = {}
    for p in paths:
        res[p] = run_path(args, G, torch, p)
        log(f"churn {p}: {res[p]['value']:.4g} pairs/s")
    head = res[paths[0]]
    achieved_gbs = head["bytes_per_step"] / (head["ms_per_step"] / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (counter-based request streams generated on the device)",
        "config": {
            "workload": "mixed-size churn: 8 GiB pool, 2^24 threads x 10 rounds, "
                        "log-uniform 1..4095 B, 50% replace per round, fill+verify",
            "path": paths[0],
            "allocator": "Gallatin<16 MiB, 16 B, 4096 B>",
            "pool_bytes": args.pool_gib << 30,
            "threads": args.threads,
            "rounds": args.rounds,
            "slots0": args.slots0,
            "slots_min": args.slots_min,
            "xcd_partitions": head["xcd_partitions"],
            "flags": args.flags,
        },
        "p50_step_ms": head["p50_step_ms"],
        "p99_step_ms": head["p99_step_ms"],
        "pairs_per_step": head["pairs_per_step"],
        "misses": head["misses"],
        "violations": head["violations"],
        "roofline": {
            "bound": "hbm",
            "achieved": achieved_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved_gbs / HBM_PEAK_GBS,
            "traffic": measured_traffic(head["kernel"], args),
            "kernel": head["kernel"],
            "algorithmic_bytes_per_launch": head["bytes_per_step"],
        },
    }
    if "plain" in res and paths[0] != "plain":
        pl = res["plain"]
        out["plain"] = {k: pl[k] for k in ("value", "ms_per_step", "p50_step_ms",
                                            "p99_step_ms", "misses", "violations", "kernel")}
        out["batched_over_plain"] = head["value"] / pl["value"]
    if not args.no_ceilings:
        ceil = atomic_ceiling(G, torch)
        # one claim + one release per pair; the ceiling is one atomic per
        # 128-byte line (lanes of one wave on consecutive words coalesce into
        # one L2 operation, so the "distinct" rate counts lanes, not operations)
        atoms = 2.0 * head["value"]
        out["atomic_roofline"] = {
            "achieved": atoms, "peak": ceil["per_line"], "unit": "atomics/s",
            "frac": atoms / ceil["per_line"], "per_pair": 2, "ceilings": ceil,
        }
        hbm = hbm_ceiling(G, torch)
        traffic = out["roofline"]["traffic"]
        if traffic:
            hbm["traffic_gbs"] = traffic / (head["ms_per_step"] / 1e3) / 1e9
            hbm["traffic_frac_of_copy"] = hbm["traffic_gbs"] / hbm["copy_gbs"]
            hbm["traffic_frac_of_scatter_2048"] = hbm["traffic_gbs"] / hbm["scatter_2048_gbs"]
        out["hbm_ceilings"] = hbm
        out["tree_search"] = tree_probe(args, G, torch)
        log("ceilings measured")
    out["plain_small"] = plain_small(args, G, torch)
    # the same workload with the id written and read device-coherently, so
    # the allocator needs no payload fences (an L2 write-back per free on a
    # pool too small for XCD partitions)
    out["plain_small"]["coherent_payload"] = plain_small(
        args, G, torch, flags=G.GAL_FLAG_NO_PAYLOAD_FENCE)
    # plain payload stores, slots filled on demand so that the 1 GiB pool can
    # take one partition per XCD (the eager boot of every class in every
    # partition needs more segments than it has)
    xcds = max(1, torch.cuda.get_device_properties(0).multi_processor_count // 32)
    out["plain_small"]["partitioned_lazy_boot"] = plain_small(
        args, G, torch, flags=G.GAL_FLAG_LAZY_BOOT, partitions=xcds)
    log("plain small done")
    if not args.no_word_churn:
        out["churn_one_word"] = churn_one_word(args, G, torch)
        log("one-word churn done")
    if not args.no_big:
        out["block_multiseg"] = block_multiseg(args, G, torch)
        log("block and multi-segment done")
    if not args.no_zipf:
        out["batched_zipf"] = zipf_sweep(args, G, torch)
        log("batched Zipf sweep done")
    if not args.no_serving:
        out["serving"] = serving_run(args, G, torch)
        log("serving done")
    if not args.no_ceilings:
        add_section_rooflines(out, ceil["per_line"])
    if not args.no_cpu_baseline:
        out.update(cpu_baselines(args))
        log("cpu baselines done")
    emit(out, args.detail_json)


if __name__ == "__main__":
    main()
