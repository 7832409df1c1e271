"""Serving loop driver (gal_serving_*): sessions on the global allocator.

Workload: arrivals per step ~ Poisson(arrivals_mean); a session lives a
geometric(p_end) number of steps; each step a live session appends
lognormal(mu, sigma) tokens x 256 B (32 words of 8 B each) to its
chunked_vector (chunks of 128 B doubling to 64 KiB, then further 64 KiB
chunks, so an append is never cut short for want of capacity); sessions are
admitted only while the allocator has more than `reserve_segments` free
segments; beside them
ht_ops operations per step (70 % insert / 30 % remove, keys uniform in
[1, 2^28]) run on an extendible hash table that grows on the same
allocator.  Step latency is measured with events around each step.
"""
import ctypes

import numpy as np

from . import GallatinError, _check, _dev_u64, _p, _stream, lib

# counter indices of gal_serving_step
ADMITTED, REFUSED, ENDED, WORDS, WORDS_REFUSED, CHUNKS, CHUNK_FAILED, CHUNKS_FREED, \
    HT_INSERTED, HT_REMOVED, HT_ABSENT, HT_FAILED = range(12)


def tokens_table(mu=3.0, sigma=1.0, n=1024):
    """Quantiles of floor(lognormal(mu, sigma)), at least 1 token."""
    from scipy.special import ndtri
    q = ndtri((np.arange(n) + 0.5) / n)
    return np.maximum(1, np.floor(np.exp(mu + sigma * q))).astype(np.uint32)


class Serving:
    def __init__(self, max_sessions, seed=42, mu=3.0, sigma=1.0):
        self.handle = _p()
        table = np.ascontiguousarray(tokens_table(mu, sigma))
        _check(lib().gal_serving_create(max_sessions, seed,
                                        table.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.byref(self.handle)))

    def step(self, step, arrivals, reserve_segments, p_end, ht_ops, counters, gate=0):
        """gate 0: admit while the segment tree has more than reserve_segments
        free segments; gate 1: while the free memory (free segments plus the
        free blocks of formatted segments, in segments) exceeds it."""
        _check(lib().gal_serving_step(self.handle, step, arrivals, reserve_segments, gate,
                                      int(p_end * 2 ** 32), ht_ops, counters.data_ptr(),
                                      _stream()))

    def gate(self):
        """(free segments, free memory in segments) at the last step."""
        v = [ctypes.c_uint64() for _ in range(2)]
        _check(lib().gal_serving_gate(self.handle, *[ctypes.byref(x) for x in v]))
        return v[0].value, v[1].value

    def check(self):
        v = [ctypes.c_uint64() for _ in range(4)]
        _check(lib().gal_serving_check(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(("live", "words", "bad", "held_bytes"), (x.value for x in v)))

    def table(self):
        v = [ctypes.c_uint64() for _ in range(3)]
        _check(lib().gal_serving_table(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(("keys", "buckets", "failed"), (x.value for x in v)))

    def destroy(self):
        if self.handle:
            _check(lib().gal_serving_destroy(self.handle))
            self.handle = _p()


def run(torch, steps, warmup, seed=42, max_sessions=1 << 19, arrivals_mean=2000,
        p_end=0.01, ht_ops=1 << 20, reserve_fraction=0.10, num_segments=None, log=None,
        trace_every=50, gate=0, head_steps=30):
    """Run warmup + steps steps on the current global allocator; returns
    per-step latencies of the timed steps, the counters of all steps, and a
    trace of the warm-up every `trace_every` steps (free segments, live
    sessions, admissions and refusals so far), also passed to `log`.  The
    latencies are summarised over all timed steps and over the first
    `head_steps` of them (the state the warm-up reached)."""
    from . import global_allocator
    if num_segments is None:
        num_segments = global_allocator().num_segments
    rng = np.random.default_rng(seed)
    arrivals = rng.poisson(arrivals_mean, warmup + steps)
    reserve = int(num_segments * reserve_fraction)
    sv = Serving(max_sessions, seed)
    ctr = torch.zeros(16, dtype=torch.int64, device="cuda")
    events0 = global_allocator().stats()
    try:
        trace = []
        for k in range(warmup):
            sv.step(k, int(arrivals[k]), reserve, p_end, ht_ops, ctr, gate)
            if trace_every and (k + 1) % trace_every == 0:
                c = ctr.cpu().tolist()
                fs, fm = sv.gate()
                t = {"step": k + 1, "segments_free": fs, "free_memory_segments": fm,
                     "live": c[ADMITTED] - c[ENDED], "admitted": c[ADMITTED],
                     "refused": c[REFUSED], "chunk_failed": c[CHUNK_FAILED]}
                trace.append(t)
                if log:
                    log(f"serving warm-up {t}")
        torch.cuda.synchronize()
        warm = ctr.cpu().tolist()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        evs[0].record()
        for k in range(steps):
            sv.step(warmup + k, int(arrivals[warmup + k]), reserve, p_end, ht_ops, ctr, gate)
            evs[k + 1].record()
        torch.cuda.synchronize()
        c = ctr.cpu().tolist()
        ms = [evs[k].elapsed_time(evs[k + 1]) for k in range(steps)]
        total = evs[0].elapsed_time(evs[-1]) / 1e3
        chk = sv.check()
        tab = sv.table()
        segments_free = global_allocator().info()["segments_free"]
        gate_end = sv.gate()
        events1 = global_allocator().stats()
    finally:
        sv.destroy()
    timed = [c[i] - warm[i] for i in range(16)]
    srt = sorted(ms)
    head = sorted(ms[:head_steps]) if head_steps else srt
    return {
        "steps": steps, "warmup": warmup, "reserve_segments": reserve, "gate": gate,
        "trace": trace,
        "p50_step_ms": srt[len(srt) // 2],
        "p99_step_ms": srt[min(len(srt) - 1, int(round(0.99 * (len(srt) - 1))))],
        "mean_step_ms": total * 1e3 / steps,
        "max_step_ms": srt[-1],
        # the first head_steps timed steps: the state right after warm-up
        "head_steps": len(head), "p50_head_ms": head[len(head) // 2],
        "p99_head_ms": head[min(len(head) - 1, int(round(0.99 * (len(head) - 1))))],
        "table_ops_per_s": steps * ht_ops / total,
        "allocator_ops_per_s": (timed[CHUNKS] + timed[CHUNKS_FREED]) / total,
        "words_per_s": timed[WORDS] / total,
        "sessions_admitted_per_s": timed[ADMITTED] / total,
        "live_sessions": chk["live"], "session_words_bad": chk["bad"],
        "session_words": chk["words"], "gb_live": chk["held_bytes"] / 1e9,
        "segments_free_end": segments_free, "free_memory_segments_end": gate_end[1],
        "table_keys": tab["keys"], "table_inserts_refused": c[HT_FAILED],
        # refusals of the timed steps (the admission gate in steady state)
        "refused_sessions": timed[REFUSED], "admitted_sessions": timed[ADMITTED],
        "words_refused": c[WORDS_REFUSED], "chunk_failures": c[CHUNK_FAILED],
        "counters": {n: c[i] for i, n in enumerate(
            ("admitted", "refused", "ended", "words", "words_refused", "chunks",
             "chunk_failed", "chunks_freed", "ht_inserted", "ht_removed", "ht_absent",
             "ht_failed"))},
        # the global allocator's event counters (gal_get_stats) over the run
        "allocator_events": {k: events1[k] - events0[k] for k in events1
                             if events1[k] != events0[k]},
    }
