"""Host generators of batched request streams (gal_batch_op arrays).

zipf_stream(): the batched persistent workload: `nbatches` batches whose
request counts follow Zipf(s) on [1, max_count]; each request is a malloc
with probability p_malloc, else a free of a pointer chosen uniformly from
those allocated in earlier batches and not yet freed (a free with nothing
live becomes a malloc).  Malloc sizes are floor(lognormal(mu, sigma)) bytes
clipped to [1, max_size].  Every malloc gets its own slot of the pointer
table, so a stream of M mallocs needs M slots.

All randomness comes from one numpy Generator seeded with `seed`, so a
stream is reproducible from its parameters.
"""
import numpy as np

from . import BATCH_OP_DTYPE


def bounded_zipf(rng, n, s, max_count):
    """n draws of Zipf(s) restricted to [1, max_count] (inverse CDF)."""
    k = np.arange(1, max_count + 1, dtype=np.float64)
    cdf = np.cumsum(k ** -s)
    cdf /= cdf[-1]
    return np.searchsorted(cdf, rng.random(n), side="left").astype(np.int64) + 1


def lognormal_sizes(rng, n, mu, sigma, max_size):
    x = np.floor(np.exp(rng.normal(mu, sigma, n)))
    return np.clip(x, 1, max_size).astype(np.uint64)


def zipf_stream(seed, nbatches=256, s=1.2, max_count=1 << 20, p_malloc=0.6,
                mu=6.0, sigma=2.0, max_size=16 << 20):
    """Returns (ops, offsets, nslots): ops in BATCH_OP_DTYPE, offsets of the
    nbatches + 1 batch boundaries, and the number of pointer-table slots."""
    rng = np.random.default_rng(seed)
    counts = bounded_zipf(rng, nbatches, s, max_count)
    parts = []
    offsets = np.zeros(nbatches + 1, np.uint64)
    live = np.empty(0, np.int64)
    next_slot = 0
    for b in range(nbatches):
        n = int(counts[b])
        is_free = rng.random(n) >= p_malloc
        nf = int(is_free.sum())
        if nf > len(live):  # not enough live pointers: the last frees become mallocs
            extra = nf - len(live)
            idx = np.nonzero(is_free)[0][-extra:]
            is_free[idx] = False
            nf = len(live)
        nm = n - nf
        ops = np.zeros(n, BATCH_OP_DTYPE)
        if nf:
            pick = rng.choice(len(live), nf, replace=False)
            ops["kind"][is_free] = 1
            ops["slot"][is_free] = live[pick].astype(np.uint32)
            keep = np.ones(len(live), bool)
            keep[pick] = False
            live = live[keep]
        new = np.arange(next_slot, next_slot + nm, dtype=np.int64)
        ops["slot"][~is_free] = new.astype(np.uint32)
        ops["arg"][~is_free] = lognormal_sizes(rng, nm, mu, sigma, max_size)
        next_slot += nm
        live = np.concatenate([live, new])
        parts.append(ops)
        offsets[b + 1] = offsets[b] + n
    ops = np.concatenate(parts) if parts else np.zeros(0, BATCH_OP_DTYPE)
    return ops, offsets, next_slot


def check_stream(ops, offsets, nslots):
    """Structural checks of a stream (host): every malloc has a fresh slot,
    every free names a slot allocated in an earlier batch and not freed
    before.  Returns (mallocs, frees)."""
    allocated_in = np.full(nslots, -1, np.int64)
    freed = np.zeros(nslots, bool)
    mallocs = frees = 0
    for b in range(len(offsets) - 1):
        o = ops[int(offsets[b]):int(offsets[b + 1])]
        m = o["kind"] == 0
        ms = o["slot"][m].astype(np.int64)
        assert np.all(allocated_in[ms] == -1), "slot reused by a malloc"
        fs = o["slot"][~m].astype(np.int64)
        assert np.all(allocated_in[fs] >= 0) and np.all(allocated_in[fs] < b), \
            "free of a slot not allocated in an earlier batch"
        assert not np.any(freed[fs]) and len(np.unique(fs)) == len(fs), "double free"
        freed[fs] = True
        allocated_in[ms] = b
        mallocs += int(m.sum())
        frees += int((~m).sum())
    assert mallocs == nslots
    return mallocs, frees
