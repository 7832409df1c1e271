"""Exhaustive interleavings of the persistent stream's chunk schedule (CPU).

k_stream (gallatin_amd/csrc/batches.hip) runs a batched request stream as
one launch: workgroups take chunks in stream order from a queue counter; a
workgroup serves a chunk's mallocs (publishing their table entries) as soon
as it takes the chunk, and runs the chunk's frees after it has served its
next chunk's mallocs.  A free waits until the entry it frees is published.
The stream guarantees that a chunk's frees only name entries of earlier
chunks (a chunk never spans batches, and a batch frees only pointers of
earlier batches).

This file models the schedule as a state machine, one step at a time, and
explores every interleaving of small streams: every reachable state must
still be able to finish (every chunk served, every free run, every
workgroup gone).  It also explores the order the kernel used before (frees
right after the chunk's own mallocs) and a variant measured and dropped
(the previous chunk's frees before the chunk's entries are published), and
checks that a stream breaking the guarantee is caught.

The model restates the schedule; the GPU tests of the batched streams pin
the compiled kernel.
"""
import itertools
import random

import pytest


def _explore(W, deps, order="deferred"):
    """deps[k]: the chunks whose entries chunk k's frees wait for.  Returns
    (states explored, states that cannot finish)."""
    C = len(deps)
    dep_mask = [sum(1 << j for j in d) for d in deps]

    # a workgroup: (pc, cur, prev); pc 0 take, 1 serve cur, 2 free prev,
    # 3 drain (queue empty: free prev, then leave), 4 gone; -1: no chunk
    def step(head, pub, wg):
        pc, cur, prev = wg
        if pc == 0:  # take the next chunk
            if head >= C:
                return head + 1, pub, (3, -1, prev)
            if order == "overlap":  # previous chunk's frees before publishing
                return head + 1, pub, (2, head, prev)
            return head + 1, pub, (1, head, prev)
        if pc == 1:  # serve and publish cur's mallocs
            pub |= 1 << cur
            if order == "immediate":  # the chunk's own frees next
                return head, pub, (2, cur, cur)
            if order == "overlap":  # frees done before: take the next chunk
                return head, pub, (0, -1, cur)
            return head, pub, (2, cur, prev)
        if pc in (2, 3):  # frees of prev: wait until their entries are published
            if prev >= 0 and (pub & dep_mask[prev]) != dep_mask[prev]:
                return None
            if pc == 3:
                return head, pub, (4, -1, -1)
            if order == "overlap":
                return head, pub, (1, cur, -1)
            if order == "immediate":
                return head, pub, (0, -1, -1)
            return head, pub, (0, -1, cur)
        return None

    init = (0, 0, tuple((0, -1, -1) for _ in range(W)))
    edges = {}
    stack = [init]
    while stack:
        st = stack.pop()
        if st in edges:
            continue
        head, pub, wgs = st
        succ = []
        for i, wg in enumerate(wgs):
            nxt = step(head, pub, wg)
            if nxt is None:
                continue
            h2, p2, w2 = nxt
            succ.append((min(h2, C + W), p2, wgs[:i] + (w2,) + wgs[i + 1:]))
        edges[st] = succ
        stack.extend(succ)
    final = [s for s, succ in edges.items() if all(w[0] == 4 for w in s[2])]
    for s in final:
        assert s[1] == (1 << C) - 1, "a chunk was never served"
    rev = {}
    for s, succ in edges.items():
        for t in succ:
            rev.setdefault(t, []).append(s)
    good = set(final)
    todo = list(final)
    while todo:
        t = todo.pop()
        for s in rev.get(t, ()):
            if s not in good:
                good.add(s)
                todo.append(s)
    return len(edges), len(edges) - len(good)


def _streams(C, n, seed):
    """n random valid dependency lists of C chunks (frees name earlier chunks)."""
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        out.append([sorted(rng.sample(range(k), rng.randint(0, k))) if k else []
                    for k in range(C)])
    return out


@pytest.mark.parametrize("order", ["deferred", "immediate", "overlap"])
@pytest.mark.parametrize("W,C", [(1, 4), (2, 4), (2, 5), (3, 4), (3, 5)])
def test_every_interleaving_drains(order, W, C):
    for deps in _streams(C, 6, seed=W * 10 + C):
        states, stuck = _explore(W, deps, order)
        assert states > 0 and stuck == 0, (order, W, deps)


def test_chains_through_every_chunk_drain():
    """Each chunk's frees wait on the chunk before it, the longest chain."""
    for W, C in itertools.product((1, 2, 3), (3, 5)):
        deps = [[k - 1] if k else [] for k in range(C)]
        assert _explore(W, deps, "deferred")[1] == 0


@pytest.mark.parametrize("order,W,deps", [
    # a free that names a chunk two ahead: one workgroup is still holding
    # those frees when it would have to serve that chunk
    ("deferred", 1, [[2], [], [], []]),
    # a free that names the next chunk, under the old immediate order
    ("immediate", 1, [[1], [], []]),
])
def test_planted_stream_violations_are_caught(order, W, deps):
    """Streams whose frees name later chunks break the guarantee the schedule
    relies on; the exploration must find states that cannot finish."""
    assert _explore(W, deps, order)[1] > 0
