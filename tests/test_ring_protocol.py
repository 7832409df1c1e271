"""Exhaustive interleavings of the segment ring's completion protocol (CPU).

return_block (include/gallatin/allocators/gallatin.hpp) completes ring
tickets through one word per segment: (frontier, written-mask over the W
tickets from the frontier, ticket x at bit x mod W).  A returner

  1. waits until its ticket is within W of the frontier (load, retry),
  2. sets its ticket's bit with an atomic or,
  3. leaves if the frontier has passed its ticket or the frontier's bit is
     clear; otherwise moves the frontier over the whole run of set bits
     with one compare-and-swap, clearing them (on failure: back to 3 with
     the value seen), and leaves.

This file models those steps as a state machine, one atomic step at a time,
with the returners' tickets already drawn (the ticket counter is a plain
fetch-and-add), and explores every interleaving for small windows and
returner counts.  It checks what the reference's in-order spin
(memory_table.cuh:906-910) guarantees and what the allocator relies on:
every ticket is completed exactly once, runs complete in ticket order (the
count never covers an unwritten entry), a mark never finds its bit already
set, a returner never changes the word after its ticket was completed by
another (a compare-and-swap it decided on before then fails; on the device
a later set-up of the segment also changes the word's generation), and from
every reachable state all returners can finish (no stuck state).

The model restates the protocol; it pins the protocol's design, and the GPU
tests (ring reuse known answers, recycling streams against the oracle,
fresh-after-drain checks) pin the compiled code.
"""
import pytest


def _explore(W, N, leave_when_passed=True, reach=None):
    """States explored and states that cannot finish.  `leave_when_passed`
    and `reach` (the distance from the frontier at which a returner may
    mark, W in the protocol) exist to plant mistakes."""
    FULL = (1 << W) - 1
    reach = W if reach is None else reach

    def rot(bits, r):
        return ((bits >> r) | (bits << (W - r))) & FULL if r else bits

    def rotl(bits, r):
        return ((bits << r) | (bits >> (W - r))) & FULL if r else bits

    # word = (frontier, bits); a returner = (pc, seen frontier, seen bits, marked,
    # pending compare-and-swap (next frontier, next bits, k)); pc 0 load, 1 decide,
    # 2 or, 3 compare-and-swap, 4 done
    def step(word, th, pos):
        pc, wf, wb, marked, nxt = th
        if pc == 0:
            return word, (1, word[0], word[1], marked, None), None
        if pc == 1:  # local decision, folded into the next atomic step
            ahead = pos - wf
            if not marked:
                if ahead >= reach:
                    return word, (0, wf, wb, marked, None), None
                return word, (2, wf, wb, marked, None), None
            if ahead < 0 and leave_when_passed:  # completed by another returner
                return word, (4, 0, 0, marked, None), None
            r = wf % W
            rb = rot(wb, r)
            if not rb & 1:
                return word, (4, 0, 0, marked, None), None
            k = W if rb == FULL else ((~rb) & -(~rb)).bit_length() - 1
            run = FULL if k == W else (1 << k) - 1
            return word, (3, wf, wb, marked, (wf + k, wb & ~rotl(run, r) & FULL, k)), None
        if pc == 2:
            assert not (word[1] >> (pos % W)) & 1, "mark found its bit set"
            nw = (word[0], word[1] | (1 << (pos % W)))
            return nw, (1, nw[0], nw[1], True, None), None
        if pc == 3:
            if word == (wf, wb):
                # a compare-and-swap decided before the ticket was completed
                # by another returner must fail (the frontier moved)
                assert word[0] <= pos, "changed the word after its ticket was completed"
                return (nxt[0], nxt[1]), (4, 0, 0, marked, None), (wf, nxt[2])
            return word, (1, word[0], word[1], marked, None), None
        raise AssertionError(pc)

    init = ((0, 0), tuple((0, 0, 0, False, None) for _ in range(N)), ())
    edges = {}
    stack = [init]
    while stack:
        st = stack.pop()
        if st in edges:
            continue
        word, ths, runs = st
        succ = []
        for i, th in enumerate(ths):
            if th[0] == 4:
                continue
            nw, nth, run = step(word, th, i)
            succ.append((nw, ths[:i] + (nth,) + ths[i + 1:], runs + ((run,) if run else ())))
        edges[st] = succ
        stack.extend(succ)
        if not succ:  # everyone left
            assert word == (N, 0), (word, runs)
            done = [t for f, k in runs for t in range(f, f + k)]
            assert done == list(range(N)), runs
    # every reachable state can still reach a final one
    rev = {}
    for s, succ in edges.items():
        for t in succ:
            rev.setdefault(t, []).append(s)
    good = {s for s, succ in edges.items() if not succ}
    frontier = list(good)
    while frontier:
        t = frontier.pop()
        for s in rev.get(t, ()):
            if s not in good:
                good.add(s)
                frontier.append(s)
    return len(edges), len(edges) - len(good)


@pytest.mark.parametrize("W,N", [(2, 3), (2, 4), (3, 4), (4, 4), (2, 5), (3, 5)])
def test_every_interleaving_completes_in_order(W, N):
    states, stuck = _explore(W, N)
    assert states > 0 and stuck == 0


@pytest.mark.parametrize("mistake", [dict(leave_when_passed=False), dict(reach=3)])
def test_planted_mistakes_are_caught(mistake):
    """Dropping the rule that a returner leaves once its ticket was completed
    (it then moves the frontier again from a later view), or letting a
    ticket mark one window too far ahead (its bit still held by the ticket W
    before it), must fail the exploration."""
    with pytest.raises(AssertionError):
        _explore(2, 4, **mistake)
