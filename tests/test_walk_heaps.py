"""CPU: the batched heap procedures of the wave-uniform graph walk, written as
list operations, against std::collections::BinaryHeap's own procedures (push
= append + sift_up; pop = last to the root, sift_down_to_bottom, sift_up;
src/lib.rs:198-303 walks both heaps), over random operation sequences with
heavy ties (21 - 31 distinct distances), comparing the whole arrays after
every step, since the array layout decides which of equal distances pops or
is evicted first.

- frontier (min-heap, Reverse<Candidate>): sift_up with every ancestor read
  at once (a ballot finds the first at or before the new entry, the ones
  below move down a level); pop whose descent reads five levels below its
  position at a time (lane t < 62: depth log2(t + 2), path t + 2 - 2^depth).
- w heap (max-heap, entries in lanes): pop finds its descent path, then every
  entry on it moves up a level at once (each takes its child toward the
  path's end), then the last entry is sifted up with the same all-ancestors
  step.
"""
import random


def _le_min(a, b):      # Reverse order: a <= b iff b.d <= a.d
    return b[0] <= a[0]


def _le_max(a, b):
    return a[0] <= b[0]


def _sift_up(h, pos, e, le):
    while pos > 0:
        par = (pos - 1) // 2
        if le(e, h[par]):
            break
        h[pos] = h[par]
        pos = par
    h[pos] = e


def _pop(h, le):
    item = h.pop()
    if not h:
        return item
    top = h[0]
    end, pos, child = len(h), 0, 1
    while end >= 2 and child <= end - 2:
        child += 1 if le(h[child], h[child + 1]) else 0
        h[pos] = h[child]
        pos, child = child, 2 * child + 1
    if child == end - 1:
        h[pos] = h[child]
        pos = child
    _sift_up(h, pos, item, le)
    return top


def _depth(x):
    return x.bit_length() - 1


def _sift_up_all(h, pos, e, stop_at):
    """All ancestors at once: lane j holds ancestor j; J = the first j whose
    entry stops the walk up (stop_at(e, ancestor)); the ancestors below J move
    down a level; e lands at ancestor J - 1."""
    I = pos + 1
    dp = _depth(I)
    anc = {j: h[(I >> j) - 1] for j in range(1, dp + 1)}
    stops = [j for j in range(1, dp + 1) if stop_at(e, anc[j])]
    J = stops[0] if stops else dp + 1
    for j in range(1, J):
        h[(I >> (j - 1)) - 1] = anc[j]
    h[(I >> (J - 1)) - 1] = e


def _fr_pop_batched(h):
    item = h.pop()
    if not h:
        return item
    top = h[0]
    end, pos, child = len(h), 0, 1
    while end >= 2 and child <= end - 2:
        v = {}
        for t in range(62):
            k = _depth(t + 2)
            idx = ((pos + 1) << k) + (t + 2 - (1 << k)) - 1
            if idx < end:
                v[t] = h[idx]
        rel, lvl = 1, 0
        while lvl < 5 and end >= 2 and child <= end - 2:
            c0, c1 = v[2 * rel - 2], v[2 * rel - 1]
            right = c1[0] <= c0[0]
            h[pos] = c1 if right else c0
            pos = child + (1 if right else 0)
            rel = 2 * rel + (1 if right else 0)
            child = 2 * pos + 1
            lvl += 1
    if child == end - 1:
        h[pos] = h[child]
        pos = child
    _sift_up_all(h, pos, item, lambda e, a: a[0] <= e[0])
    return top


def _w_pop_lanes(L, n):
    item = L[n - 1]
    n -= 1
    if n == 0:
        return n
    end, pos, child = n, 0, 1
    while end >= 2 and child <= end - 2:
        right = L[child][0] <= L[child + 1][0]
        pos = child + (1 if right else 0)
        child = 2 * pos + 1
    if child == end - 1:
        pos = child
    I = pos + 1
    dp = _depth(I)
    old = list(L)
    for lane in range(64):
        k = dp - _depth(lane + 1)
        if k >= 1 and (I >> k) == lane + 1:
            L[lane] = old[(I >> (k - 1)) - 1]
    _sift_up_all(L, pos, item, lambda e, a: e[0] <= a[0])
    return n


def test_frontier_batched_procedures_match_binaryheap():
    rng = random.Random(1)
    for _ in range(400):
        a, b, nid = [], [], 0
        for _ in range(rng.randint(1, 700)):
            if a and rng.random() < 0.4:
                assert _pop(a, _le_min) == _fr_pop_batched(b)
            else:
                c = (float(rng.randint(0, 20)), nid)
                nid += 1
                a.append(c)
                _sift_up(a, len(a) - 1, c, _le_min)
                b.append(None)
                _sift_up_all(b, len(b) - 1, c, lambda e, x: x[0] <= e[0])
            assert a == b


def test_w_heap_lane_procedures_match_binaryheap():
    rng = random.Random(2)
    for _ in range(400):
        ref, L, n, nid = [], [(0.0, -1)] * 64, 0, 0
        wb = rng.randint(1, 64)
        for _ in range(rng.randint(1, 300)):
            c = (float(rng.randint(0, 30)), nid)
            nid += 1
            if n < wb:
                ref.append(c)
                _sift_up(ref, len(ref) - 1, c, _le_max)
                n += 1
                _sift_up_all(L, n - 1, c, lambda e, a: e[0] <= a[0])
            elif c[0] < ref[0][0]:   # the walk's replace: pop the top, push c
                _pop(ref, _le_max)
                ref.append(c)
                _sift_up(ref, len(ref) - 1, c, _le_max)
                n = _w_pop_lanes(L, n) + 1
                _sift_up_all(L, n - 1, c, lambda e, a: e[0] <= a[0])
            assert ref == L[:n]


def test_a_planted_wrong_tie_rule_is_caught():
    # the same check with the frontier's tie rule flipped (a.d < e.d stops the
    # walk up instead of a.d <= e.d) must fail somewhere
    rng = random.Random(3)
    caught = False
    for _ in range(200):
        a, b, nid = [], [], 0
        for _ in range(200):
            c = (float(rng.randint(0, 5)), nid)
            nid += 1
            a.append(c)
            _sift_up(a, len(a) - 1, c, _le_min)
            b.append(None)
            _sift_up_all(b, len(b) - 1, c, lambda e, x: x[0] < e[0])
            if a != b:
                caught = True
                break
        if caught:
            break
    assert caught
