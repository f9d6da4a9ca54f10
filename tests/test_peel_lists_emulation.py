"""CPU emulation of the segmented removal lists of the peeling ticks
(peel_claim, peel_slot, peel_view, peel_item, peel_slot_lane and
peel_lists_geometry in graphalgo_amd/csrc/algorithms.hip).

A list is PEEL_SEGS segments of segcap items, each with its own counter, and
an overflow segment of ovcap items; a workgroup claims n slots in segment
blockIdx % PEEL_SEGS and spills what does not fit into the overflow.  Readers
number the items segment by segment (the overflow's last) from the clamped
counters.  The emulation replays random claim sequences -- including
segments that overflow and claims that straddle a segment's end -- and checks
that reading items 0..total-1 back, by either lookup, returns every written
item exactly once and nothing else.  A planted mistake (the overflow numbered
from the unclamped counters) must fail.
"""
import numpy as np
import pytest

SEGS = 32


def geometry(nnz):
    lcap = nnz // 2 + 64
    segcap = lcap // (2 * SEGS) + 64
    return segcap, lcap, SEGS * segcap + lcap


def claim(cnt, seg, n, segcap):
    base = cnt[seg]
    cnt[seg] += n
    room = max(segcap - base, 0)
    obase = 0
    if room < n:
        obase = cnt[SEGS]
        cnt[SEGS] += n - room
    return base, obase


def slot(seg, base, obase, r, segcap, ovcap):
    room = max(segcap - base, 0)
    if r < room:
        return seg * segcap + base + r
    o = obase + r - room
    return SEGS * segcap + o if o < ovcap else -1


def view(cnt, segcap, ovcap, clamp_overflow=True):
    sz = [min(c, segcap) for c in cnt[:SEGS]] + [min(cnt[SEGS], ovcap) if clamp_overflow else cnt[SEGS]]
    pre = np.concatenate([[0], np.cumsum(sz)[:-1]])
    return pre, int(sum(sz))


def item_ballot(pre, t, segcap):
    """peel_item: the highest lane s with pre[s] <= t (wave-uniform t)."""
    s = max(i for i in range(SEGS + 1) if pre[i] <= t)
    return s * segcap + (t - pre[s])


def item_lane(pre, t, segcap):
    """peel_slot_lane: binary lifting over the prefix, steps 32..1 clamped to PEEL_SEGS."""
    lo = 0
    step = 32
    while step >= 1:
        cand = min(lo + step, SEGS)
        if pre[cand] <= t:
            lo = cand
        step >>= 1
    return lo * segcap + (t - pre[lo])


@pytest.mark.parametrize("nnz,claims,seed", [(20_000, 300, 1), (2_000, 60, 2), (1_818_806, 2000, 3),
                                              (600, 40, 4)])
def test_claimed_items_read_back_once(nnz, claims, seed):
    rng = np.random.default_rng(seed)
    segcap, ovcap, stride = geometry(nnz)
    lcap = nnz // 2 + 64
    cnt = [0] * (SEGS + 1)
    store = {}
    written = 0
    for c in range(claims):
        # a skewed mix: some workgroups claim far more than a segment holds
        n = int(min(rng.geometric(1 / max(2, lcap // claims)), lcap - written))
        if n <= 0:
            break
        seg = int(rng.integers(0, 4)) if rng.random() < 0.5 else int(rng.integers(0, SEGS))
        base, obase = claim(cnt, seg, n, segcap)
        for r in range(n):
            s = slot(seg, base, obase, r, segcap, ovcap)
            assert 0 <= s < stride and s not in store
            store[s] = written + r
        written += n
    pre, total = view(cnt, segcap, ovcap)
    assert total == written
    got_b = [store[item_ballot(pre, t, segcap)] for t in range(total)]
    got_l = [store[item_lane(pre, t, segcap)] for t in range(total)]
    assert sorted(got_b) == list(range(written)) and got_b == got_l


def test_overflow_is_used_and_planted_mistake_fails():
    segcap, ovcap, _ = geometry(2_000)
    cnt = [0] * (SEGS + 1)
    store = {}
    base, obase = claim(cnt, 0, 3 * segcap, segcap)  # one workgroup overfills segment 0
    for r in range(3 * segcap):
        store[slot(0, base, obase, r, segcap, ovcap)] = r
    assert cnt[SEGS] == 2 * segcap
    pre, total = view(cnt, segcap, ovcap)
    assert sorted(store[item_lane(pre, t, segcap)] for t in range(total)) == list(range(3 * segcap))
    # numbering segment 0 by its raw counter instead of the clamped one loses the overflow's items
    raw = [cnt[0]] + cnt[1:SEGS] + [cnt[SEGS]]
    bad_pre = np.concatenate([[0], np.cumsum(raw)[:-1]])
    with pytest.raises((KeyError, IndexError)):
        [store[item_lane(bad_pre, t, segcap)] for t in range(int(sum(raw)))]
