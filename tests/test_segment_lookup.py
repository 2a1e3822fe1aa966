"""CPU model of the product-segment lookup in the row path's expansion
(for_products, tsg_expand.h).  A batch covers a window of G consecutive
products of a group of G A entries; every non-empty segment (the products of
one A entry) that overlaps the window writes its lane at its first position
in the window (position 0 for a segment that started earlier), and a running
maximum over the window gives each position its segment.  This checks the
rule against a binary search over the segment ends, including empty segments,
windows at every offset and batch striding, and that no window position ever
gets two writers (the kernel's LDS writes would race)."""
from __future__ import annotations

import numpy as np
import pytest


def lookup(cnt, qb, G):
    incl = np.cumsum(cnt)
    start = incl - cnt
    mark = np.full(G, -1)
    for s in range(G):
        if cnt[s] > 0 and start[s] < qb + G and incl[s] > qb:
            pos = start[s] - qb if start[s] > qb else 0
            assert mark[pos] == -1, "two writers at one window position"
            mark[pos] = s
    return np.maximum.accumulate(mark)


@pytest.mark.parametrize("G", [16, 32, 64])
def test_segment_lookup_matches_search(G):
    rng = np.random.default_rng(G)
    for trial in range(1500):
        p0 = rng.random()
        hi = [1, 3, 40, 200][trial % 4]
        cnt = np.where(rng.random(G) < p0, 0, rng.integers(0, hi + 1, G))
        incl = np.cumsum(cnt)
        total = int(incl[-1])
        bstep = 1 + trial % 2
        for base in range(0, total, G * bstep):
            seg = lookup(cnt, base, G)
            for l in range(G):
                q = base + l
                if q < total:
                    assert seg[l] == int(np.searchsorted(incl, q, side="right"))


def test_all_products_in_one_segment():
    cnt = np.zeros(64, dtype=np.int64)
    cnt[5] = 1000
    for qb in (0, 64, 960):
        assert (lookup(cnt, qb, 64) == 5).all()
