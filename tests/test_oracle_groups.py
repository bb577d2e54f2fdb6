"""Target-group splitting of the oracle (no GPU): the reference's known answers (groups.cu:224-297, the
distance and interaction-radius criteria) and closed forms for the window-extent criterion."""

import math

import numpy as np
import pytest

import oracle as O


def _reference_case():
    """groups.cu:224-251: particles 4..127 on the diagonal of Box(0, 128); leaves of OctreeMaker
    divide().divide(2) with counts {4,1,8,8,8,8,31,8,8,8,16,16,16,0,0}"""
    first, last = 4, 128
    lvl1 = [i << 60 for i in range(8)]
    leaves = sorted(set(lvl1[:2] + [(2 << 60) + (j << 57) for j in range(8)] + lvl1[3:])) + [1 << 63]
    leaves = np.array(leaves, np.uint64)
    counts = np.array([4, 1, 8, 8, 8, 8, 31, 8, 8, 8, 16, 16, 16, 0, 0], np.uint32)
    assert leaves.size == counts.size + 1
    layout = O.layout_from_counts(counts)
    x = np.arange(last, dtype=np.float32)
    y, z = x.copy(), x.copy()
    h = np.full(last, 128.0, np.float32)
    h[first + 64 + 6] = 0.99 * math.sqrt(3.0) / 2  # just small enough to split
    h[first + 64 + 7] = 1.01 * math.sqrt(3.0) / 2  # just large enough not to
    for a in (x, y, z):
        a[5] -= 0.01  # a larger distance between particles 5 and 6
    lim = np.array([0, 128] * 3, np.float32)
    dist_crit = (128.0 ** 3 / 64) ** (1 / 3)
    return first, last, x, y, z, h, leaves, layout, lim, dist_crit


def test_reference_known_answers():
    first, last, x, y, z, h, leaves, layout, lim, dc = _reference_case()
    g = O.group_splits(first, last, x, y, z, h, leaves, layout, lim, math.sqrt(3.0) / dc * 1.01, 0.0)
    assert list(g) == [4, 6, 68, 75, 128]
    # a tolerance below the particle spacing splits every pair: groups of one, 64 + 60 of them (the reference
    # uses 0.99, which puts particles 4 and 5, 0.99 sqrt(3) apart, exactly on the threshold; in float the tie
    # falls the other way, so 0.98)
    g = O.group_splits(first, last, x, y, z, h, leaves, layout, lim, math.sqrt(3.0) / dc * 0.98, 0.0)
    assert list(g) == list(range(first, last + 1))


@pytest.mark.parametrize("ratio,expect", [(0.0, 64), (50.0, 64), (8.0, 8), (20.0, 16), (25.0, 32), (49.0, 32)])
def test_extent_criterion_closed_form(ratio, expect):
    """64 particles spaced 1/128 along x in the unit box, one leaf (edge 1), tolerance 0.01: the window extent
    is 63/128 = 49.2 allowed extents, so it is cut into chunks of 64 / 2^k with chunk * 49.2 <= 64 * ratio"""
    n = 64
    x = (np.arange(n, dtype=np.float32) / 128).astype(np.float32)
    y = np.full(n, 0.5, np.float32)
    z = np.full(n, 0.5, np.float32)
    leaves = np.array([0, 1 << 63], np.uint64)
    layout = np.array([0, n], np.uint32)
    lim = np.array([0, 1] * 3, np.float32)
    g = O.group_splits(0, n, x, y, z, None, leaves, layout, lim, 0.01, ratio)
    assert list(g) == list(range(0, n + 1, expect))


def test_groups_partition_the_targets():
    rng = np.random.default_rng(3)
    from ryoanji_amd import workloads as W

    s = W.plummer(20000, seed=9)
    sk, perm = O.sort_by_key(O.sfc_keys(s.x, s.y, s.z, s.limits))
    x, y, z, m, h = (a[perm] for a in s.arrays())
    leaves, counts = O.compute_octree(sk, 64)
    layout = O.layout_from_counts(counts)
    first, last = 100, 19000
    g = O.group_splits(first, last, x, y, z, h * 50, leaves, layout, s.limits)
    d = np.diff(g.astype(np.int64))
    assert g[0] == first and g[-1] == last and np.all(d >= 1) and np.all(d <= 64)
    # no group crosses a window boundary
    win = (g[:-1].astype(np.int64) - first) // 64
    assert np.all((g[1:].astype(np.int64) - 1 - first) // 64 == win)
    assert g.size - 1 > (last - first + 63) // 64  # the Plummer tails do get split
    del rng
