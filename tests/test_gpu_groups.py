"""Split target groups on the GPU (ryo_group_splits_f32) against the oracle, bit-exact, and the traversal over
split groups against the oracle traversal over the same groups."""

import math

import numpy as np
import pytest

import oracle as O
from ryoanji_amd import workloads as W

from gpu_util import gpu_group_splits, gpu_traverse, oracle_tree, rel_err
from test_oracle_groups import _reference_case

pytestmark = pytest.mark.gpu


def test_reference_known_answers(ctx):
    first, last, x, y, z, h, leaves, layout, lim, dc = _reference_case()
    g = gpu_group_splits(ctx, first, last, x, y, z, h, leaves, layout, lim, math.sqrt(3.0) / dc * 1.01, 0.0)
    assert list(g) == [4, 6, 68, 75, 128]
    g = gpu_group_splits(ctx, first, last, x, y, z, h, leaves, layout, lim, math.sqrt(3.0) / dc * 0.98, 0.0)
    assert list(g) == list(range(first, last + 1))


@pytest.mark.parametrize("ratio,expect", [(0.0, 64), (8.0, 8), (20.0, 16), (25.0, 32)])
def test_extent_criterion_closed_form(ctx, ratio, expect):
    n = 64
    x = (np.arange(n, dtype=np.float32) / 128).astype(np.float32)
    y = np.full(n, 0.5, np.float32)
    z = np.full(n, 0.5, np.float32)
    g = gpu_group_splits(ctx, 0, n, x, y, z, None, np.array([0, 1 << 63], np.uint64), np.array([0, n], np.uint32),
                         np.array([0, 1] * 3, np.float32), 0.01, ratio)
    assert list(g) == list(range(0, n + 1, expect))


CASES = {
    "uniform": lambda: W.uniform_cube(50000, seed=31),
    "plummer": lambda: W.plummer(50000, seed=32),
    "hernquist": lambda: W.hernquist_mix(50000, seed=33),
    "box": lambda: W.uniform_box(50000, seed=34, extents=(2.0, 1.0, 0.25)),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("hscale,tol,ratio", [(1.0, 2.0, 8.0), (30.0, 2.0, 8.0), (30.0, 2.0, 0.0), (None, 1.0, 8.0)])
def test_splits_bit_exact(ctx, name, hscale, tol, ratio):
    s = CASES[name]()
    t = oracle_tree(s)
    h = None if hscale is None else (t["h"] * np.float32(hscale)).astype(np.float32)
    first, last = 37, s.n - 11
    o = O.group_splits(first, last, t["x"], t["y"], t["z"], h, t["leaves"], t["layout"], s.limits, tol, ratio)
    g = gpu_group_splits(ctx, first, last, t["x"], t["y"], t["z"], h, t["leaves"], t["layout"], s.limits, tol,
                         ratio)
    assert np.array_equal(g, o), f"{np.count_nonzero(g[:min(g.size, o.size)] != o[:min(g.size, o.size)])} differ"


def test_capacity_error(ctx):
    import ryoanji_amd as R

    s = CASES["plummer"]()
    t = oracle_tree(s)
    o = O.group_splits(0, s.n, t["x"], t["y"], t["z"], t["h"], t["leaves"], t["layout"], s.limits)
    with pytest.raises(R.RyoError) as e:
        gpu_group_splits(ctx, 0, s.n, t["x"], t["y"], t["z"], t["h"], t["leaves"], t["layout"], s.limits,
                         capacity=o.size - 1)
    assert e.value.code == R.RYO_ERR_CAPACITY


@pytest.mark.parametrize("name", ["plummer", "box"])
def test_traversal_over_split_groups(ctx, name):
    """the traversal with split groups (MultipoleHolder's grouping) equals the oracle over the same groups"""
    s = CASES[name]()
    t = oracle_tree(s)
    bounds = gpu_group_splits(ctx, 0, s.n, t["x"], t["y"], t["z"], t["h"] * np.float32(30), t["leaves"], t["layout"],
                              s.limits)
    assert bounds.size - 1 > (s.n + 63) // 64
    args = (t["x"], t["y"], t["z"], t["h"], t["m"])
    otot, _, oax, oay, oaz, ost = O.compute_gravity_groups(t["leaves"], t["layout"], t["centers"], t["mp"], *args,
                                                           s.limits, bounds)
    gtot, _, gax, gay, gaz, gst = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"], t["layout"],
                                               t["centers"], t["mp"], 64, bounds=bounds)
    e = np.sort(rel_err(gax, gay, gaz, oax, oay, oaz))
    assert e[int(0.99 * e.size)] <= 1e-5, e[int(0.99 * e.size)]
    assert gst["sumP2P"] == ost[0]
    assert abs(gtot - otot) <= 1e-5 * abs(otot)
