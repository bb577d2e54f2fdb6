"""The allocation-free traversal (ryo_traverse_f32_noalloc) against ryo_traverse_f32 through the C ABI: two calls in
a row on the same GPU (the second with fewer particles and groups), each equal bit for bit to the allocating
entry point on the same inputs; sizes beyond its fixed work array are refused before any launch."""

import ctypes as C

import numpy as np
import pytest

import ryoanji_amd as R
from ryoanji_amd import workloads as W
from gpu_util import dev, host, oracle_tree, zeros

pytestmark = pytest.mark.gpu


def _inputs(n, seed, group_size):
    s = W.plummer(n, seed=seed)
    t = oracle_tree(s)
    ngroups = (n + group_size - 1) // group_size
    bounds = np.minimum(np.arange(ngroups + 1, dtype=np.int64) * group_size, n).astype(np.uint32)
    d = dict(groups=dev(bounds), arrs=[dev(t[k]) for k in ("x", "y", "z", "m", "h")],
             child=dev(t["tree"]["childOffsets"]), i2l=dev(t["tree"]["internalToLeaf"]), layout=dev(t["layout"]),
             centers=dev(np.ascontiguousarray(t["centers"], np.float32)),
             mp=dev(np.ascontiguousarray(t["mp"], np.float32)))
    return d, ngroups


def _call(ctx, entry, d, ngroups, n, init_node):
    import torch

    p, ax, ay, az = (zeros(n, torch.float32) for _ in range(4))
    pot = C.c_double()
    arrs = [R.ptr(a) for a in d["arrs"]]
    ctx.call(entry, ngroups, R.ptr(d["groups"]), R.ptr(d["groups"]) + 4, init_node, *arrs, *arrs, R.ptr(d["child"]),
             R.ptr(d["i2l"]), R.ptr(d["layout"]), R.ptr(d["centers"]), R.ptr(d["mp"]), 1.0, 0,
             R.host_f32([0, 0, 0]), R.ptr(p), R.ptr(ax), R.ptr(ay), R.ptr(az), C.byref(pot))
    return pot.value, [host(a) for a in (p, ax, ay, az)], ctx.stats()


@pytest.mark.parametrize("init_node", [0, 1])
def test_noalloc_traversal_equals_the_allocating_entry_two_calls(ctx, init_node):
    for n, seed, gsz in ((40000, 11, 64), (6000, 12, 16)):  # the second call: fewer particles and groups
        print(f"noalloc traversal: n={n} seed={seed} group_size={gsz} init_node={init_node}", flush=True)
        d, ngroups = _inputs(n, seed, gsz)
        ref_pot, ref, ref_stats = _call(ctx, "ryo_traverse_f32", d, ngroups, n, init_node)
        pot, out, stats = _call(ctx, "ryo_traverse_f32_noalloc", d, ngroups, n, init_node)
        for a, b in zip(out, ref):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert pot == ref_pot
        assert stats == ref_stats
        assert np.all(np.isfinite(out[1])) and np.any(out[1] != 0)


def test_noalloc_traversal_refuses_more_groups_than_its_array(ctx):
    d, ngroups = _inputs(4096, 13, 64)
    too_many = 1 << 20  # RYO_NOALLOC_MAX_GROUPS
    with pytest.raises(R.RyoError) as e:
        # refused on the host before any launch (the group arrays are never read)
        _call(ctx, "ryo_traverse_f32_noalloc", d, too_many + 1, 4096, 0)
    assert e.value.code == R.RYO_ERR_CAPACITY
