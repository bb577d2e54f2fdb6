"""Helpers for the GPU parity tests: device buffers via torch, calls through the C ABI."""

from __future__ import annotations

import ctypes as C

import numpy as np

import oracle as O
import ryoanji_amd as R


def dev(a):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    elif a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a.copy()).cuda()


def host(t, dtype=None):
    a = t.detach().cpu().numpy()
    if dtype is not None:
        a = a.view(dtype)
    return a


def zeros(n, dtype):
    import torch

    return torch.zeros(int(n), dtype=dtype, device="cuda")


def gpu_keys(ctx, s, kind=0, keys=None, skip_removed=0):
    import torch

    x, y, z = dev(s.x), dev(s.y), dev(s.z)
    k = zeros(s.n, torch.int64) if keys is None else dev(keys)
    ctx.call("ryo_sfc_keys_f32", R.ptr(x), R.ptr(y), R.ptr(z), R.ptr(k), s.n, R.host_f32(s.limits), kind,
             skip_removed)
    return host(k, np.uint64)


def gpu_sort(ctx, keys, values=None):
    import torch

    n = keys.size
    k = dev(keys)
    v = dev(np.arange(n, dtype=np.int32) if values is None else values)
    kb, vb = zeros(n, torch.int64), zeros(n, torch.int32)
    tb = R.lib().ryo_sort_temp_bytes(n)
    tmp = zeros(tb, torch.uint8)
    ctx.call("ryo_sort_pairs_u64_i32", R.ptr(k), R.ptr(v), n, R.ptr(kb), R.ptr(vb), R.ptr(tmp), tb)
    ctx.call("ryo_sort_check", R.ptr(tmp), n)  # no kernel met offsets or lists inconsistent with the keys
    return host(k, np.uint64), host(v)


def gpu_leaves(ctx, sorted_keys, bucket):
    import torch

    n = sorted_keys.size
    k = dev(sorted_keys)
    cap = max(16, n)
    leaves, counts, layout = zeros(cap + 1, torch.int64), zeros(cap, torch.int32), zeros(cap + 1, torch.int32)
    L = C.c_int()
    ctx.call("ryo_build_leaves_u64", R.ptr(k), n, bucket, R.ptr(leaves), R.ptr(counts), R.ptr(layout), cap,
             C.byref(L))
    L = L.value
    return host(leaves[: L + 1], np.uint64), host(counts[:L], np.uint32), host(layout[: L + 1], np.uint32)


def gpu_link(ctx, leaves):
    import torch

    L = leaves.size - 1
    nI = (L - 1) // 7
    n = L + nI
    lv = dev(leaves)
    out = dict(prefixes=zeros(n, torch.int64), childOffsets=zeros(n + 1, torch.int32),
               parents=zeros(max(1, (n - 1) // 8), torch.int32), levelRange=zeros(23, torch.int32),
               internalToLeaf=zeros(n, torch.int32), leafToInternal=zeros(n, torch.int32))
    lrh = (C.c_int32 * 23)()
    ctx.call("ryo_link_octree_u64", R.ptr(lv), L, *(R.ptr(out[k]) for k in ("prefixes", "childOffsets", "parents",
                                                                           "levelRange", "internalToLeaf",
                                                                           "leafToInternal")), lrh)
    res = {k: host(v) for k, v in out.items()}
    res["prefixes"] = res["prefixes"].view(np.uint64)
    res["childOffsets"] = res["childOffsets"][:n]
    res["levelRangeHost"] = np.array(list(lrh), np.int32)
    return res


def gpu_upsweep(ctx, x, y, z, m, layout, tree, limits, inv_theta, mdq=False):
    """float upsweep; mdq=True: CartesianMDQpole multipoles (12 per node)"""
    import torch

    L = tree["numLeaves"] if "numLeaves" in tree else layout.size - 1
    n = L + (L - 1) // 7
    d = {k: dev(tree[k]) for k in ("prefixes", "childOffsets", "parents", "leafToInternal", "levelRange")}
    xs = [dev(a) for a in (x, y, z, m)]
    lay = dev(layout)
    w = 12 if mdq else 8
    c32, q32 = zeros(4 * n, torch.float32), zeros(w * n, torch.float32)
    c64, q64 = zeros(4 * n, torch.float64), zeros(w * n, torch.float64)
    ctx.call("ryo_upsweep_mdq_f32" if mdq else "ryo_upsweep_f32", *(R.ptr(a) for a in xs), R.ptr(lay), L, R.ptr(d["prefixes"]),
             R.ptr(d["childOffsets"]), R.ptr(d["parents"]), R.ptr(d["leafToInternal"]), R.ptr(d["levelRange"]),
             R.host_f32(limits),
             float(inv_theta), R.ptr(c32), R.ptr(q32), R.ptr(c64), R.ptr(q64))
    return (host(c32).reshape(n, 4), host(q32).reshape(n, w), host(c64).reshape(n, 4), host(q64).reshape(n, w))


def gpu_traverse(ctx, x, y, z, m, h, tree, layout, centers, mp, group_size, G=1.0, num_shells=0, box_l=None,
                 init_node=0, stack_cap=None, want_pot=True, bounds=None, mdq=False):
    """group_size-target groups, or the group boundaries `bounds` (G+1 entries) when given; mdq=True: the
    CartesianMDQpole traversal (mp with 12 values per node)"""
    import torch

    n = x.size
    if bounds is None:
        ngroups = (n + group_size - 1) // group_size
        bounds = np.minimum(np.arange(ngroups + 1, dtype=np.int64) * group_size, n)
    ngroups = len(bounds) - 1
    groups = dev(np.asarray(bounds).astype(np.uint32))
    arrs = [dev(a) for a in (x, y, z, m, h)]
    child, i2l = dev(tree["childOffsets"]), dev(tree["internalToLeaf"])
    lay = dev(layout)
    c, q = dev(np.ascontiguousarray(centers, np.float32)), dev(np.ascontiguousarray(mp, np.float32))
    p, ax, ay, az = (zeros(n, torch.float32) for _ in range(4))
    pot = C.c_double()
    bl = R.host_f32(box_l if box_l is not None else [0, 0, 0])
    common = [R.ptr(groups), R.ptr(groups) + 4, init_node, *(R.ptr(a) for a in arrs), *(R.ptr(a) for a in arrs),
              R.ptr(child), R.ptr(i2l), R.ptr(lay), R.ptr(c), R.ptr(q), float(G), num_shells, bl,
              R.ptr(p) if want_pot else None, R.ptr(ax), R.ptr(ay), R.ptr(az), C.byref(pot)]
    if mdq and stack_cap is None:
        ctx.call("ryo_traverse_mdq_f32", ngroups, *common)
    elif mdq:
        ctx.call("ryo_traverse_mdq_f32_stackcap", stack_cap, ngroups, *common)
    elif stack_cap is None:
        ctx.call("ryo_traverse_f32", ngroups, *common)
    else:
        ctx.call("ryo_traverse_f32_stackcap", stack_cap, ngroups, *common)
    return pot.value, host(p), host(ax), host(ay), host(az), ctx.stats()


def gpu_group_splits(ctx, first, last, x, y, z, h, leaves, layout, limits, tol_factor=2.0, extent_ratio=8.0,
                     group_size=64, capacity=None):
    """ryo_group_splits_f32 on device copies of the inputs -> group boundaries (host)"""
    cap = capacity if capacity is not None else int(last - first) + 2
    arrs = [dev(np.ascontiguousarray(a, np.float32)) for a in (x, y, z)]
    hh = None if h is None else dev(np.ascontiguousarray(h, np.float32))
    lv, lay = dev(np.ascontiguousarray(leaves, np.uint64)), dev(np.ascontiguousarray(layout, np.uint32))
    out = dev(np.zeros(cap, np.uint32))
    G = C.c_int()
    ctx.call("ryo_group_splits_f32", int(first), int(last), *(R.ptr(a) for a in arrs), R.ptr(hh), R.ptr(lv),
             len(leaves) - 1, R.ptr(lay), R.host_f32(limits), group_size, float(tol_factor), float(extent_ratio),
             R.ptr(out), cap, C.byref(G))
    return host(out, np.uint32)[:G.value + 1]


def rel_err(ax, ay, az, bx, by, bz):
    d = np.sqrt((ax - bx) ** 2 + (ay - by) ** 2 + (az - bz) ** 2)
    r = np.sqrt(bx.astype(np.float64) ** 2 + by.astype(np.float64) ** 2 + bz.astype(np.float64) ** 2)
    return d / np.maximum(r, 1e-300)


def oracle_tree(s, bucket=64, inv_theta=2.0):
    """Full oracle pipeline: keys, sort, permute, leaves, linked tree, float upsweep."""
    keys = O.sfc_keys(s.x, s.y, s.z, s.limits)
    sk, perm = O.sort_by_key(keys)
    x, y, z, m, h = (a[perm] for a in s.arrays())
    leaves, counts = O.compute_octree(sk, bucket)
    layout = O.layout_from_counts(counts)
    tree = O.link_octree(leaves)
    c, q = O.upsweep(x, y, z, m, leaves, layout, s.limits, inv_theta)
    return dict(keys=keys, sorted_keys=sk, perm=perm, x=x, y=y, z=z, m=m, h=h, leaves=leaves, counts=counts,
                layout=layout, tree=tree, centers=c, mp=q)
