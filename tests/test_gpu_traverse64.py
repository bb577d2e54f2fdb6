"""The double-coordinate traversal instantiations through the C ABI against the reference's own CPU computeGravity
instantiated for the same types (oracle/_ref): (d,d,d,d,d) with CartesianQuadrupole<double> and (d,f,f,f,d) with
CartesianQuadrupole<float>. Trees come from the oracle, double centers and multipoles from the reference's FP64
upsweep on double coordinates; parity uses the reference CPU's 16-target groups."""

import ctypes as C

import numpy as np
import pytest

import oracle as O
import ryoanji_amd as R
from ryoanji_amd import workloads as W

from gpu_util import dev, host, oracle_tree, rel_err, zeros

pytestmark = pytest.mark.gpu


def _double_system(s, seed):
    """the system's float coordinates moved by ~1e-9 of the box, so that they need double precision"""
    t = oracle_tree(s)
    rng = np.random.default_rng(seed)
    ext = float(np.max(s.limits[1::2] - s.limits[0::2]))
    xyz = [t[k].astype(np.float64) + rng.uniform(-1e-9, 1e-9, s.n) * ext for k in ("x", "y", "z")]
    lim = s.limits.astype(np.float64)
    c64, q64 = O.upsweep(*xyz, t["m"].astype(np.float64), t["leaves"], t["layout"], lim, 2.0, precision="f64",
                         lib="ref")
    return t, xyz, lim, c64, q64


def _gpu64(ctx, t, xyz, h, m, c64, mp, group_size, mixed, num_shells=0, box_l=None, mdq=False):
    import torch

    n = xyz[0].size
    ft = np.float32 if mixed else np.float64
    tt = torch.float32 if mixed else torch.float64
    ng = (n + group_size - 1) // group_size
    bounds = dev(np.minimum(np.arange(ng + 1, dtype=np.int64) * group_size, n).astype(np.uint32))
    pos = [dev(a) for a in xyz]
    hm = [dev(np.asarray(a, ft)) for a in (h, m)]
    child, i2l, lay = dev(t["tree"]["childOffsets"]), dev(t["tree"]["internalToLeaf"]), dev(t["layout"])
    c, q = dev(np.ascontiguousarray(c64, np.float64)), dev(np.ascontiguousarray(mp, ft))
    p, ax, ay, az = (zeros(n, tt) for _ in range(4))
    pot = C.c_double()
    bl = R.host_f64(box_l if box_l is not None else [0.0, 0.0, 0.0])
    args = [ng, R.ptr(bounds), R.ptr(bounds) + 4, 0, *(R.ptr(a) for a in pos), R.ptr(hm[1]), R.ptr(hm[0]),
            *(R.ptr(a) for a in pos), R.ptr(hm[1]), R.ptr(hm[0]), R.ptr(child), R.ptr(i2l), R.ptr(lay), R.ptr(c),
            R.ptr(q), 1.0, num_shells, bl, R.ptr(p), R.ptr(ax), R.ptr(ay), R.ptr(az), C.byref(pot)]
    name = "ryo_traverse_mdq_f64" if mdq else "ryo_traverse_f64"
    ctx.call(name + "_f32" if mixed else name, *args)
    return pot.value, host(p), host(ax), host(ay), host(az), ctx.stats()


@pytest.mark.parametrize("name,mk", [("uniform", lambda: W.uniform_cube(20000, seed=41)),
                                     ("plummer", lambda: W.plummer(20000, seed=42)),
                                     ("box", lambda: W.uniform_box(20000, seed=43, extents=(2.0, 1.0, 0.5)))])
def test_double_traversal_equals_reference(ctx, name, mk):
    s = mk()
    t, xyz, lim, c64, q64 = _double_system(s, 7)
    h, m = t["h"].astype(np.float64), t["m"].astype(np.float64)
    rtot, rug, rax, ray, raz = O.ref_compute_gravity64(t["leaves"], t["layout"], c64, q64, *xyz, h, m, lim)
    gtot, gp, gax, gay, gaz, st = _gpu64(ctx, t, xyz, h, m, c64, q64, 16, mixed=False)
    # the same interactions summed in another order: equal to FP64 rounding (relative to the RMS acceleration
    # for the few targets whose net force nearly cancels)
    e = np.sort(rel_err(gax, gay, gaz, rax, ray, raz))
    rms = np.sqrt(np.mean(rax ** 2 + ray ** 2 + raz ** 2))
    d = np.sqrt((gax - rax) ** 2 + (gay - ray) ** 2 + (gaz - raz) ** 2) / rms
    assert e[int(0.99 * e.size)] <= 1e-11 and d.max() <= 1e-10, (e[int(0.99 * e.size)], d.max())
    assert np.allclose(gp, rug, rtol=1e-11, atol=0)
    assert abs(gtot - rtot) <= 1e-12 * abs(rtot)
    assert st["sumP2P"] > 0 and st["sumM2P"] > 0


def test_mixed_traversal_equals_reference(ctx):
    """(d,f,f,f,d): float h, m, multipoles and outputs, double coordinates, centers and interactions"""
    s = W.plummer(20000, seed=44)
    t, xyz, lim, c64, q64 = _double_system(s, 8)
    q32 = q64.astype(np.float32)
    rtot, rug, rax, ray, raz = O.ref_compute_gravity64(t["leaves"], t["layout"], c64, q32, *xyz, t["h"], t["m"], lim,
                                                       mixed=True)
    gtot, gp, gax, gay, gaz, _ = _gpu64(ctx, t, xyz, t["h"], t["m"], c64, q32, 16, mixed=True)
    assert gax.dtype == np.float32
    # both accumulate in double and round once to float: equal up to the final rounding
    e = np.sort(rel_err(gax, gay, gaz, rax, ray, raz))
    assert e[int(0.99 * e.size)] <= 2e-7, e[int(0.99 * e.size)]
    assert abs(gtot - rtot) <= 1e-11 * abs(rtot)


def test_double_traversal_periodic_images(ctx):
    s = W.uniform_cube(8192, seed=45, lo=0.0, hi=1.0, periodic=True)
    t, xyz, lim, c64, q64 = _double_system(s, 9)
    h, m = t["h"].astype(np.float64), t["m"].astype(np.float64)
    rtot, rug, rax, ray, raz = O.ref_compute_gravity64(t["leaves"], t["layout"], c64, q64, *xyz, h, m, lim,
                                                       num_shells=1)
    gtot, gp, gax, gay, gaz, _ = _gpu64(ctx, t, xyz, h, m, c64, q64, 16, mixed=False, num_shells=1,
                                        box_l=[1.0, 1.0, 1.0])
    # 27 image sums that cancel to a small residual: compare relative to the RMS acceleration
    rms = np.sqrt(np.mean(rax ** 2 + ray ** 2 + raz ** 2))
    d = np.sqrt((gax - rax) ** 2 + (gay - ray) ** 2 + (gaz - raz) ** 2) / rms
    assert d.max() <= 1e-10, d.max()
    assert abs(gtot - rtot) <= 1e-11 * abs(rtot)


def test_double_traversal_64_target_groups(ctx):
    """one wavefront per 64-target group: the reference bounds against the FP64 direct sum, and the same forces as
    the 16-target grouping to the accuracy of the method"""
    s = W.uniform_cube(16384, seed=46)
    t, xyz, lim, c64, q64 = _double_system(s, 10)
    h, m = t["h"].astype(np.float64), t["m"].astype(np.float64)
    _, _, ax, ay, az, st = _gpu64(ctx, t, xyz, h, m, c64, q64, 64, mixed=False)
    ref = O.direct_sum(*(a.astype(np.float32) for a in xyz), t["h"], t["m"])
    e = np.sort(rel_err(ax, ay, az, ref[:, 1], ref[:, 2], ref[:, 3]))
    assert e[int(0.99 * e.size)] < 2e-3 and e[-1] < 3e-2
    _, _, bx, by, bz, _ = _gpu64(ctx, t, xyz, h, m, c64, q64, 16, mixed=False)
    assert np.quantile(rel_err(ax, ay, az, bx, by, bz), 0.99) < 2e-3
    assert st["maxStack"] <= 22 * 64


@pytest.mark.parametrize("name,mk", [("plummer", lambda: W.plummer(30000, seed=47)),
                                     ("box", lambda: W.uniform_box(30000, seed=48, extents=(0.5, 2.0, 1.0)))])
def test_double_pipeline_equals_reference(ctx, name, mk):
    """the whole path on double particles through the C ABI (keys, sort, gather, leaves, linked tree, FP64 upsweep,
    traversal) against the reference's own pipeline on the same doubles (oracle/_ref for keys, sort, upsweep and
    gravity): integer arrays bit-exact, centers and multipoles to FP64 rounding, forces to FP64 rounding"""
    import torch

    from gpu_util import gpu_leaves, gpu_link, gpu_sort

    s = mk()
    rng = np.random.default_rng(5)
    lim = s.limits.astype(np.float64)
    xyz = [np.clip(a.astype(np.float64) + rng.uniform(-1e-9, 1e-9, s.n) * (lim[2 * d + 1] - lim[2 * d]),
                   lim[2 * d], lim[2 * d + 1]) for d, a in enumerate((s.x, s.y, s.z))]
    m, h = s.m.astype(np.float64), s.h.astype(np.float64)
    # the reference
    rkeys = O.ref_sfc_keys64(*xyz, lim)
    rsk, rperm = O.sort_by_key(rkeys, lib="ref")
    rx = [a[rperm] for a in (*xyz, m, h)]
    leaves, counts = O.compute_octree(rsk, 64, lib="ref")
    layout = O.layout_from_counts(counts)
    tree = O.link_octree(leaves, lib="ref")
    c64, q64 = O.upsweep(*rx[:4], leaves, layout, lim, 2.0, precision="f64", lib="ref")
    rtot, rug, rax, ray, raz = O.ref_compute_gravity64(leaves, layout, c64, q64, *rx[:3], rx[4], rx[3], lim)
    # the device path
    pos = [dev(a) for a in xyz]
    k = zeros(s.n, torch.int64)
    ctx.call("ryo_sfc_keys_f64", *(R.ptr(a) for a in pos), R.ptr(k), s.n, R.host_f64(lim), 0, 0)
    gkeys = host(k, np.uint64)
    assert np.array_equal(gkeys, rkeys)
    gsk, gperm = gpu_sort(ctx, gkeys)
    assert np.array_equal(gsk, rsk) and np.array_equal(gperm, rperm)
    src = [dev(a) for a in (*xyz, m, h)]
    dst = [zeros(s.n, torch.float64) for _ in range(5)]
    perm = dev(gperm)
    ins, outs = (C.c_void_p * 5)(*(R.ptr(a) for a in src)), (C.c_void_p * 5)(*(R.ptr(a) for a in dst))
    ctx.call("ryo_gather_f64", R.ptr(perm), s.n, 5, ins, outs)
    gx = [host(a) for a in dst]
    for a, b in zip(gx, rx):
        assert np.array_equal(a, b)
    gl, gc, glay = gpu_leaves(ctx, gsk, 64)
    assert np.array_equal(gl, leaves) and np.array_equal(glay, layout)
    gt = gpu_link(ctx, gl)
    for key in ("childOffsets", "internalToLeaf", "leafToInternal"):
        assert np.array_equal(gt[key], tree[key]), key
    L = gl.size - 1
    n = L + (L - 1) // 7
    d = {key: dev(gt[key]) for key in ("prefixes", "childOffsets", "parents", "leafToInternal", "levelRange")}
    gdx = [dev(a) for a in gx[:4]]
    lay = dev(glay)
    cen, mp, mp32 = zeros(4 * n, torch.float64), zeros(8 * n, torch.float64), zeros(8 * n, torch.float32)
    ctx.call("ryo_upsweep_f64", *(R.ptr(a) for a in gdx), R.ptr(lay), L, R.ptr(d["prefixes"]),
             R.ptr(d["childOffsets"]), R.ptr(d["parents"]), R.ptr(d["leafToInternal"]), R.ptr(d["levelRange"]),
             R.host_f64(lim), 2.0, R.ptr(cen), R.ptr(mp), R.ptr(mp32))
    gcen, gmp = host(cen).reshape(n, 4), host(mp).reshape(n, 8)
    assert np.allclose(gcen[:, :3], c64[:, :3], rtol=1e-12, atol=1e-14 * np.abs(c64[:, :3]).max())
    assert np.allclose(gcen[:, 3], c64[:, 3], rtol=1e-12, atol=0)
    assert np.allclose(gmp, q64, rtol=1e-10, atol=1e-13 * np.abs(q64).max())
    assert np.array_equal(host(mp32).reshape(n, 8), gmp.astype(np.float32))
    t = {"tree": {"childOffsets": gt["childOffsets"], "internalToLeaf": gt["internalToLeaf"]}, "layout": glay}
    gtot, gp, gax, gay, gaz, _ = _gpu64(ctx, t, gx[:3], gx[4], gx[3], gcen, gmp, 16, mixed=False)
    e = np.sort(rel_err(gax, gay, gaz, rax, ray, raz))
    assert e[int(0.99 * e.size)] <= 1e-10, e[int(0.99 * e.size)]
    assert abs(gtot - rtot) <= 1e-10 * abs(rtot)


@pytest.mark.parametrize("mixed", [False, True])
def test_double_mdq_upsweep_and_traversal(ctx, mixed):
    """CartesianMDQpole with double coordinates (signed masses, so the dipoles are nonzero): the GPU upsweep against
    the reference's MDQpole upsweep on the same double particles, then (d,d,d,d,d) / (d,f,f,f,d) traversal against
    the reference's computeGravity with the same multipoles"""
    import torch

    s = W.plummer(20000, seed=46)
    rng = np.random.default_rng(9)
    s.m = np.where(rng.random(s.n) < 1 / 3, -s.m, s.m).astype(s.m.dtype)
    t, xyz, lim, _, _ = _double_system(s, 10)
    m64 = t["m"].astype(np.float64)
    cm, qm = O.ref_upsweep_mdq(*xyz, m64, t["leaves"], t["layout"], lim, 2.0, precision="f64")
    L = t["leaves"].size - 1
    n = L + (L - 1) // 7
    tr = t["tree"]
    d = {key: dev(tr[key]) for key in ("prefixes", "childOffsets", "parents", "leafToInternal", "levelRange")}
    gdx = [dev(a) for a in (*xyz, m64)]
    lay = dev(t["layout"])
    cen, mp, mp32 = zeros(4 * n, torch.float64), zeros(12 * n, torch.float64), zeros(12 * n, torch.float32)
    ctx.call("ryo_upsweep_mdq_f64", *(R.ptr(a) for a in gdx), R.ptr(lay), L, R.ptr(d["prefixes"]),
             R.ptr(d["childOffsets"]), R.ptr(d["parents"]), R.ptr(d["leafToInternal"]), R.ptr(d["levelRange"]),
             R.host_f64(lim), 2.0, R.ptr(cen), R.ptr(mp), R.ptr(mp32))
    gcen, gmp = host(cen).reshape(n, 4), host(mp).reshape(n, 12)
    assert np.abs(qm[:, 8:11]).max() > 1e-3 * np.abs(qm[:, 1:7]).max()
    assert np.allclose(gcen[:, :3], cm[:, :3], rtol=1e-12, atol=1e-14 * np.abs(cm[:, :3]).max())
    assert np.allclose(gcen[:, 3], cm[:, 3], rtol=1e-12, atol=0)
    scale = np.abs(qm).max(axis=0) + 1e-300
    assert np.all(np.abs(gmp - qm) <= 1e-11 * scale)
    assert np.array_equal(host(mp32).reshape(n, 12), gmp.astype(np.float32))

    ft = np.float32 if mixed else np.float64
    q = qm.astype(ft)
    h, m = t["h"].astype(ft), t["m"].astype(ft)
    rtot, rug, rax, ray, raz = O.ref_compute_gravity64(t["leaves"], t["layout"], cm, q, *xyz, h, m, lim,
                                                       mixed=mixed, mdq=True)
    gtot, gp, gax, gay, gaz, _ = _gpu64(ctx, t, xyz, h, m, cm, q, 16, mixed=mixed, mdq=True)
    e = np.sort(rel_err(gax, gay, gaz, rax, ray, raz))
    rms = np.sqrt(np.mean(rax.astype(np.float64) ** 2 + ray ** 2 + raz ** 2))
    dd = np.sqrt((gax - rax).astype(np.float64) ** 2 + (gay - ray) ** 2 + (gaz - raz) ** 2) / rms
    if mixed:
        assert e[int(0.99 * e.size)] <= 2e-7, e[int(0.99 * e.size)]
    else:
        assert e[int(0.99 * e.size)] <= 1e-11 and dd.max() <= 1e-10, (e[int(0.99 * e.size)], dd.max())
    # signed masses: the total may cancel, so the bound also scales with the sum of |potential| terms
    assert abs(gtot - rtot) <= 1e-11 * abs(rtot) + 1e-12 * float(np.abs(gp.astype(np.float64)).sum())
