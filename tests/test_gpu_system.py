"""The resident-system pipeline (keys -> sort -> permute -> leaves -> octree -> FP64 upsweep -> traversal [+ Ewald])
against the oracle, the FP64 direct sum and the reference accuracy bounds."""

import numpy as np
import pytest

import oracle as O
import ryoanji_amd as R
from ryoanji_amd import workloads as W

from gpu_util import dev, host, oracle_tree, rel_err

pytestmark = pytest.mark.gpu


def _solve(ctx, s, **kw):
    sysm = R.System(ctx, s.n)
    sysm.load(*(dev(a) for a in s.arrays()), s.limits, periodic=kw.pop("periodic", False))
    prm = R.default_params(**kw)
    sysm.build(prm)
    pot = sysm.gravity(prm)
    p, ax, ay, az, order = (host(t) for t in sysm.results())
    return sysm, pot, p, ax, ay, az, order


@pytest.mark.parametrize("name,mk", [("uniform", lambda: W.uniform_cube(65536, seed=1)),
                                     ("plummer", lambda: W.plummer(65536, seed=2))])
def test_system_tree_and_forces(ctx, name, mk):
    s = mk()
    sysm, pot, p, ax, ay, az, order = _solve(ctx, s)
    t = oracle_tree(s)
    assert np.array_equal(order, t["perm"])  # the stable sort permutation
    tr = sysm.tree()
    assert np.array_equal(host(tr["keys"], np.uint64), t["sorted_keys"])
    assert np.array_equal(host(tr["leaves"], np.uint64), t["leaves"])
    assert np.array_equal(host(tr["layout"], np.uint32), t["layout"])
    for k in ("childOffsets", "internalToLeaf", "leafToInternal", "parents"):
        assert np.array_equal(host(tr[k]), t["tree"][k]), k
    # forces against the oracle traversal with the same grouping (64) on the same tree
    args = (t["x"], t["y"], t["z"], t["h"], t["m"])
    otot, oug, oax, oay, oaz, _ = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], *args, s.limits,
                                                    group_size=64)
    e = np.sort(rel_err(ax, ay, az, oax, oay, oaz))
    assert e[int(0.99 * e.size)] <= 1e-5, e[int(0.99 * e.size)]
    assert abs(pot - otot) <= 1e-5 * abs(otot)


def test_accuracy_vs_fp64_direct(ctx):
    """reference bounds at theta 0.5 (global_forces_gpu.cpp:218-220): 1st pct < 2e-3, max < 3e-2, pot < 1e-2"""
    s = W.uniform_cube(65536, seed=3)
    sysm, pot, p, ax, ay, az, order = _solve(ctx, s)
    x, y, z, m, h = (a[order] for a in s.arrays())
    tgt = np.random.default_rng(0).choice(s.n, 4096, replace=False)
    import torch

    out = torch.zeros(4096 * 4, dtype=torch.float64, device="cuda")
    ctx.call("ryo_direct_sum_f64", *(R.ptr(dev(a)) for a in (x, y, z, h, m)), s.n, R.ptr(dev(tgt.astype(np.int64))),
             4096, 1.0, None, 0, R.ptr(out))
    ref = host(out).reshape(4096, 4)
    e = np.sort(rel_err(ax[tgt], ay[tgt], az[tgt], ref[:, 1], ref[:, 2], ref[:, 3]))
    assert e[int(0.99 * e.size)] < 2e-3 and e[-1] < 3e-2
    # total potential over all particles by the FP64 direct sum on the GPU
    out2 = torch.zeros(s.n * 4, dtype=torch.float64, device="cuda")
    ctx.call("ryo_direct_sum_f64", *(R.ptr(dev(a)) for a in (x, y, z, h, m)), s.n, None, s.n, 1.0, None, 0,
             R.ptr(out2))
    dpot = 0.5 * host(out2).reshape(-1, 4)[:, 0].sum()
    assert abs(pot - dpot) / abs(dpot) < 1e-2


def test_direct_sum_matches_oracle(ctx):
    s = W.plummer(3000, seed=4)
    import torch

    tgt = np.arange(0, 3000, 7, dtype=np.int64)
    out = torch.zeros(tgt.size * 4, dtype=torch.float64, device="cuda")
    ctx.call("ryo_direct_sum_f64", *(R.ptr(dev(a)) for a in (s.x, s.y, s.z, s.h, s.m)), s.n, R.ptr(dev(tgt)),
             tgt.size, 1.0, None, 0, R.ptr(out))
    g = host(out).reshape(-1, 4)
    o = O.direct_sum(s.x, s.y, s.z, s.h, s.m, targets=tgt)
    assert np.allclose(g, o, rtol=1e-11, atol=1e-13 * np.abs(o).max())
    # periodic images
    out1 = torch.zeros(tgt.size * 4, dtype=torch.float64, device="cuda")
    bl = R.host_f64([60.0, 60.0, 60.0])
    ctx.call("ryo_direct_sum_f64", *(R.ptr(dev(a)) for a in (s.x, s.y, s.z, s.h, s.m)), s.n, R.ptr(dev(tgt)),
             tgt.size, 1.0, bl, 1, R.ptr(out1))
    o1 = O.direct_sum(s.x, s.y, s.z, s.h, s.m, targets=tgt, box_l=[60.0, 60.0, 60.0], num_shells=1)
    assert np.allclose(host(out1).reshape(-1, 4), o1, rtol=1e-11, atol=1e-13 * np.abs(o1).max())


def test_ewald_matches_oracle(ctx):
    s = W.uniform_cube(20000, seed=6, lo=0.0, hi=1.0, periodic=True)
    t = oracle_tree(s)
    c64, q64 = O.upsweep(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"], s.limits.astype(np.float64), 2.0,
                         precision="f64")
    root, center = q64[0], c64[0, :3]
    otot, oug, oax, oay, oaz = O.ewald(root, center, t["x"], t["y"], t["z"], t["m"], 1.0)
    import torch

    n = s.n
    ug, ax, ay, az = (torch.zeros(n, dtype=torch.float32, device="cuda") for _ in range(4))
    import ctypes

    tot = ctypes.c_double()
    ctx.call("ryo_ewald_f32", R.host_f64(root), R.host_f64(center), 0, n,
             *(R.ptr(dev(a)) for a in (t["x"], t["y"], t["z"], t["m"])), R.host_f32(s.limits), 1.0,
             R.host_f64(O.EWALD_DEFAULT), R.ptr(ug), R.ptr(ax), R.ptr(ay), R.ptr(az), ctypes.byref(tot))
    gax, gay, gaz = host(ax), host(ay), host(az)
    e = rel_err(gax, gay, gaz, oax, oay, oaz)
    assert np.quantile(e, 0.99) < 1e-5
    assert abs(tot.value - otot) <= 1e-9 * abs(otot)


def test_periodic_system_with_ewald(ctx):
    s = W.uniform_cube(32768, seed=7, lo=0.0, hi=1.0, periodic=True)
    sysm, pot, p, ax, ay, az, order = _solve(ctx, s, periodic=True, num_shells=1, ewald=1)
    t = oracle_tree(s)
    args = (t["x"], t["y"], t["z"], t["h"], t["m"])
    otot, oug, oax, oay, oaz, _ = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], *args, s.limits,
                                                    group_size=64, num_shells=1)
    c64, q64 = O.upsweep(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"], s.limits.astype(np.float64), 2.0,
                         precision="f64")
    etot, eug, eax, eay, eaz = O.ewald(q64[0], c64[0, :3], t["x"], t["y"], t["z"], t["m"], 1.0)
    rx, ry, rz = oax + eax, oay + eay, oaz + eaz
    e = np.sort(rel_err(ax, ay, az, rx, ry, rz))
    assert e[int(0.99 * e.size)] <= 1e-4, e[int(0.99 * e.size)]
    assert abs(pot - (otot + etot)) <= 1e-5 * abs(otot + etot)
