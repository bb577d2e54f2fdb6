"""CartesianMDQpole (monopole + dipole + quadrupole) upsweep and traversal against the reference's MDQpole
instantiation (ryoanji/nbody/cartesian_qpole.hpp:237-417), through the C ABI.

Signed masses make the dipole about the |m|-weighted center nonzero, so every dipole term of P2M, M2M and M2P
is exercised. The FP64 multipoles are compared with the reference's double instantiation (1e-11 of each
column's scale: only the summation order differs); the traversal runs on the reference's own float centers and
multipoles with 16-target groups (per-particle relative acceleration error p99 <= 1e-5).
"""

import numpy as np
import pytest

import oracle as O
from ryoanji_amd import workloads as W

import ryoanji_amd as R

from gpu_util import dev, gpu_traverse, gpu_upsweep, host, oracle_tree, rel_err

pytestmark = pytest.mark.gpu


def _signed(s, seed):
    """the system with a random sign on a third of the masses"""
    rng = np.random.default_rng(seed)
    s.m = np.where(rng.random(s.m.size) < 1 / 3, -s.m, s.m).astype(s.m.dtype)
    return s


SYSTEMS = {
    "uniform65536": lambda: _signed(W.uniform_cube(65536, seed=31), 1),
    "plummer32768": lambda: _signed(W.plummer(32768, seed=32), 2),
    "box40000": lambda: _signed(W.uniform_box(40000, seed=33, extents=(2.0, 1.0, 0.25)), 3),
}


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_upsweep_mdq(ctx, name):
    s = SYSTEMS[name]()
    t = oracle_tree(s)
    c32, q32, c64, q64 = gpu_upsweep(ctx, t["x"], t["y"], t["z"], t["m"], t["layout"], t["tree"], s.limits, 2.0,
                                     mdq=True)
    oc, oq = O.ref_upsweep_mdq(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"],
                               s.limits.astype(np.float64), 2.0, precision="f64")
    assert q64.shape == oq.shape
    # the dipole is genuinely nonzero in this system
    assert np.abs(oq[:, 8:11]).max() > 1e-3 * np.abs(oq[:, 1:7]).max()
    scale = np.abs(oq).max(axis=0) + 1e-300
    assert np.allclose(c64[:, :3], oc[:, :3], rtol=1e-12, atol=1e-12 * np.abs(oc[:, :3]).max())
    assert np.all(np.abs(q64 - oq) <= 1e-11 * scale)
    assert np.all(q64[:, 11] == 0)
    assert np.all(np.abs(q32 - oq.astype(np.float32)) <= 1e-6 * scale)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_traversal_mdq_parity(ctx, name):
    s = SYSTEMS[name]()
    t = oracle_tree(s)
    fc, fq = O.ref_upsweep_mdq(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"], s.limits, 2.0)
    args = (t["x"], t["y"], t["z"], t["h"], t["m"])
    otot, oug, oax, oay, oaz = O.ref_compute_gravity_mdq(t["leaves"], t["layout"], fc, fq, *args, s.limits)
    gtot, gp, gax, gay, gaz, _ = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"], t["layout"],
                                              fc, fq, 16, mdq=True)
    e = np.sort(rel_err(gax, gay, gaz, oax, oay, oaz))
    assert e[int(0.99 * e.size)] <= 1e-5, f"p99 {e[int(0.99 * e.size)]:.3e} max {e[-1]:.3e}"
    assert abs(gtot - otot) <= 1e-5 * abs(otot) + 1e-6 * float(np.abs(gp).sum())
    pe = np.abs(gp - oug) / np.maximum(np.abs(oug), 1e-30)
    assert np.quantile(pe, 0.99) <= 1e-4


def test_mdq_dipole_lowers_error(ctx):
    """with signed masses the MDQ far field is closer to the direct sum than the quadrupole's"""
    s = SYSTEMS["plummer32768"]()
    t = oracle_tree(s)
    c, q, _, _ = gpu_upsweep(ctx, t["x"], t["y"], t["z"], t["m"], t["layout"], t["tree"], s.limits, 2.0)
    cm, qm, _, _ = gpu_upsweep(ctx, t["x"], t["y"], t["z"], t["m"], t["layout"], t["tree"], s.limits, 2.0, mdq=True)
    sel = np.arange(0, t["x"].size, 64)
    ref = O.direct_sum(t["x"], t["y"], t["z"], t["h"], t["m"], targets=sel)
    errs = []
    for cc, qq, mdq in ((c, q, False), (cm, qm, True)):
        _, _, ax, ay, az, _ = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"], t["layout"], cc,
                                           qq, 16, mdq=mdq)
        errs.append(np.median(rel_err(ax[sel], ay[sel], az[sel], ref[:, 1], ref[:, 2], ref[:, 3])))
    assert errs[1] < 0.5 * errs[0], errs


@pytest.mark.parametrize("group_size,num_shells", [(64, 0), (16, 1), (64, 1)])
def test_traversal_mdq_groups_and_images(ctx, group_size, num_shells):
    """64-target groups (the GPU demo's), and one shell of periodic images in a unit periodic box, against the
    reference's MDQpole computeGravity on the same grouping and images"""
    if num_shells:
        s = _signed(W.uniform_cube(32768, seed=34, lo=0.0, hi=1.0, periodic=True), 4)
    else:
        s = SYSTEMS["plummer32768"]()
    t = oracle_tree(s)
    fc, fq = O.ref_upsweep_mdq(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"], s.limits, 2.0)
    args = (t["x"], t["y"], t["z"], t["h"], t["m"])
    if group_size == 16:
        otot, oug, oax, oay, oaz = O.ref_compute_gravity_mdq(t["leaves"], t["layout"], fc, fq, *args, s.limits,
                                                             num_shells=num_shells)
    else:
        otot, oug, oax, oay, oaz = O.ref_compute_gravity_mdq(t["leaves"], t["layout"], fc, fq, *args, s.limits,
                                                             num_shells=num_shells, group_size=group_size)
    box = [1.0, 1.0, 1.0] if num_shells else None
    gtot, gp, gax, gay, gaz, _ = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"], t["layout"],
                                              fc, fq, group_size, num_shells=num_shells, box_l=box, mdq=True)
    if num_shells:
        # 27 float image sums of forces that largely cancel: relative to the RMS acceleration
        rms = np.sqrt(np.mean(oax ** 2 + oay ** 2 + oaz ** 2))
        d = np.sqrt((gax - oax) ** 2 + (gay - oay) ** 2 + (gaz - oaz) ** 2) / rms
        assert np.quantile(d, 0.99) <= 2e-5, np.quantile(d, 0.99)
    else:
        e = np.sort(rel_err(gax, gay, gaz, oax, oay, oaz))
        assert e[int(0.99 * e.size)] <= 1e-5, e[int(0.99 * e.size)]


def test_traversal_mdq_overflow_retry(ctx):
    """a LIFO too shallow for most groups: overflowing groups are split and traversed again (splitFailedGroups),
    and every result stays within the float tolerance of the full-depth MDQ traversal"""
    s = SYSTEMS["plummer32768"]()
    t = oracle_tree(s)
    fc, fq = O.ref_upsweep_mdq(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"], s.limits, 2.0)
    full = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"], t["layout"], fc, fq, 64, mdq=True)
    assert R.lib().ryo_traversal_retries(ctx.h) == 0
    cap = max(9, int(full[5]["maxStack"]) - 1)  # one entry short of the deepest walk
    small = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"], t["layout"], fc, fq, 64, mdq=True,
                         stack_cap=cap)
    assert R.lib().ryo_traversal_retries(ctx.h) > 0
    # split groups have smaller bounding boxes, so some cells are accepted that the full group opened: the
    # difference is the Barnes-Hut error level, not float noise
    e = np.sort(rel_err(small[2], small[3], small[4], full[2], full[3], full[4]))
    assert e[int(0.99 * e.size)] <= 2e-3, e[int(0.99 * e.size)]


def test_resident_system_mdq(ctx):
    """ryo_system with multipole = 1 (MultipoleHolder's CartesianMDQpole instantiation): tree, FP64 MDQ upsweep
    and MDQ traversal against the reference's MDQpole upsweep and computeGravity; Ewald is refused for MDQpole
    as the reference has no such instantiation"""
    s = SYSTEMS["uniform65536"]()
    t = oracle_tree(s)
    sysm = R.System(ctx, s.n)
    sysm.load(*(dev(a) for a in s.arrays()), s.limits)
    prm = R.default_params(group_size=16, multipole=1)
    sysm.build(prm)
    pot = sysm.gravity(prm)
    p, ax, ay, az, order = (host(a) for a in sysm.results())
    assert np.array_equal(order, t["perm"])
    tr = sysm.tree()
    q64 = host(tr["multipoles64"]).reshape(-1, 12)
    oc, oq = O.ref_upsweep_mdq(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"],
                               s.limits.astype(np.float64), 2.0, precision="f64")
    scale = np.abs(oq).max(axis=0) + 1e-300
    assert np.all(np.abs(q64 - oq) <= 1e-11 * scale)
    fc, fq = O.ref_upsweep_mdq(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"], s.limits, 2.0)
    otot, oug, oax, oay, oaz = O.ref_compute_gravity_mdq(t["leaves"], t["layout"], fc, fq, t["x"], t["y"], t["z"],
                                                         t["h"], t["m"], s.limits)
    e = np.sort(rel_err(ax, ay, az, oax, oay, oaz))
    assert e[int(0.99 * e.size)] <= 1e-5, e[int(0.99 * e.size)]
    assert abs(pot - otot) <= 1e-5 * abs(otot) + 1e-6 * float(np.abs(p).sum())
    sysm.close()
    per = W.uniform_cube(4096, seed=35, lo=0.0, hi=1.0, periodic=True)
    sp = R.System(ctx, per.n)
    sp.load(*(dev(a) for a in per.arrays()), per.limits, periodic=True)
    pp = R.default_params(multipole=1, num_shells=1, ewald=1)
    sp.build(pp)
    with pytest.raises(R.RyoError) as err:
        sp.gravity(pp)
    assert err.value.code == R.RYO_ERR_UNSUPPORTED
    sp.close()
