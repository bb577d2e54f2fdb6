"""The oracle's FP64 Ewald pair sum (the periodic accuracy reference) pinned by closed forms: the Madelung
constant of the simple cubic lattice, independence of the splitting parameter alpha, and the traversal + Ewald
correction in the one configuration where the root multipole is exact (a single particle)."""

import numpy as np
import pytest

import oracle as O
from ryoanji_amd import workloads as W

# potential at a lattice site of the simple cubic lattice (unit spacing, unit mass) due to all other sites and
# a neutralizing background, for the potential -1/r: + 2.837297479480... (the Madelung constant of the
# simple cubic jellium)
MADELUNG_SC = 2.8372974794806


@pytest.mark.parametrize("L,lo", [(1.0, 0.0), (3.0, -1.0)])
def test_single_particle_is_the_madelung_constant(L, lo):
    x = np.array([lo + 0.3 * L], np.float32)
    y = np.array([lo + 0.71 * L], np.float32)
    z = np.array([lo + 0.05 * L], np.float32)
    m = np.array([2.0], np.float32)
    h = np.array([1e-3], np.float32)
    for a in (5.5, 6.5, 9.0):
        out = O.ewald_pair_sum(x, y, z, h, m, L, alpha_scale=a, G=0.5)
        assert abs(out[0, 0] - 0.5 * 2.0 * 2.0 * MADELUNG_SC / L) < 1e-11, (a, out[0, 0])
        assert np.abs(out[0, 1:]).max() < 1e-12


def test_sub_lattice_of_eight():
    """eight equal masses on the sites of a lattice of spacing L/2: each sees that lattice, potential
    m * Madelung / (L/2), no force"""
    L, m0 = 2.0, 0.25
    g = np.array([(i, j, k) for i in (0, 1) for j in (0, 1) for k in (0, 1)], np.float64) * (L / 2) + 0.125  # exact in float32
    x, y, z = (g[:, d].astype(np.float32) for d in range(3))
    m = np.full(8, m0, np.float32)
    h = np.full(8, 1e-3, np.float32)
    out = O.ewald_pair_sum(x, y, z, h, m, L)
    assert np.allclose(out[:, 0], m0 * m0 * MADELUNG_SC / (L / 2), rtol=1e-11, atol=0)
    assert np.abs(out[:, 1:]).max() < 1e-11


def _system_with_close_pairs(n, seed):
    s = W.uniform_cube(n, seed=seed, lo=0.0, hi=1.0)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    # a few pairs well inside the softening length (h_i + h_j = 2e-3), one exact duplicate
    for k, d in enumerate((3e-4, 1e-3, 1.7e-3)):
        x[2 * k + 1], y[2 * k + 1], z[2 * k + 1] = x[2 * k] + d, y[2 * k], z[2 * k] - 0.5 * d
    x[9], y[9], z[9] = x[8], y[8], z[8]
    return x, y, z, s.h, s.m


def test_independent_of_alpha():
    x, y, z, h, m = _system_with_close_pairs(300, 5)
    targets = np.arange(0, 300, 3)
    ref = O.ewald_pair_sum(x, y, z, h, m, 1.0, targets=targets, alpha_scale=6.5)
    rms = np.sqrt(np.mean(np.sum(ref[:, 1:] ** 2, axis=1)))
    for a in (5.2, 8.0, 11.0):
        out = O.ewald_pair_sum(x, y, z, h, m, 1.0, targets=targets, alpha_scale=a)
        assert np.abs(out[:, 1:] - ref[:, 1:]).max() < 1e-10 * rms, a
        assert np.abs(out[:, 0] - ref[:, 0]).max() < 1e-10 * np.abs(ref[:, 0]).max(), a


def test_softened_pairs_follow_the_traversal_law():
    """moving a close pair's partner changes the pair sum by exactly the softened-law difference of the two
    positions (the rest of the Ewald sum is smooth and nearly unchanged for a tiny move)"""
    x, y, z, h, m = _system_with_close_pairs(200, 8)
    t = np.array([0])
    a = O.ewald_pair_sum(x, y, z, h, m, 1.0, targets=t)
    x2 = x.copy()
    x2[1] = x[0] + np.float32(2e-4)  # partner from 3e-4 to 2e-4 along x (still inside the softening)
    b = O.ewald_pair_sum(x2, y, z, h, m, 1.0, targets=t)
    hij = float(h[0]) + float(h[1])

    def soft(px):
        r = np.array([float(px) - float(x[0]), float(y[1]) - float(y[0]), float(z[1]) - float(z[0])])
        R2 = r @ r
        return -float(m[1]) * R2 / hij ** 3, float(m[1]) * r / hij ** 3

    (pa, fa), (pb, fb) = soft(x[1]), soft(x2[1])
    # the smooth remainder changes by ~ m a^3 |dr| ~ 1e-7 of the softened force change
    assert abs((b[0, 0] - a[0, 0]) / float(m[0]) - (pb - pa)) < 1e-6 * abs(pb - pa)
    assert np.abs((b[0, 1:] - a[0, 1:]) - (fb - fa)).max() < 1e-3 * np.abs(fb - fa).max()


def _pipeline(x, y, z, h, m, L, consistent=False):
    """the reference's periodic path in the oracle: traversal over 1 replica shell (opening every cell: one leaf)
    plus the Ewald correction from the root multipole (default settings: alpha 2/L, hCut 2.8); consistent=True
    takes the k-space potential consistent with the k-space acceleration"""
    lim = np.array([-L / 2, L / 2] * 3, np.float32)
    leaves = np.array([0, 1 << 63], np.uint64)
    lay = np.array([0, x.size], np.uint32)
    c, q = O.upsweep(x, y, z, m, leaves, lay, lim, 2.0)
    _, ug, ax, ay, az = O.compute_gravity_f64(leaves, lay, c, q, x, y, z, h, m, lim, num_shells=1)
    c64, q64 = O.upsweep(x, y, z, m, leaves, lay, lim.astype(np.float64), 2.0, precision="f64")
    _, eug, eax, eay, eaz = O.ewald(q64[0], c64[0, :3], x, y, z, m, L, consistent=consistent)
    return ug + eug, np.stack([ax + eax, ay + eay, az + eaz], axis=1)


@pytest.mark.parametrize("L", [1.0, 2.0])
def test_single_particle_traversal_plus_ewald(L):
    """one particle: the root multipole is the particle itself, so the traversal over the replica images plus the
    Ewald correction is the exact periodic sum up to the correction's cutoffs, except for one term: the
    reference's k-space potential (ewald.hpp:315-320 with hfac_cos = -g0 M, :111 and :180) has the opposite sign
    of its k-space acceleration, so the path's potential exceeds the exact one by twice the k-space sum
    2 M sum_{0 < |h| <= 2.8} exp(-pi^2 h^2 / (alpha L)^2) / (pi h^2 L) (alpha L = 2); its accelerations are exact"""
    x = np.array([0.1 * L], np.float32)
    y = np.array([-0.2 * L], np.float32)
    z = np.array([0.05 * L], np.float32)
    m = np.array([1.0], np.float32)
    h = np.array([1e-3], np.float32)
    pot, acc = _pipeline(x, y, z, h, m, L)
    out = O.ewald_pair_sum(x, y, z, h, m, L)
    hv = np.array([(a, b, c) for a in range(-3, 4) for b in range(-3, 4) for c in range(-3, 4)], np.float64)
    h2 = (hv ** 2).sum(1)
    h2 = h2[(h2 > 0) & (h2 <= 2.8 ** 2)]
    kspace = np.sum(np.exp(-np.pi ** 2 * h2 / 4.0) / (np.pi * h2 * L))
    assert abs(pot[0] - out[0, 0] - 2 * kspace) < 1e-9 * abs(out[0, 0])
    assert np.abs(acc).max() < 1e-9 and np.abs(out[0, 1:]).max() < 1e-12


def test_close_pair_accelerations_equal_traversal_plus_ewald():
    """two particles 6e-3 L apart: the root quadrupole about their center represents their far images to
    (d/L)^4, so the path's accelerations equal the exact periodic ones to ~1e-9"""
    L = 1.0
    d = np.array([0.003, 0.0018, -0.0012])
    x, y, z = (np.array([0.02 + d[k], 0.02 - d[k]], np.float32) for k in range(3))
    m = np.array([0.5, 0.5], np.float32)
    h = np.array([1e-4, 1e-4], np.float32)
    _, acc = _pipeline(x, y, z, h, m, L)
    out = O.ewald_pair_sum(x, y, z, h, m, L)
    scale = np.abs(out[:, 1:]).max()
    assert np.abs(acc - out[:, 1:]).max() < 1e-9 * scale
    # and the periodic part matters at that level: without the images the accelerations differ by far more
    iso = O.direct_sum(x, y, z, h, m)
    assert np.abs(iso[:, 1:] - out[:, 1:]).max() > 1e-6 * scale


@pytest.mark.parametrize("L", [1.0, 2.0])
def test_consistent_potential_is_exact_for_one_particle(L):
    """with the k-space potential consistent with the k-space acceleration, the path's potential of one particle is
    the exact periodic potential (the Madelung constant) and its acceleration is unchanged"""
    x = np.array([0.1 * L], np.float32)
    y = np.array([-0.2 * L], np.float32)
    z = np.array([0.05 * L], np.float32)
    m = np.array([1.0], np.float32)
    h = np.array([1e-3], np.float32)
    pot, acc = _pipeline(x, y, z, h, m, L, consistent=True)
    pot_ref, acc_ref = _pipeline(x, y, z, h, m, L)
    out = O.ewald_pair_sum(x, y, z, h, m, L)
    assert abs(pot[0] - out[0, 0]) < 1e-9 * abs(out[0, 0])
    assert abs(pot[0] - 2.8372974794806 / L) < 5e-9  # the correction's k-space cut (hCut 2.8) leaves ~3e-9
    assert np.array_equal(acc, acc_ref)


def test_consistent_potential_of_a_close_pair():
    """two particles: the consistent potential matches the exact sum to the root quadrupole's accuracy, the
    reference's does not; accelerations are the same in both"""
    d = np.array([0.003, 0.0018, -0.0012])
    x, y, z = (np.array([0.02 + d[k], 0.02 - d[k]], np.float32) for k in range(3))
    m = np.array([0.5, 0.5], np.float32)
    h = np.array([1e-4, 1e-4], np.float32)
    pc, ac = _pipeline(x, y, z, h, m, 1.0, consistent=True)
    pr, ar = _pipeline(x, y, z, h, m, 1.0)
    out = O.ewald_pair_sum(x, y, z, h, m, 1.0)
    assert np.array_equal(ac, ar)
    assert np.abs(pc - out[:, 0]).max() < 1e-8 * np.abs(out[:, 0]).max()
    assert np.abs(pr - out[:, 0]).max() > 1e-3 * np.abs(out[:, 0]).max()
