"""GPU parity of every stage of the path against the oracle, through the C ABI.

Integer stages (keys, sort permutation, leaves/counts/layout, linked-octree arrays) must be bit-exact.
The FP64 upsweep is compared with the oracle's FP64 instantiation (1e-12 relative) and with the float
reference pipeline (1e-5). Traversal parity uses identical float inputs and the reference grouping of 16
targets: per-particle relative acceleration error p99 <= 1e-5 (the path's stated tolerance).
"""

import numpy as np
import pytest

import oracle as O
import ryoanji_amd as R
from ryoanji_amd import workloads as W

from gpu_util import (dev, gpu_keys, gpu_leaves, gpu_link, gpu_sort, gpu_traverse, gpu_upsweep, host, oracle_tree,
                      rel_err, zeros)

pytestmark = pytest.mark.gpu

SYSTEMS = {
    "uniform65536": lambda: W.uniform_cube(65536, seed=1),
    "plummer32768": lambda: W.plummer(32768, seed=2),
    "hernquist40000": lambda: W.hernquist_mix(40000, seed=3),
    # non-cubic boxes: mixed-dimension Hilbert keys (bits 21/20/18 and 21/21/20)
    "box40000": lambda: W.uniform_box(40000, seed=24, extents=(2.0, 1.0, 0.25)),
    "disk32768": lambda: W.plummer_disk(32768, seed=25, flattening=0.4),
}


def _dups(n=5000, seed=4):
    """many particles on a few identical positions: level-21 leaves with more than `bucket` particles"""
    s = W.uniform_cube(n, seed=seed)
    k = np.arange(n) % 37
    for a in (s.x, s.y, s.z):
        a[:] = a[k]
    return s


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_keys_bit_exact(ctx, name):
    s = SYSTEMS[name]()
    for kind, hil in ((0, True), (1, False)):
        g = gpu_keys(ctx, s, kind=kind)
        o = O.sfc_keys(s.x, s.y, s.z, s.limits, hilbert=hil)
        assert np.array_equal(g, o), f"{name} kind {kind}: {np.count_nonzero(g != o)} keys differ"


def test_keys_remove_key_untouched(ctx):
    s = W.uniform_cube(10000, seed=5)
    pre = np.zeros(s.n, np.uint64)
    pre[::7] = np.uint64(1 << 63)
    g = gpu_keys(ctx, s, keys=pre, skip_removed=1)
    o = O.sfc_keys(s.x, s.y, s.z, s.limits, keys=pre)
    assert np.array_equal(g, o)
    assert np.all(g[::7] == np.uint64(1 << 63))


# up to 8192 pairs the sort is one block running every pass; beyond, most-significant-digit rounds and LDS sorts
@pytest.mark.parametrize("n", [1, 2, 100, 4095, 4096, 4097, 8191, 8192, 8193, 32767, 32768, 32769, 70001, 1 << 20,
                               3_000_017])
def test_sort_stable_bit_exact(ctx, n):
    rng = np.random.default_rng(n)
    # few distinct values in the high bytes and ties: stability matters
    keys = (rng.integers(0, 1 << 20, n, dtype=np.uint64) << np.uint64(40)) | rng.integers(0, 4, n, dtype=np.uint64)
    vals = rng.permutation(n).astype(np.int32)
    gk, gv = gpu_sort(ctx, keys, vals)
    ok, ov = O.sort_by_key(keys, vals)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv, ov)


@pytest.mark.parametrize("n", [200000, 30000])
def test_sort_full_64bit(ctx, n):
    rng = np.random.default_rng(9)
    keys = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    keys[::3] = keys[1::3][: keys[::3].size]
    gk, gv = gpu_sort(ctx, keys)
    ok, ov = O.sort_by_key(keys)
    assert np.array_equal(gk, ok) and np.array_equal(gv, ov)


def _clustered_keys(kind, n, rng):
    """key sets that drive the large-sort path through its rarer branches"""
    u64 = np.uint64
    if kind == "equal":  # one key: every round keeps a single large segment, nothing is left for the LDS sorts
        return np.full(n, u64(0x1234_5678_9ABC_DEF0), dtype=np.uint64)
    if kind == "deep":  # top 40 bits shared, a few hundred distinct low values: rounds 1..5 split nothing
        return u64(0x5A5A_5A5A_5A00_0000) | rng.integers(0, 300, n, dtype=np.uint64)
    if kind == "mixed_depth":  # a dense cluster beside sparse keys: large segments at several depths at once
        k = rng.integers(0, np.iinfo(np.uint64).max >> np.uint64(1), n, dtype=np.uint64)
        m = rng.random(n) < 0.6
        k[m] = u64(0x0123_4567_0000_0000) | rng.integers(0, 1 << 22, int(m.sum()), dtype=np.uint64)
        m2 = rng.random(n) < 0.1
        k[m2] = u64(0x0123_4567_89AB_CD00) | rng.integers(0, 16, int(m2.sum()), dtype=np.uint64)
        return k
    if kind == "removed":  # removal sentinels (2^63) among 63-bit keys
        k = rng.integers(0, np.iinfo(np.uint64).max >> np.uint64(1), n, dtype=np.uint64)
        k[rng.random(n) < 0.3] = u64(1 << 63)
        return k
    if kind == "small_buckets":  # many top-digit buckets of a few hundred keys: packed ranges of several buckets
        return (rng.integers(0, 1 << 16, n, dtype=np.uint64) << np.uint64(48)) | rng.integers(0, 1 << 12, n,
                                                                                                dtype=np.uint64)
    if kind == "sizes":  # second-level buckets of every LDS size class (< 512, < 2048, < 4096, < 8192 keys)
        top = rng.choice([1, 2, 3, 4], n, p=[0.1, 0.2, 0.3, 0.4]).astype(np.uint64)
        return (top << np.uint64(56)) | (rng.integers(0, 64, n, dtype=np.uint64) << np.uint64(48)) | \
            rng.integers(0, 1 << 30, n, dtype=np.uint64)
    if kind.startswith("clusters"):  # clusters of keys equal in their top 40 bits beside random keys: runs longer
        # than the insertion limit after a range's top digits (`clusters_<count>_<size>`)
        _, count, size = kind.split("_")
        count, size = int(count), int(size)
        k = rng.integers(0, np.iinfo(np.uint64).max >> np.uint64(1), n, dtype=np.uint64)
        tops = rng.integers(0, 1 << 39, count, dtype=np.uint64) << np.uint64(24)
        at = rng.permutation(n)[: count * size].reshape(count, size)
        for c in range(count):
            k[at[c]] = tops[c] | rng.integers(0, 1 << 24, size, dtype=np.uint64)
        return k
    raise ValueError(kind)


@pytest.mark.parametrize("kind,n", [("equal", 300000), ("deep", 250000), ("mixed_depth", 1 << 20),
                                    ("removed", 200001), ("small_buckets", 1 << 21), ("sizes", 400000),
                                    ("deep", 8193), ("clusters_3_1000", 6000), ("clusters_8_500", 8000),
                                    ("clusters_9_400", 8000), ("clusters_20_5000", 200000),
                                    ("clusters_4_3000", 1 << 20)])
def test_sort_clustered_bit_exact(ctx, kind, n):
    """stable order and permutation on key sets that need deep rounds, pack buckets or keep equal-key runs"""
    rng = np.random.default_rng(n + len(kind))
    keys = _clustered_keys(kind, n, rng)
    vals = rng.permutation(n).astype(np.int32)
    gk, gv = gpu_sort(ctx, keys, vals)
    ok, ov = O.sort_by_key(keys, vals)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv, ov)


def test_sort_sfc_keys_large(ctx):
    """Hilbert keys of a clustered (Plummer) system of 2^22 particles, bit-exact keys and permutation"""
    s = W.plummer(1 << 22, seed=5)
    keys = O.sfc_keys(s.x, s.y, s.z, s.limits)
    gk, gv = gpu_sort(ctx, keys)
    ok, ov = O.sort_by_key(keys)
    assert np.array_equal(gk, ok) and np.array_equal(gv, ov)


CASES = list(SYSTEMS) + ["dups", "n1", "n64", "n65", "n1000"]


def _case(name):
    if name in SYSTEMS:
        return SYSTEMS[name]()
    if name == "dups":
        return _dups()
    return W.uniform_cube(int(name[1:]), seed=11)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("bucket", [64, 16])
def test_leaves_and_layout_bit_exact(ctx, name, bucket):
    s = _case(name)
    sk, _ = O.sort_by_key(O.sfc_keys(s.x, s.y, s.z, s.limits))
    leaves, counts = O.compute_octree(sk, bucket)
    layout = O.layout_from_counts(counts)
    gl, gc, gy = gpu_leaves(ctx, sk, bucket)
    assert gl.size == leaves.size, f"{gl.size - 1} leaves vs {leaves.size - 1}"
    assert np.array_equal(gl, leaves)
    assert np.array_equal(gc, counts)
    assert np.array_equal(gy, layout)


@pytest.mark.parametrize("name", CASES)
def test_linked_octree_bit_exact(ctx, name):
    s = _case(name)
    sk, _ = O.sort_by_key(O.sfc_keys(s.x, s.y, s.z, s.limits))
    leaves, _ = O.compute_octree(sk, 64)
    o = O.link_octree(leaves)
    g = gpu_link(ctx, leaves)
    for k in ("prefixes", "childOffsets", "parents", "levelRange", "internalToLeaf", "leafToInternal"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(g["levelRangeHost"], o["levelRange"])


@pytest.mark.parametrize("name", list(SYSTEMS) + ["dups"])
def test_upsweep_fp64(ctx, name):
    s = _case(name)
    t = oracle_tree(s)
    c32, q32, c64, q64 = gpu_upsweep(ctx, t["x"], t["y"], t["z"], t["m"], t["layout"], t["tree"], s.limits, 2.0)
    oc, oq = O.upsweep(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"], s.limits.astype(np.float64), 2.0,
                       precision="f64")
    # FP64 against the FP64 instantiation: only the summation order differs
    scale = np.abs(oq).max(axis=0) + 1e-300
    assert np.allclose(c64[:, :3], oc[:, :3], rtol=1e-12, atol=1e-12 * np.abs(oc[:, :3]).max())
    assert np.array_equal(c64[:, 3] > 0, oc[:, 3] > 0)
    assert np.all(np.abs(q64 - oq) <= 1e-11 * scale)
    # the FP32 copies are the FP64 values rounded once (mac^2 in column 3 of the float centers)
    assert np.allclose(c32, oc.astype(np.float32), rtol=1e-6, atol=1e-7 * np.abs(oc[:, :3]).max())
    assert np.all(np.abs(q32 - oq.astype(np.float32)) <= 1e-6 * scale)
    # against the reference float pipeline: centers and quadrupoles agree to float accumulation error; its MAC
    # radius carries the float error of the cell geometry and center, an absolute error of a few float ulps of
    # the box (large relative to the radius of deep cells), so the radii are compared with that absolute term
    fc, fq = t["centers"], t["mp"]
    assert np.allclose(c32[:, :3], fc[:, :3], rtol=0, atol=5e-6 * np.abs(fc[:, :3]).max())
    fscale = np.abs(fq).max(axis=0) + 1e-30
    assert np.all(np.abs(q32 - fq) <= 1e-4 * fscale)
    box = float(max(s.limits[1] - s.limits[0], s.limits[3] - s.limits[2], s.limits[5] - s.limits[4]))
    r_gpu, r_ref = np.sqrt(c32[:, 3].astype(np.float64)), np.sqrt(fc[:, 3].astype(np.float64))
    assert np.all(np.abs(r_gpu - r_ref) <= 1e-4 * r_ref + 4e-6 * box)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_traversal_parity_16_targets(ctx, name):
    """same tree, same float centers/multipoles, reference grouping of 16: <= 1e-5 at p99"""
    s = SYSTEMS[name]()
    t = oracle_tree(s)
    args = (t["x"], t["y"], t["z"], t["h"], t["m"])
    otot, oug, oax, oay, oaz, ost = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], *args, s.limits,
                                                      group_size=16)
    gtot, gp, gax, gay, gaz, gst = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"], t["layout"],
                                                t["centers"], t["mp"], 16)
    e = np.sort(rel_err(gax, gay, gaz, oax, oay, oaz))
    assert e[int(0.99 * e.size)] <= 1e-5, f"p99 {e[int(0.99 * e.size)]:.3e} max {e[-1]:.3e}"
    assert abs(gtot - otot) <= 1e-5 * abs(otot)
    pe = np.abs(gp - oug) / np.abs(oug)
    assert np.quantile(pe, 0.99) <= 1e-5
    # interaction counts: the oracle also counts M2P with empty cells; P2P counts agree exactly
    assert gst["sumP2P"] == ost[0]
    assert gst["sumM2P"] <= ost[1]


@pytest.mark.parametrize("name", ["uniform65536", "plummer32768"])
def test_traversal_parity_64_targets(ctx, name):
    """the production grouping (one wave = 64 targets) against the oracle run with the same grouping"""
    s = SYSTEMS[name]()
    t = oracle_tree(s)
    args = (t["x"], t["y"], t["z"], t["h"], t["m"])
    otot, _, oax, oay, oaz, ost = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], *args, s.limits,
                                                    group_size=64)
    for init in (0, 1):
        gtot, _, gax, gay, gaz, gst = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"],
                                                   t["layout"], t["centers"], t["mp"], 64, init_node=init)
        e = np.sort(rel_err(gax, gay, gaz, oax, oay, oaz))
        assert e[int(0.99 * e.size)] <= 1e-5, (init, e[int(0.99 * e.size)])
        assert gst["sumP2P"] == ost[0]


def test_traversal_split_retry(ctx):
    """traversal_overflow_gpu.cpp:99-146: a stack capacity one below the largest stack a group needs forces the
    overflow path; failed groups are split along the SFC and traversed again, every target is written once,
    and the result keeps the Barnes-Hut accuracy (split groups have smaller boxes, hence other interactions)"""
    import ryoanji_amd as R

    s = W.plummer(20000, seed=8)
    t = oracle_tree(s)
    a = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"], t["layout"], t["centers"], t["mp"], 64)
    assert R.lib().ryo_traversal_retries(ctx.h) == 0
    cap = max(9, int(a[5]["maxStack"]) - 1)
    b = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"], t["layout"], t["centers"], t["mp"], 64,
                     stack_cap=cap)
    assert R.lib().ryo_traversal_retries(ctx.h) > 0
    ref = O.direct_sum(t["x"], t["y"], t["z"], t["h"], t["m"])
    for r in (a, b):
        e = np.sort(rel_err(r[2], r[3], r[4], ref[:, 1], ref[:, 2], ref[:, 3]))
        assert e[int(0.99 * e.size)] < 1e-2 and np.isfinite(e).all()
    assert abs(a[0] - b[0]) <= 1e-3 * abs(a[0])


def test_single_leaf_and_tiny(ctx):
    for n in (1, 2, 30, 64):
        s = W.uniform_cube(n, seed=n)
        t = oracle_tree(s)
        args = (t["x"], t["y"], t["z"], t["h"], t["m"])
        _, _, oax, oay, oaz, _ = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], *args, s.limits,
                                                   group_size=16)
        _, _, gax, gay, gaz, _ = gpu_traverse(ctx, t["x"], t["y"], t["z"], t["m"], t["h"], t["tree"], t["layout"],
                                              t["centers"], t["mp"], 16, init_node=1)
        assert np.allclose(gax, oax, rtol=1e-5, atol=1e-6 * np.abs(oax).max() + 1e-30)


def _faces(limits):
    """particles on the lower and upper faces, one ulp inside them and at the middle, along every axis (125)"""
    lim = np.asarray(limits, np.float32)
    vals = []
    for d in range(3):
        lo, hi = lim[2 * d], lim[2 * d + 1]
        vals.append(np.array([lo, np.nextafter(lo, hi), np.float32(0.5) * (lo + hi), np.nextafter(hi, lo), hi],
                             np.float32))
    g = np.array(np.meshgrid(*vals, indexing="ij")).reshape(3, -1)
    n = g.shape[1]
    return W.System(g[0], g[1], g[2], np.full(n, 1.0 / n, np.float32), np.full(n, 1e-3, np.float32), lim)


@pytest.mark.parametrize("limits", [[-1, 1, -1, 1, -1, 1], [0, 2, -0.5, 0.5, 0, 0.25], [-30, 30, -30, 30, -30, 30]])
def test_keys_on_box_faces_bit_exact(ctx, limits):
    """cell indices at the box faces (the upper face clamps into the last cell): keys equal the oracle's"""
    s = _faces(limits)
    for kind, hil in ((0, True), (1, False)):
        assert np.array_equal(gpu_keys(ctx, s, kind=kind), O.sfc_keys(s.x, s.y, s.z, s.limits, hilbert=hil)), kind


def test_empty_inputs(ctx):
    """no particles: keys and sort do nothing, the leaves are the root alone with no particles (the reference's
    computeOctree of an empty key range), a system refuses to build"""
    import ryoanji_amd as R

    e = np.zeros(0, np.float32)
    s = W.System(e, e, e, e, e, [-1, 1, -1, 1, -1, 1])
    assert gpu_keys(ctx, s).size == 0
    k, v = gpu_sort(ctx, np.zeros(0, np.uint64))
    assert k.size == 0 and v.size == 0
    leaves, counts, layout = gpu_leaves(ctx, np.zeros(0, np.uint64), 64)
    ol, oc = O.compute_octree(np.zeros(0, np.uint64), 64)
    assert np.array_equal(leaves, ol) and np.array_equal(counts, oc) and layout.tolist() == [0, 0]
    assert leaves.tolist() == [0, 1 << 63]
    sysm = R.System(ctx, 16)
    sysm.load(*(dev(a) for a in s.arrays()), s.limits)
    with pytest.raises(R.RyoError):
        sysm.build(R.default_params())
    sysm.close()


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_keys_f64_bit_exact(ctx, name):
    """keys of double coordinates (Box<double>) equal the reference's computeSfcKeys<double> (oracle/_ref), for
    coordinates that need double precision and for points exactly on the box faces"""
    import torch

    s = SYSTEMS[name]()
    rng = np.random.default_rng(11)
    lim = s.limits.astype(np.float64)
    xyz = []
    for d, a in enumerate((s.x, s.y, s.z)):
        v = a.astype(np.float64) + rng.uniform(-1e-9, 1e-9, s.n) * (lim[2 * d + 1] - lim[2 * d])
        v[:3] = (lim[2 * d], lim[2 * d + 1], 0.5 * (lim[2 * d] + lim[2 * d + 1]))
        xyz.append(np.clip(v, lim[2 * d], lim[2 * d + 1]))
    for kind, hil in ((0, True), (1, False)):
        pos = [dev(a) for a in xyz]
        k = zeros(s.n, torch.int64)
        ctx.call("ryo_sfc_keys_f64", *(R.ptr(a) for a in pos), R.ptr(k), s.n, R.host_f64(lim), kind, 0)
        g = host(k, np.uint64)
        o = O.ref_sfc_keys64(*xyz, lim, hilbert=hil)
        assert np.array_equal(g, o), f"{name} kind {kind}: {np.count_nonzero(g != o)} keys differ"


@pytest.mark.parametrize("name", list(SYSTEMS))
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_keys_u32_bit_exact(ctx, name, precision):
    """32-bit keys (10 levels) of float or double coordinates equal the reference's computeSfcKeys with
    HilbertKey<uint32_t> / MortonKey<uint32_t> (oracle/_ref), box faces included; keys equal to removeKey (2^30)
    are left untouched with skip_removed"""
    import torch

    s = SYSTEMS[name]()
    dt = np.float32 if precision == "f32" else np.float64
    lim = s.limits.astype(dt)
    xyz = []
    for d, a in enumerate((s.x, s.y, s.z)):
        v = a.astype(dt).copy()
        v[:3] = (lim[2 * d], lim[2 * d + 1], dt(0.5) * (lim[2 * d] + lim[2 * d + 1]))
        xyz.append(v)
    host_lim = R.host_f32(lim) if precision == "f32" else R.host_f64(lim)
    for kind, hil in ((0, True), (1, False)):
        pos = [dev(a) for a in xyz]
        k = zeros(s.n, torch.int32)
        ctx.call(f"ryo_sfc_keys_u32_{precision}", *(R.ptr(a) for a in pos), R.ptr(k), s.n, host_lim, kind, 0)
        g = host(k, np.uint32)
        o = O.ref_sfc_keys32(*xyz, lim, hilbert=hil, precision=precision)
        assert np.array_equal(g, o), f"{name} {precision} kind {kind}: {np.count_nonzero(g != o)} keys differ"
    # removeKey<uint32_t> entries are skipped
    pre = np.zeros(s.n, np.uint32)
    pre[::5] = np.uint32(1 << 30)
    k = dev(pre)
    ctx.call(f"ryo_sfc_keys_u32_{precision}", *(R.ptr(a) for a in pos), R.ptr(k), s.n, host_lim, 0, 1)
    g = host(k, np.uint32)
    assert np.all(g[::5] == np.uint32(1 << 30))
    mask = np.ones(s.n, bool)
    mask[::5] = False
    assert np.array_equal(g[mask], O.ref_sfc_keys32(*xyz, lim, precision=precision)[mask])
