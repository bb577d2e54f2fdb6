"""Mixed-dimension Hilbert keys of the oracle (no GPU): the reference's bit-count rule, bijectivity, and equality
with the reference's own functions (oracle/_ref) for encode, decode, validity and the non-cubic pipeline."""

import numpy as np
import pytest

import oracle as O
from ryoanji_amd import workloads as W

BITS = [(21, 21, 21), (21, 20, 18), (18, 21, 20), (20, 18, 21), (21, 21, 19), (19, 21, 21), (21, 19, 21),
        (21, 17, 17), (16, 16, 21), (21, 20, 20), (20, 21, 20), (3, 2, 1), (1, 3, 2), (2, 2, 3)]


def test_box_bits_closed_form():
    """hilbert.hpp:110-116: a box with edge ratios (8, 4, 1) keeps 21 bits on x and loses 1 on y, 3 on z;
    box.hpp:193-198: reduction floor(log2 ratio + 0.585), i.e. one bit from a ratio of 4/3, two from 8/3"""
    assert O.box_bits([0, 8, 0, 4, 0, 1]) == (21, 20, 18)
    assert O.box_bits([-1, 1, -1, 1, -1, 1]) == (21, 21, 21)
    assert O.box_bits([0, 1, 0, 1.3, 0, 1]) == (21, 21, 21)
    assert O.box_bits([0, 1, 0, 1.34, 0, 1]) == (20, 21, 20)
    assert O.box_bits([0, 1, 0, 2.6, 0, 1]) == (20, 21, 20)
    assert O.box_bits([0, 1, 0, 2.7, 0, 1]) == (19, 21, 19)


@pytest.mark.parametrize("bits", [(3, 2, 1), (1, 3, 2), (2, 3, 3), (3, 3, 1), (2, 1, 2)])
def test_mixd_is_a_bijection_on_small_grids(bits):
    """every cell of a 2^bx x 2^by x 2^bz grid gets its own valid key, which decodes back to the cell"""
    pts = [(x, y, z) for x in range(1 << bits[0]) for y in range(1 << bits[1]) for z in range(1 << bits[2])]
    keys = [O.hilbert_mixd(*p, bits=bits) for p in pts]
    assert len(set(keys)) == len(pts)
    for p, k in zip(pts, keys):
        ok, q = O.decode_hilbert_mixd(k, bits)
        assert ok and q == p
    # the keys only use 3 bits per level of the longest axis
    assert max(keys) < 8 ** max(bits)


def test_mixd_decode_inverts_encode():
    rng = np.random.default_rng(4)
    for bits in BITS:
        for _ in range(300):
            p = tuple(int(rng.integers(0, 1 << b)) for b in bits)
            ok, q = O.decode_hilbert_mixd(O.hilbert_mixd(*p, bits=bits), bits)
            assert ok and q == p, (bits, p)


@pytest.mark.skipif(not O.have_ref(), reason="reference harness not built (no reference checkout)")
def test_mixd_encode_decode_match_reference():
    rng = np.random.default_rng(5)
    for bits in BITS:
        for _ in range(500):
            p = tuple(int(rng.integers(0, 1 << b)) for b in bits)
            assert O.hilbert_mixd(*p, bits=bits) == O.hilbert_mixd(*p, bits=bits, lib="ref"), (bits, p)
        # arbitrary keys: the same validity verdict, and the same point for reachable keys
        top = 3 * max(bits)
        for k in rng.integers(0, 1 << top, 500, dtype=np.uint64):
            ok_o, q_o = O.decode_hilbert_mixd(int(k), bits)
            ok_r, q_r = O.decode_hilbert_mixd(int(k), bits, lib="ref")
            assert ok_o == ok_r, (bits, int(k))
            if ok_o:
                assert q_o == q_r, (bits, int(k))


@pytest.mark.skipif(not O.have_ref(), reason="reference harness not built (no reference checkout)")
def test_box_bits_match_reference():
    rng = np.random.default_rng(6)
    for _ in range(300):
        e = np.exp(rng.uniform(-3, 3, 3)).astype(np.float32)
        lo = rng.uniform(-2, 2, 3).astype(np.float32)
        lim = np.array([lo[0], lo[0] + e[0], lo[1], lo[1] + e[1], lo[2], lo[2] + e[2]], np.float32)
        assert O.box_bits(lim) == O.box_bits(lim, lib="ref"), lim


@pytest.mark.skipif(not O.have_ref(), reason="reference harness not built (no reference checkout)")
@pytest.mark.parametrize("mk", [lambda: W.uniform_box(40000, seed=21, extents=(2.0, 1.0, 0.25)),
                                lambda: W.uniform_box(40000, seed=22, extents=(0.5, 2.0, 1.5)),
                                lambda: W.plummer_disk(40000, seed=23, flattening=0.2)],
                         ids=["box_2_1_0.25", "box_0.5_2_1.5", "plummer_disk"])
def test_noncubic_pipeline_equals_reference(mk):
    """keys, sort, leaves, linked tree, upsweep (MAC radii from the mixed-dimension cell boxes) and traversal"""
    s = mk()
    assert O.box_bits(s.limits) != (21, 21, 21)
    ko = O.sfc_keys(s.x, s.y, s.z, s.limits)
    assert np.array_equal(ko, O.sfc_keys(s.x, s.y, s.z, s.limits, lib="ref"))
    so, po = O.sort_by_key(ko)
    lo, co = O.compute_octree(so, 32)
    lr, cr = O.compute_octree(so, 32, lib="ref")
    assert np.array_equal(lo, lr) and np.array_equal(co, cr)
    to, tr = O.link_octree(lo), O.link_octree(lo, lib="ref")
    for k in ("prefixes", "childOffsets", "parents", "levelRange", "internalToLeaf", "leafToInternal"):
        assert np.array_equal(to[k], tr[k]), k
    x, y, z, m, h = (a[po] for a in s.arrays())
    lay = O.layout_from_counts(co)
    c, q = O.upsweep(x, y, z, m, lo, lay, s.limits, 2.0)
    c_r, q_r = O.upsweep(x, y, z, m, lo, lay, s.limits, 2.0, lib="ref")
    assert np.array_equal(c, c_r) and np.array_equal(q, q_r)
    go = O.compute_gravity(lo, lay, c, q, x, y, z, h, m, s.limits)
    gr = O.compute_gravity(lo, lay, c, q, x, y, z, h, m, s.limits, lib="ref")
    for i in (1, 2, 3, 4):
        assert np.array_equal(go[i], gr[i])
