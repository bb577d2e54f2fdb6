"""A C++ caller with the call sequence of the reference's single-GPU demonstrator (ryoanji/test/demo.cu:55-219) and
of TreeBuilder's cornerstone loop (interface/treebuilder.cuh:57-101), compiled against include/ryoanji_amd.hpp
(ryoanji_amd/demo/demo_ref.cpp): TreeBuilder<uint64_t>::update(x, y, z, n, box), the five-step upsweep with the
reference's names, computeFixedGroups into GroupData, stackSize, traverse from node 1, readBhStats, directSum<T>
and updateOctreeGpu iterated from {0, 2^63}. Its outputs are checked against the oracle (float) and the reference's
own double instantiations (oracle/_ref): bit-exact keys order, node keys, child offsets, layout and cornerstone
leaves and counts; centers and multipoles to the float pipeline's rounding; forces p99 <= 1e-5 on the same grouping;
the direct sum to FP64 rounding of its outputs."""

import os
import subprocess

import numpy as np
import pytest

import oracle as O
from ryoanji_amd import workloads as W

from gpu_util import oracle_tree, rel_err

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ryoanji_amd", "demo", "demo_ref")


def _run(tmp_path, typ, power, theta, group_size):
    assert os.path.exists(DEMO), "build the caller with make -C ryoanji_amd/csrc"
    dump = tmp_path / f"demo_ref_{typ}.bin"
    print(f"demo_ref {typ} 2^{power} theta {theta} groups {group_size}", flush=True)
    out = subprocess.run([DEMO, typ, str(power), str(theta), str(group_size), str(dump)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    raw = dump.read_bytes()
    n, nodes, leaves, tsize, iters, loop_leaves = (int(v) for v in np.frombuffer(raw[:48], np.int64))
    ft = np.float32 if tsize == 4 else np.float64
    off = 48

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off)
        off += a.nbytes
        return a

    d = dict(n=n, nodes=nodes, leaves=leaves, iterations=iters)
    d["x0"], d["y0"], d["z0"] = (take(ft, n) for _ in range(3))
    d["x"], d["y"], d["z"], d["m"], d["h"] = (take(ft, n) for _ in range(5))
    d["prefixes"], d["child"], d["layout"] = take(np.uint64, nodes), take(np.int32, nodes), take(np.uint32, leaves + 1)
    d["centers"], d["mp"] = take(ft, 4 * nodes).reshape(nodes, 4), take(ft, 8 * nodes).reshape(nodes, 8)
    d["p"], d["ax"], d["ay"], d["az"] = (take(ft, n) for _ in range(4))
    d["dp"], d["dax"], d["day"], d["daz"] = (take(ft, n) for _ in range(4))
    d["loop_leaves"], d["loop_counts"] = take(np.uint64, loop_leaves + 1), take(np.uint32, loop_leaves)
    d["stats"], d["pot"] = take(np.uint64, 5), float(take(np.float64, 1)[0])
    assert off == len(raw)
    d["summary"] = out.stdout.split()
    return d


def test_reference_call_sequence_float(tmp_path):
    d = _run(tmp_path, "f32", 15, 0.5, 64)
    n = d["n"]
    s = W.System(d["x0"], d["y0"], d["z0"], np.full(n, 1.0 / n, np.float32), np.full(n, 1e-3, np.float32),
                 [-1, 1, -1, 1, -1, 1])
    t = oracle_tree(s)
    # TreeBuilder::update sorts x, y, z only (the reference's treebuilder.cuh:74-79); m and h stay as given
    assert np.array_equal(d["x"], t["x"]) and np.array_equal(d["y"], t["y"]) and np.array_equal(d["z"], t["z"])
    assert np.array_equal(d["m"], s.m) and np.array_equal(d["h"], s.h)
    assert np.array_equal(d["prefixes"], t["tree"]["prefixes"])
    assert np.array_equal(d["child"], t["tree"]["childOffsets"][:d["nodes"]])
    assert np.array_equal(d["layout"], t["layout"])
    # updateOctreeGpu iterated from {0, 2^63}: the converged cornerstone leaves and counts
    assert np.array_equal(d["loop_leaves"], t["leaves"]) and np.array_equal(d["loop_counts"], t["counts"])
    assert d["iterations"] >= 1
    # the five-step upsweep against the reference float pipeline (centers with mac^2, quadrupoles); like the
    # reference demo the multipole loop stops at level 1, so the root's multipole is not formed (node 0 skipped)
    c, q = t["centers"], t["mp"]
    assert np.allclose(d["centers"][:, :3], c[:, :3], rtol=0, atol=2e-6)
    assert np.allclose(d["centers"][:, 3], c[:, 3], rtol=2e-5, atol=0)
    scale = np.abs(q[1:]).max(axis=0) + 1e-30
    assert np.all(np.abs(d["mp"][1:] - q[1:]) <= 2e-5 * scale)
    # traverse (64-target groups from node 1) against the oracle traversal on the same grouping
    _, _, oax, oay, oaz, _ = O.compute_gravity(t["leaves"], t["layout"], c, q, t["x"], t["y"], t["z"], t["h"], t["m"],
                                               s.limits, group_size=64)
    e = np.sort(rel_err(d["ax"], d["ay"], d["az"], oax, oay, oaz))
    assert e[int(0.99 * n)] <= 1e-5, e[int(0.99 * n)]
    assert d["stats"][0] > 0 and d["stats"][2] > 0
    # directSum<float>: FP64 accumulation rounded to float, against the oracle's FP64 direct sum
    ref = O.direct_sum(t["x"], t["y"], t["z"], t["h"], t["m"])
    assert np.allclose(d["dax"], ref[:, 1], rtol=1e-6, atol=1e-6 * np.abs(ref[:, 1]).max())
    assert np.allclose(d["dp"], ref[:, 0], rtol=1e-6, atol=0)
    # the accuracy bounds of the reference's global_forces_gpu test (theta 0.5)
    e1, emax = float(d["summary"][6]), float(d["summary"][7])
    assert e1 < 2e-3 and emax < 3e-2


def test_reference_call_sequence_double(tmp_path):
    """(d,d,d,d,d): the same sequence with double bodies, against the reference's double instantiations"""
    d = _run(tmp_path, "f64", 14, 0.5, 16)
    n = d["n"]
    lim = np.array([-1, 1, -1, 1, -1, 1], np.float64)
    keys = O.ref_sfc_keys64(d["x0"], d["y0"], d["z0"], lim)
    sk, perm = O.sort_by_key(keys, lib="ref")
    xs, ys, zs = (a[perm] for a in (d["x0"], d["y0"], d["z0"]))
    assert np.array_equal(d["x"], xs) and np.array_equal(d["y"], ys) and np.array_equal(d["z"], zs)
    leaves, counts = O.compute_octree(sk, 64, lib="ref")
    tree = O.link_octree(leaves, lib="ref")
    layout = O.layout_from_counts(counts)
    assert np.array_equal(d["prefixes"], tree["prefixes"])
    assert np.array_equal(d["child"], tree["childOffsets"][:d["nodes"]])
    assert np.array_equal(d["layout"], layout)
    assert np.array_equal(d["loop_leaves"], leaves) and np.array_equal(d["loop_counts"], counts)
    m = d["m"]
    c64, q64 = O.upsweep(xs, ys, zs, m, leaves, layout, lim, 2.0, precision="f64", lib="ref")
    assert np.allclose(d["centers"][:, :3], c64[:, :3], rtol=0, atol=1e-12)
    assert np.allclose(d["centers"][:, 3], c64[:, 3], rtol=1e-10, atol=0)
    assert np.allclose(d["mp"][1:], q64[1:], rtol=0, atol=1e-11 * np.abs(q64[1:]).max())
    h = d["h"]
    rtot, rug, rax, ray, raz = O.ref_compute_gravity64(leaves, layout, c64, q64, xs, ys, zs, h, m, lim)
    e = np.sort(rel_err(d["ax"], d["ay"], d["az"], rax, ray, raz))
    assert e[int(0.99 * n)] <= 1e-10, e[int(0.99 * n)]
    # directSum<double> on sampled targets against a plain FP64 evaluation of P2P<Exact>
    tg = np.random.default_rng(3).choice(n, 64, replace=False)
    for i in tg:
        dx, dy, dz = xs - xs[i], ys - ys[i], zs - zs[i]
        r2 = dx * dx + dy * dy + dz * dz
        r2e = np.maximum(r2, (h[i] + h) ** 2)
        w = m / (r2e * np.sqrt(r2e))
        assert np.isclose(d["dax"][i], np.sum(dx * w), rtol=1e-12, atol=1e-14)
        assert np.isclose(d["dp"][i], -m[i] * np.sum(w * r2), rtol=1e-12, atol=0)
