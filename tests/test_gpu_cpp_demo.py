"""The C++ caller (ryoanji_amd/demo/demo.cpp), written against the reference's entry-point names in
include/ryoanji_amd.hpp, runs the whole path through the C ABI, meets the reference accuracy bounds and
equals the oracle on its own input."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ryoanji_amd", "demo", "demo")


def test_cpp_demo_runs_and_is_accurate(gpu):
    assert os.path.exists(DEMO), "build the demo with make -C ryoanji_amd/csrc"
    out = subprocess.run([DEMO, "16", "0.5"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    n, cells, p2p, m2p, pot, e50, e1, emax = out.stdout.split()
    assert int(n) == 65536 and int(cells) > 65536 // 64
    assert float(p2p) > 0 and float(m2p) > 0
    assert float(e1) < 2e-3 and float(emax) < 3e-2
    assert float(pot) < 0


def test_cpp_demo_equals_oracle(gpu, tmp_path):
    """the C++ drop-in path (TreeBuilder::update, upsweep, computeFixedGroups, traverse through the reference's
    names) gives the oracle's permutation, node keys, child offsets and layout bit for bit and its forces to
    1e-5 (p99, the path's tolerance), on the demo's own input"""
    import numpy as np

    import oracle as O
    from ryoanji_amd import workloads as W

    from gpu_util import oracle_tree, rel_err

    dump = tmp_path / "demo.bin"
    out = subprocess.run([DEMO, "15", "0.5", str(dump)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    raw = dump.read_bytes()
    n, nodes, leaves = np.frombuffer(raw[:24], np.int64)
    off = 24

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off)
        off += a.nbytes
        return a

    x0, y0, z0 = (take(np.float32, n) for _ in range(3))
    xs, ys, zs, ms, hs, p, ax, ay, az = (take(np.float32, n) for _ in range(9))
    prefixes, child, layout = take(np.uint64, nodes), take(np.int32, nodes), take(np.uint32, leaves + 1)
    assert off == len(raw)
    s = W.System(x0, y0, z0, np.full(n, 1.0 / n, np.float32), np.full(n, 1e-3, np.float32), [-1, 1, -1, 1, -1, 1])
    t = oracle_tree(s)
    assert np.array_equal(xs, t["x"]) and np.array_equal(ys, t["y"]) and np.array_equal(zs, t["z"])
    assert np.array_equal(prefixes, t["tree"]["prefixes"]) and np.array_equal(child, t["tree"]["childOffsets"])
    assert np.array_equal(layout, t["layout"])
    _, _, oax, oay, oaz, _ = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], t["x"], t["y"], t["z"],
                                               t["h"], t["m"], s.limits, group_size=64)
    e = np.sort(rel_err(ax, ay, az, oax, oay, oaz))
    assert e[int(0.99 * e.size)] <= 1e-5, e[int(0.99 * e.size)]
