"""The C++ caller (ryoanji_amd/demo/demo.cpp), written against the reference's entry-point names in
include/ryoanji_amd.hpp, runs the whole path through the C ABI and meets the reference accuracy bounds."""

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
