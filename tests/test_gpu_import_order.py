"""The library and torch share one HIP runtime whichever the caller loads first: a process that creates a server
(native streams only) before torch has touched the GPU can still use torch and a torch-backed context after it."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
from ryoanji_amd import serving as S
srv = S.Server(1, 4096, batch_max_n=0)          # the library's own HIP streams, before torch initialises the GPU
import torch
import ryoanji_amd as R
from ryoanji_amd import workloads as W
ctx = R.Context()                                 # torch's current stream
s = W.plummer(2048, seed=3)
sysm = R.System(ctx, s.n)
sysm.load(*[torch.from_numpy(a).cuda() for a in s.arrays()], s.limits)
prm = R.default_params()
sysm.build(prm)
pot = sysm.gravity(prm)
assert np.isfinite(pot) and pot < 0, pot
sysm.close(); ctx.close(); srv.close()
print("ok", flush=True)
"""


@pytest.mark.gpu
def test_library_before_torch():
    r = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
