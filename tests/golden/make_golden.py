"""Write the golden fixtures of tests/golden/ from the REFERENCE CPU path (oracle/_ref/libryoref.so).

Run in a container where the reference checkout exists and `make -C oracle ref` has built the harness:
    python tests/golden/make_golden.py
Each fixture is a compressed .npz of inputs and the reference's outputs (keys, stable-sort permutation,
cornerstone leaves and counts, linked-octree arrays, float centers/MAC and quadrupoles, the CPU traversal's
accelerations and potentials with its 16-target groups, and the Ewald correction).
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from ryoanji_amd import workloads as W  # noqa: E402


def fixture(s, name, theta=0.5, num_shells=0, ewald=False, bucket=64):
    assert O.have_ref(), "the reference harness is not built"
    keys = O.sfc_keys(s.x, s.y, s.z, s.limits, lib="ref")
    sk, perm = O.sort_by_key(keys, lib="ref")
    x, y, z, m, h = (a[perm] for a in s.arrays())
    leaves, counts = O.compute_octree(sk, bucket, lib="ref")
    layout = O.layout_from_counts(counts)
    tree = O.link_octree(leaves, lib="ref")
    c, q = O.upsweep(x, y, z, m, leaves, layout, s.limits, 1.0 / theta, lib="ref")
    tot, ug, ax, ay, az, _ = O.compute_gravity(leaves, layout, c, q, x, y, z, h, m, s.limits, lib="ref",
                                               num_shells=num_shells)
    out = dict(x=s.x, y=s.y, z=s.z, m=s.m, h=s.h, limits=s.limits, theta=np.float32(theta),
               num_shells=np.int32(num_shells), bucket=np.int32(bucket), keys=keys, sorted_keys=sk, perm=perm,
               leaves=leaves, counts=counts, **{k: tree[k] for k in ("prefixes", "childOffsets", "parents",
                                                                     "levelRange", "internalToLeaf",
                                                                     "leafToInternal")},
               centers=c, multipoles=q, pot_total=np.float64(tot), ugrav=ug, ax=ax, ay=ay, az=az)
    if ewald:
        c64, q64 = O.upsweep(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64), m.astype(np.float64),
                             leaves, layout, s.limits.astype(np.float64), 1.0 / theta, precision="f64", lib="ref")
        etot, eug, eax, eay, eaz = O.ref_ewald(q64[0], c64[0, :3], x, y, z, m, s.limits.astype(np.float64))
        out.update(root_multipole=q64[0], root_center=c64[0, :3], ewald_total=np.float64(etot), ewald_ugrav=eug,
                   ewald_ax=eax, ewald_ay=eay, ewald_az=eaz)
    path = os.path.join(HERE, f"ref_{name}.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes;", leaves.size - 1, "leaves")


def main():
    fixture(W.uniform_cube(16384, seed=101), "uniform16384")
    fixture(W.plummer(8192, seed=102), "plummer8192", theta=0.6)
    fixture(W.hernquist_mix(8192, seed=103), "hernquist8192", bucket=32)
    fixture(W.uniform_cube(4096, seed=104, lo=0.0, hi=1.0, periodic=True), "periodic4096", num_shells=1, ewald=True)


if __name__ == "__main__":
    main()
