"""CPU checker for the Barnes-Hut path: TEST INFRASTRUCTURE ONLY.

Python bindings (ctypes, numpy) over two CPU libraries built by ``oracle/Makefile``:

* ``_build/liboracle.so`` -- the restatement of the reference CPU path (``ryo_oracle.cpp``),
* ``_ref/libryoref.so``   -- the reference's own CPU functions, compiled from its headers
  (present only where the reference checkout was available at build time).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg use this
package, and only as the checker / the timed CPU baseline; the product library never loads it.
Both libraries expose the same C signatures (prefix ``orc_`` / ``ref_``), so every wrapper takes
``lib="orc"`` or ``lib="ref"``.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libryoref.so")

KEY_END = 1 << 63
MAX_LEVEL = 21

_libs: dict[str, C.CDLL] = {}

_P = C.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def lib(which: str = "orc") -> C.CDLL:
    if which in _libs:
        return _libs[which]
    path = ORACLE_SO if which == "orc" else REF_SO
    if not os.path.exists(path):
        raise RuntimeError(f"checker library {path} is not built (run `make -C oracle`)")
    L = C.CDLL(path)
    p = which + "_"
    sig = {
        "sfc_keys": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, C.c_int]),
        "sort_by_key": (None, [_P, _P, C.c_size_t]),
        "compute_octree": (C.c_long, [_P, C.c_size_t, C.c_uint, _P, _P, C.c_long]),
        "link_octree": (C.c_int, [_P, C.c_int, _P, _P, _P, _P, _P, _P]),
        "upsweep_f32": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, _P, _P, C.c_float, _P, _P]),
        "upsweep_f64": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, _P, _P, C.c_float, _P, _P]),
    }
    if which == "orc":
        sig.update({
            "compute_gravity_f32": (C.c_double, [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_float,
                                                 C.c_int, C.c_int, _P, _P, _P, _P, _P, C.c_int, _P]),
            "compute_gravity_f64": (C.c_double, [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_double,
                                                 C.c_int, C.c_int, _P, _P, _P, _P]),
            "direct_sum_f64": (C.c_int, [_P, _P, _P, _P, _P, C.c_size_t, _P, C.c_size_t, C.c_double, _P,
                                         C.c_int, _P]),
            "direct_sum_f32": (C.c_int, [_P, _P, _P, _P, _P, C.c_size_t, _P, C.c_size_t, C.c_double, _P,
                                         C.c_int, _P]),
            "ewald_f64": (C.c_double, [_P, _P, _P, _P, _P, _P, C.c_size_t, C.c_double, C.c_double, _P, _P, _P,
                                       _P, _P]),
            "ewald_consistent_f64": (C.c_double, [_P, _P, _P, _P, _P, _P, C.c_size_t, C.c_double, C.c_double, _P,
                                                  _P, _P, _P, _P]),
            "ewald_pair_sum_f64": (C.c_int, [_P, _P, _P, _P, _P, C.c_size_t, _P, C.c_size_t, C.c_double, C.c_double,
                                             C.c_double, _P]),
            "hilbert3d": (C.c_uint64, [C.c_uint, C.c_uint, C.c_uint, C.c_int]),
            "hilbert_mixd": (C.c_uint64, [C.c_uint, C.c_uint, C.c_uint, _P]),
            "decode_hilbert_mixd": (C.c_int, [C.c_uint64, _P, _P]),
            "box_bits": (None, [_P, _P]),
            "group_splits": (C.c_long, [C.c_uint, C.c_uint, _P, _P, _P, _P, _P, C.c_int, _P, _P, C.c_float,
                                        C.c_float, _P, C.c_long]),
            "compute_gravity_groups_f32": (C.c_double, [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                                        C.c_float, C.c_int, _P, C.c_long, _P, _P, _P, _P, _P]),
            "decode_hilbert3d": (None, [C.c_uint64, _P]),
            "p2p_f64": (None, [_P, _P, _P, C.c_double, C.c_double, C.c_double, C.c_int]),
            "m2p_f64": (None, [_P, _P, _P, _P]),
            "p2m_f64": (None, [_P, _P, _P, _P, C.c_uint, _P, _P]),
            "m2m_f64": (None, [_P, _P, _P, C.c_int, _P]),
        })
    else:
        sig.update({
            "compute_gravity_f32": (C.c_double, [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_float,
                                                 C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
            "compute_gravity_f64": (C.c_double, [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_float,
                                                 C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
            "sfc_keys_f64": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, C.c_int]),
            "sfc_keys_u32": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, C.c_int]),
            "upsweep_mdq_f32": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, _P, _P, C.c_float, _P, _P]),
            "upsweep_mdq_f64": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, _P, _P, C.c_float, _P, _P]),
            "compute_gravity_mdq_f32": (C.c_double, [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_float,
                                                     C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
            "compute_gravity_mdq_g64_f32": (C.c_double, [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                                         C.c_float, C.c_int, _P, _P, _P, _P]),
            "compute_gravity_mdq_f64": (C.c_double, [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_float,
                                                     C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
            "compute_gravity_mdq_df": (C.c_double, [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_float,
                                                    C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
            "sfc_keys_u32_f64": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, C.c_int]),
            "compute_gravity_df": (C.c_double, [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_float,
                                                C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
            "direct_sum_f64": (C.c_int, [_P, _P, _P, _P, _P, C.c_size_t, C.c_double, _P, C.c_int, _P, _P, _P,
                                         _P]),
            "ewald_f64": (C.c_double, [_P, _P, _P, _P, _P, _P, C.c_size_t, _P, C.c_double, _P, _P, _P, _P, _P]),
            "ewald_f32": (C.c_double, [_P, _P, _P, _P, _P, _P, C.c_size_t, _P, C.c_double, _P, _P, _P, _P, _P]),
            "hilbert_mixd": (C.c_uint64, [C.c_uint, C.c_uint, C.c_uint, _P]),
            "decode_hilbert_mixd": (C.c_int, [C.c_uint64, _P, _P]),
            "box_bits": (None, [_P, _P]),
            "hilbert_ibox": (None, [C.c_uint64, C.c_uint, _P, _P]),
        })
    for name, (res, args) in sig.items():
        f = getattr(L, p + name)
        f.restype = res
        f.argtypes = args
    _libs[which] = L
    return L


def _f(which, name):
    return getattr(lib(which), f"{which}_{name}")


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ----------------------------------------------------------------------------- pipeline


def sfc_keys(x, y, z, limits, hilbert=True, lib="orc", keys=None):
    x, y, z = (_c(a, np.float32) for a in (x, y, z))
    n = x.size
    keys = np.zeros(n, np.uint64) if keys is None else _c(keys, np.uint64).copy()
    lim = _c(limits, np.float32)
    rc = _f(lib, "sfc_keys")(_ptr(x), _ptr(y), _ptr(z), _ptr(keys), n, _ptr(lim), int(hilbert))
    if rc != 0:
        raise ValueError("sfc_keys: non-cubic boxes are not supported")
    return keys


def sort_by_key(keys, values=None, lib="orc"):
    k = _c(keys, np.uint64).copy()
    v = np.arange(k.size, dtype=np.int32) if values is None else _c(values, np.int32).copy()
    _f(lib, "sort_by_key")(_ptr(k), _ptr(v), k.size)
    return k, v


def compute_octree(sorted_keys, bucket, lib="orc"):
    k = _c(sorted_keys, np.uint64)
    cap = max(64, k.size // 2 + 64)
    while True:
        leaves = np.zeros(cap + 1, np.uint64)
        counts = np.zeros(cap, np.uint32)
        L = _f(lib, "compute_octree")(_ptr(k), k.size, bucket, _ptr(leaves), _ptr(counts), cap)
        if L >= 0:
            return leaves[: L + 1].copy(), counts[:L].copy()
        cap = -L


def link_octree(leaves, lib="orc"):
    leaves = _c(leaves, np.uint64)
    L = leaves.size - 1
    nI = (L - 1) // 7
    n = L + nI
    out = dict(
        prefixes=np.zeros(n, np.uint64),
        childOffsets=np.zeros(n + 1, np.int32),
        parents=np.zeros(max(1, (n - 1) // 8), np.int32),
        levelRange=np.zeros(MAX_LEVEL + 2, np.int32),
        internalToLeaf=np.zeros(n, np.int32),
        leafToInternal=np.zeros(n, np.int32),
    )
    _f(lib, "link_octree")(_ptr(leaves), L, *(_ptr(out[k]) for k in
                                                ("prefixes", "childOffsets", "parents", "levelRange",
                                                 "internalToLeaf", "leafToInternal")))
    out["childOffsets"] = out["childOffsets"][:n]
    out.update(numLeaves=L, numInternal=nI, numNodes=n)
    return out


def layout_from_counts(counts):
    lay = np.zeros(counts.size + 1, np.uint32)
    lay[1:] = np.cumsum(counts, dtype=np.uint64).astype(np.uint32)
    return lay


def upsweep(x, y, z, m, leaves, layout, limits, inv_theta, precision="f32", lib="orc"):
    """centers [n,4] (x,y,z,mac^2) and quadrupoles [n,8] (M,qxx,qxy,qxz,qyy,qyz,qzz,trace)."""
    leaves = _c(leaves, np.uint64)
    layout = _c(layout, np.uint32)
    L = leaves.size - 1
    n = L + (L - 1) // 7
    if precision == "f32":
        dt = np.float32
        xs = [_c(a, np.float32) for a in (x, y, z, m)]
    else:
        dt = np.float64
        xs = [_c(a, np.float32 if lib == "orc" else np.float64) for a in (x, y, z, m)]
    lim = _c(limits, dt)
    centers = np.zeros((n, 4), dt)
    mp = np.zeros((n, 8), dt)
    _f(lib, "upsweep_" + precision)(*(_ptr(a) for a in xs), _ptr(leaves), L, _ptr(layout), _ptr(lim),
                                     float(inv_theta), _ptr(centers), _ptr(mp))
    return centers, mp


def compute_gravity(tree_leaves, layout, centers, mp, x, y, z, h, m, limits, G=1.0, num_shells=0,
                    group_size=16, lib="orc", group_list=None, leaf_range=None, want_pot=True):
    """Returns (pot_total, ugrav, ax, ay, az, stats). lib="ref" is the reference computeGravity (16 targets)."""
    leaves = _c(tree_leaves, np.uint64)
    layout = _c(layout, np.uint32)
    L = leaves.size - 1
    arrs = [_c(a, np.float32) for a in (x, y, z, h, m)]
    n = arrs[0].size
    ug = np.zeros(n, np.float32) if want_pot else None
    ax, ay, az = (np.zeros(n, np.float32) for _ in range(3))
    c = _c(centers, np.float32)
    q = _c(mp, np.float32)
    lim = _c(limits, np.float32)
    stats = np.zeros(2, np.uint64)
    if lib == "orc":
        gl = None if group_list is None else _c(group_list, np.int32)
        tot = _f("orc", "compute_gravity_f32")(_ptr(leaves), L, _ptr(layout), _ptr(c), _ptr(q),
                                              *(_ptr(a) for a in arrs), _ptr(lim), float(G), num_shells,
                                              group_size, _ptr(ug), _ptr(ax), _ptr(ay), _ptr(az), _ptr(gl),
                                              0 if gl is None else gl.size, _ptr(stats))
    else:
        fl, ll = (0, L) if leaf_range is None else leaf_range
        tot = _f("ref", "compute_gravity_f32")(_ptr(leaves), L, _ptr(layout), _ptr(c), _ptr(q),
                                              *(_ptr(a) for a in arrs), _ptr(lim), float(G), num_shells, fl, ll,
                                              _ptr(ug), _ptr(ax), _ptr(ay), _ptr(az))
    return tot, ug, ax, ay, az, stats


def compute_gravity_groups(tree_leaves, layout, centers, mp, x, y, z, h, m, limits, bounds, G=1.0, num_shells=0):
    """the oracle traversal over arbitrary target groups bounds[g]..bounds[g+1] -> (pot, ugrav, ax, ay, az, stats)"""
    leaves = _c(tree_leaves, np.uint64)
    layout = _c(layout, np.uint32)
    arrs = [_c(a, np.float32) for a in (x, y, z, h, m)]
    n = arrs[0].size
    ug, ax, ay, az = (np.zeros(n, np.float32) for _ in range(4))
    b = _c(bounds, np.uint32)
    stats = np.zeros(2, np.uint64)
    tot = _f("orc", "compute_gravity_groups_f32")(_ptr(leaves), leaves.size - 1, _ptr(layout),
                                                 _ptr(_c(centers, np.float32)), _ptr(_c(mp, np.float32)),
                                                 *(_ptr(a) for a in arrs), _ptr(_c(limits, np.float32)), float(G),
                                                 num_shells, _ptr(b), b.size - 1, _ptr(ug), _ptr(ax), _ptr(ay),
                                                 _ptr(az), _ptr(stats))
    return tot, ug, ax, ay, az, stats


def compute_gravity_f64(tree_leaves, layout, centers, mp, x, y, z, h, m, limits, G=1.0, num_shells=0,
                        group_size=16):
    """The oracle traversal evaluated in double on the same (float) inputs -> (pot_total, ugrav, ax, ay, az)."""
    leaves = _c(tree_leaves, np.uint64)
    layout = _c(layout, np.uint32)
    L = leaves.size - 1
    arrs = [_c(a, np.float32) for a in (x, y, z, h, m)]
    n = arrs[0].size
    ug, ax, ay, az = (np.zeros(n, np.float64) for _ in range(4))
    tot = _f("orc", "compute_gravity_f64")(_ptr(leaves), L, _ptr(layout), _ptr(_c(centers, np.float32)),
                                          _ptr(_c(mp, np.float32)), *(_ptr(a) for a in arrs),
                                          _ptr(_c(limits, np.float32)), float(G), num_shells, group_size, _ptr(ug),
                                          _ptr(ax), _ptr(ay), _ptr(az))
    return tot, ug, ax, ay, az


def ref_sfc_keys64(x, y, z, limits, hilbert=True):
    """the reference computeSfcKeys for double coordinates and Box<double> (64-bit Hilbert or Morton keys)"""
    x, y, z = (_c(a, np.float64) for a in (x, y, z))
    keys = np.zeros(x.size, np.uint64)
    _f("ref", "sfc_keys_f64")(_ptr(x), _ptr(y), _ptr(z), _ptr(keys), x.size, _ptr(_c(limits, np.float64)),
                              int(hilbert))
    return keys


def ref_sfc_keys32(x, y, z, limits, hilbert=True, precision="f32"):
    """the reference computeSfcKeys with 32-bit keys (10 levels) for float or double coordinates and box"""
    dt = np.float32 if precision == "f32" else np.float64
    x, y, z = (_c(a, dt) for a in (x, y, z))
    keys = np.zeros(x.size, np.uint32)
    _f("ref", "sfc_keys_u32" if precision == "f32" else "sfc_keys_u32_f64")(
        _ptr(x), _ptr(y), _ptr(z), _ptr(keys), x.size, _ptr(_c(limits, dt)), int(hilbert))
    return keys


def ref_upsweep_mdq(x, y, z, m, leaves, layout, limits, inv_theta, precision="f32"):
    """the reference's upsweep with CartesianMDQpole (monopole, dipole, quadrupole): centers [n,4] (x,y,z,mac^2) and
    multipoles [n,12] (M, qxx, qxy, qxz, qyy, qyz, qzz, trace, px, py, pz, 0)"""
    dt = np.float32 if precision == "f32" else np.float64
    leaves = _c(leaves, np.uint64)
    layout = _c(layout, np.uint32)
    L = leaves.size - 1
    n = L + (L - 1) // 7
    xs = [_c(a, dt) for a in (x, y, z, m)]
    centers = np.zeros((n, 4), dt)
    mp = np.zeros((n, 12), dt)
    _f("ref", "upsweep_mdq_" + precision)(*(_ptr(a) for a in xs), _ptr(leaves), L, _ptr(layout),
                                          _ptr(_c(limits, dt)), float(inv_theta), _ptr(centers), _ptr(mp))
    return centers, mp


def ref_compute_gravity_mdq(tree_leaves, layout, centers, mp, x, y, z, h, m, limits, G=1.0, num_shells=0,
                            group_size=16):
    """the reference computeGravity (16-target groups) with CartesianMDQpole<float> -> (pot, ugrav, ax, ay, az);
    group_size=64 drives the reference's computeGravityGroup over 64-target groups (n a multiple of 64)"""
    leaves = _c(tree_leaves, np.uint64)
    layout = _c(layout, np.uint32)
    L = leaves.size - 1
    arrs = [_c(a, np.float32) for a in (x, y, z, h, m)]
    n = arrs[0].size
    ug, ax, ay, az = (np.zeros(n, np.float32) for _ in range(4))
    if group_size == 64:
        assert n % 64 == 0
        tot = _f("ref", "compute_gravity_mdq_g64_f32")(_ptr(leaves), L, _ptr(layout), _ptr(_c(centers, np.float32)),
                                                       _ptr(_c(mp, np.float32)), *(_ptr(a) for a in arrs),
                                                       _ptr(_c(limits, np.float32)), float(G), num_shells, _ptr(ug),
                                                       _ptr(ax), _ptr(ay), _ptr(az))
        return tot, ug, ax, ay, az
    assert group_size == 16
    tot = _f("ref", "compute_gravity_mdq_f32")(_ptr(leaves), L, _ptr(layout), _ptr(_c(centers, np.float32)),
                                               _ptr(_c(mp, np.float32)), *(_ptr(a) for a in arrs),
                                               _ptr(_c(limits, np.float32)), float(G), num_shells, 0, L, _ptr(ug),
                                               _ptr(ax), _ptr(ay), _ptr(az))
    return tot, ug, ax, ay, az


def ref_compute_gravity64(tree_leaves, layout, centers, mp, x, y, z, h, m, limits, G=1.0, num_shells=0, mixed=False,
                          mdq=False):
    """The reference computeGravity (16-target groups) with double coordinates and centers: mixed=False is the
    all-double instantiation (CartesianQuadrupole<double>, double outputs), mixed=True has float h, m, multipoles
    and outputs; mdq=True takes CartesianMDQpole multipoles ([n,12]). -> (pot_total, ugrav, ax, ay, az)"""
    leaves = _c(tree_leaves, np.uint64)
    layout = _c(layout, np.uint32)
    L = leaves.size - 1
    ft = np.float32 if mixed else np.float64
    xyz = [_c(a, np.float64) for a in (x, y, z)]
    hm = [_c(a, ft) for a in (h, m)]
    n = xyz[0].size
    ug, ax, ay, az = (np.zeros(n, ft) for _ in range(4))
    name = ("compute_gravity_mdq_" if mdq else "compute_gravity_") + ("df" if mixed else "f64")
    tot = _f("ref", name)(
        _ptr(leaves), L, _ptr(layout), _ptr(_c(centers, np.float64)), _ptr(_c(mp, ft)), *(_ptr(a) for a in xyz),
        *(_ptr(a) for a in hm), _ptr(_c(limits, np.float64)), float(G), num_shells, 0, L, _ptr(ug), _ptr(ax), _ptr(ay),
        _ptr(az))
    return tot, ug, ax, ay, az


def direct_sum(x, y, z, h, m, targets=None, G=1.0, box_l=None, num_shells=0, precision="f64"):
    """FP64 direct sum (P2P<Exact>) over all sources for the given target indices -> [nt, 4]."""
    arrs = [_c(a, np.float32) for a in (x, y, z, h, m)]
    n = arrs[0].size
    t = None if targets is None else _c(targets, np.int64)
    nt = n if t is None else t.size
    dt = np.float64 if precision == "f64" else np.float32
    out = np.zeros((nt, 4), dt)
    bl = None if box_l is None else _c(box_l, np.float64)
    _f("orc", "direct_sum_" + precision)(*(_ptr(a) for a in arrs), n, _ptr(t), nt, float(G), _ptr(bl),
                                         num_shells, _ptr(out))
    return out


def ewald_pair_sum(x, y, z, h, m, L, targets=None, G=1.0, alpha_scale=6.5):
    """exact periodic potential/acceleration (FP64 Ewald pair sum, softened pairs) -> [nt, 4] {G m phi, G a}"""
    arrs = [_c(a, np.float32) for a in (x, y, z, h, m)]
    n = arrs[0].size
    t = None if targets is None else _c(targets, np.int64)
    nt = n if t is None else t.size
    out = np.zeros((nt, 4), np.float64)
    rc = _f("orc", "ewald_pair_sum_f64")(*(_ptr(a) for a in arrs), n, _ptr(t), nt, float(G), float(L),
                                         float(alpha_scale), _ptr(out))
    if rc:
        raise ValueError("ewald_pair_sum: alpha_scale must exceed 5")
    return out


EWALD_DEFAULT = (1, 2.6, 2.8, 2.0, 3.0e-3)


def ewald(mroot, center, x, y, z, m, L, G=1.0, settings=EWALD_DEFAULT, consistent=False):
    """the reference's Ewald correction; consistent=True takes the k-space potential consistent with the k-space
    acceleration (the reference's potential has the opposite sign there; accelerations are the same)"""
    arrs = [_c(a, np.float32) for a in (x, y, z, m)]
    n = arrs[0].size
    ug, ax, ay, az = (np.zeros(n, np.float64) for _ in range(4))
    s = _c(settings, np.float64)
    tot = _f("orc", "ewald_consistent_f64" if consistent else "ewald_f64")(_ptr(_c(mroot, np.float64)), _ptr(_c(center, np.float64)), *(_ptr(a) for a in arrs),
                                 n, float(L), float(G), _ptr(s), _ptr(ug), _ptr(ax), _ptr(ay), _ptr(az))
    return tot, ug, ax, ay, az


def ref_ewald(mroot, center, x, y, z, m, limits, G=1.0, settings=EWALD_DEFAULT, precision="f64"):
    dt = np.float64 if precision == "f64" else np.float32
    arrs = [_c(a, dt) for a in (x, y, z, m)]
    n = arrs[0].size
    ug, ax, ay, az = (np.zeros(n, dt) for _ in range(4))
    s = _c(settings, np.float64)
    tot = _f("ref", "ewald_" + precision)(_ptr(_c(mroot, dt)), _ptr(_c(center, dt)), *(_ptr(a) for a in arrs), n,
                                          _ptr(_c(limits, dt)), float(G), _ptr(s), _ptr(ug), _ptr(ax), _ptr(ay),
                                          _ptr(az))
    return tot, ug, ax, ay, az


def hilbert3d(ix, iy, iz, order=MAX_LEVEL):
    return int(_f("orc", "hilbert3d")(ix, iy, iz, order))


def decode_hilbert3d(key):
    out = np.zeros(3, np.uint32)
    _f("orc", "decode_hilbert3d")(int(key), _ptr(out))
    return tuple(int(v) for v in out)


def group_splits(first, last, x, y, z, h, leaves, layout, limits, tol_factor=2.0, extent_ratio=8.0):
    """computeGroupSplits with 64-target windows -> group boundaries (G+1 entries)"""
    arrs = [_c(a, np.float32) for a in (x, y, z)]
    hh = None if h is None else _c(h, np.float32)
    lv = _c(leaves, np.uint64)
    lay = _c(layout, np.uint32)
    cap = int(last - first) + 2
    out = np.zeros(cap, np.uint32)
    G = _f("orc", "group_splits")(int(first), int(last), *(_ptr(a) for a in arrs), _ptr(hh), _ptr(lv), lv.size - 1,
                                  _ptr(lay), _ptr(_c(limits, np.float32)), float(tol_factor), float(extent_ratio),
                                  _ptr(out), cap)
    return out[:G + 1].copy()


def hilbert_mixd(ix, iy, iz, bits, lib="orc"):
    return int(_f(lib, "hilbert_mixd")(ix, iy, iz, _ptr(_c(bits, np.uint32))))


def decode_hilbert_mixd(key, bits, lib="orc"):
    """(valid, (x, y, z)) of a mixed-dimension Hilbert key"""
    out = np.zeros(3, np.uint32)
    ok = _f(lib, "decode_hilbert_mixd")(int(key), _ptr(_c(bits, np.uint32)), _ptr(out))
    return bool(ok), tuple(int(v) for v in out)


def box_bits(limits, lib="orc"):
    out = np.zeros(3, np.uint32)
    _f(lib, "box_bits")(_ptr(_c(limits, np.float32)), _ptr(out))
    return tuple(int(v) for v in out)


def ref_hilbert_ibox(key, level, bits):
    out = np.zeros(6, np.int32)
    _f("ref", "hilbert_ibox")(int(key), int(level), _ptr(_c(bits, np.uint32)), _ptr(out))
    return tuple(int(v) for v in out)


def p2p(acc, pi, pj, mj, hi, hj, exact=False):
    a = _c(acc, np.float64).copy()
    _f("orc", "p2p_f64")(_ptr(a), _ptr(_c(pi, np.float64)), _ptr(_c(pj, np.float64)), mj, hi, hj, int(exact))
    return a


def m2p(acc, target, center, q):
    a = _c(acc, np.float64).copy()
    _f("orc", "m2p_f64")(_ptr(a), _ptr(_c(target, np.float64)), _ptr(_c(center, np.float64)),
                         _ptr(_c(q, np.float64)))
    return a


def p2m(x, y, z, m):
    arrs = [_c(a, np.float64) for a in (x, y, z, m)]
    com = np.zeros(4)
    q = np.zeros(8)
    _f("orc", "p2m_f64")(*(_ptr(a) for a in arrs), arrs[0].size, _ptr(com), _ptr(q))
    return com, q


def m2m(xout, xsrc, msrc):
    xsrc = _c(xsrc, np.float64)
    msrc = _c(msrc, np.float64)
    out = np.zeros(8)
    _f("orc", "m2m_f64")(_ptr(_c(xout, np.float64)), _ptr(xsrc), _ptr(msrc), xsrc.shape[0], _ptr(out))
    return out


def full_tree(x, y, z, m, limits, bucket=64, inv_theta=2.0, lib="orc"):
    """keys -> stable sort -> permute -> leaves -> linked tree -> layout -> float upsweep (reference sequence)."""
    keys = sfc_keys(x, y, z, limits, lib=lib)
    skeys, perm = sort_by_key(keys, lib=lib)
    leaves, counts = compute_octree(skeys, bucket, lib=lib)
    layout = layout_from_counts(counts)
    return dict(keys=keys, sorted_keys=skeys, perm=perm, leaves=leaves, counts=counts, layout=layout)
