"""Synthetic particle systems for the Barnes-Hut workloads (bit-reproducible from a seed).

One counter-based generator (splitmix64) drives every workload, so a system is a pure function
of (kind, n, seed); floats take the top 24 bits of a 64-bit draw, giving values in [0, 1).
"""

from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """count draws of the splitmix64 stream `seed`, starting at counter `start`."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, start: int, count: int) -> np.ndarray:
    """float64 in [0,1) with 24 random bits (exactly representable in float32)."""
    return (splitmix64(seed, start, count) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


def _stream(seed: int, k: int) -> int:
    return (seed * 0x100000001B3 + 0x51ED27 * (k + 1)) & 0xFFFFFFFFFFFFFFFF


class System:
    """SoA float32 particle arrays plus the box they live in ({xmin, xmax, ymin, ymax, zmin, zmax})."""

    def __init__(self, x, y, z, m, h, limits, periodic=False, name=""):
        self.x, self.y, self.z, self.m, self.h = (np.ascontiguousarray(a, np.float32) for a in (x, y, z, m, h))
        self.limits = np.asarray(limits, np.float32)
        self.periodic = periodic
        self.name = name

    @property
    def n(self):
        return self.x.size

    def arrays(self):
        return self.x, self.y, self.z, self.m, self.h


def uniform_cube(n: int, seed: int = 42, lo: float = -1.0, hi: float = 1.0, soft: float = 1e-3, periodic=False):
    """n particles uniform in [lo, hi)^3, mass 1/n, softening `soft`."""
    ext = hi - lo
    xyz = [np.float32(lo) + np.float32(ext) * uniform01(_stream(seed, d), 0, n).astype(np.float32) for d in range(3)]
    m = np.full(n, 1.0 / n, np.float32)
    h = np.full(n, soft, np.float32)
    return System(*xyz, m, h, [lo, hi] * 3, periodic=periodic, name=f"uniform{n}")


def uniform_box(n: int, seed: int = 42, extents=(2.0, 1.0, 0.25), soft: float = 1e-3, periodic=False):
    """n particles uniform in the box [0, e_x) x [0, e_y) x [0, e_z), mass 1/n. Axes shorter than 3/4 of the
    longest one get fewer key bits (mixed-dimension Hilbert keys)."""
    xyz = [np.float32(e) * uniform01(_stream(seed, d), 0, n).astype(np.float32) for d, e in enumerate(extents)]
    m = np.full(n, 1.0 / n, np.float32)
    h = np.full(n, soft, np.float32)
    lim = [0.0, extents[0], 0.0, extents[1], 0.0, extents[2]]
    return System(*xyz, m, h, lim, periodic=periodic, name=f"box{n}")


def plummer_disk(n: int, seed: int = 42, flattening: float = 0.2, rmax: float = 30.0, soft: float = 1e-3):
    """A Plummer sphere squashed along z by `flattening`, in the box [-rmax, rmax]^2 x [-f rmax, f rmax]."""
    s = plummer(n, seed, rmax, soft)
    z = (s.z * np.float32(flattening)).astype(np.float32)
    f = float(np.float32(flattening) * np.float32(rmax))
    return System(s.x, s.y, z, s.m, s.h, [-rmax, rmax, -rmax, rmax, -f, f], name=f"disk{n}")


def _plummer_radii(n: int, seed: int, rmax: float) -> np.ndarray:
    """Plummer radii (scale 1) by inverse-CDF r = (u^(-2/3) - 1)^(-1/2), redrawn while r > rmax."""
    r = np.empty(n)
    todo = np.arange(n)
    attempt = 0
    while todo.size:
        u = uniform01(_stream(seed, 10 + attempt), 0, n)[todo]
        u = np.maximum(u, 1.0 / (1 << 25))
        rr = 1.0 / np.sqrt(u ** (-2.0 / 3.0) - 1.0 + 1e-300)
        ok = rr <= rmax
        r[todo[ok]] = rr[ok]
        todo = todo[~ok]
        attempt += 1
    return r


def _isotropic(n: int, seed: int):
    v = uniform01(_stream(seed, 3), 0, n)
    w = uniform01(_stream(seed, 4), 0, n)
    cz = 1.0 - 2.0 * v
    sz = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
    phi = 2.0 * np.pi * w
    return sz * np.cos(phi), sz * np.sin(phi), cz


def plummer(n: int, seed: int = 42, rmax: float = 30.0, soft: float = 1e-3):
    """Plummer sphere, scale radius 1, truncated at rmax, equal masses 1/n, cubic box [-rmax, rmax]."""
    r = _plummer_radii(n, seed, rmax)
    ux, uy, uz = _isotropic(n, seed)
    x, y, z = ((r * u).astype(np.float32) for u in (ux, uy, uz))
    m = np.full(n, 1.0 / n, np.float32)
    h = np.full(n, soft, np.float32)
    return System(x, y, z, m, h, [-rmax, rmax] * 3, name=f"plummer{n}")


def hernquist_mix(n: int, seed: int = 42, halos: int = 64, background: float = 0.1, soft: float = 1e-3):
    """`halos` Hernquist halos (masses log-uniform over 1e4, scale radii log-uniform in [1e-3, 1e-1], centres
    uniform in [-1,1]^3, particle counts proportional to mass, truncated at 100 a) plus a `background`
    fraction of uniform particles in [-1,1]^3; equal masses 1/n; a cubic box around all particles."""
    nb = int(round(background * n))
    nh = n - nb
    hm = 10.0 ** (4.0 * uniform01(_stream(seed, 20), 0, halos))
    a = 10.0 ** (-3.0 + 2.0 * uniform01(_stream(seed, 21), 0, halos))
    cen = [-1.0 + 2.0 * uniform01(_stream(seed, 22 + d), 0, halos) for d in range(3)]
    counts = np.floor(nh * hm / hm.sum()).astype(np.int64)
    counts[np.argmax(counts)] += nh - counts.sum()
    owner = np.repeat(np.arange(halos), counts)
    r = np.empty(nh)
    todo = np.arange(nh)
    attempt = 0
    while todo.size:
        u = uniform01(_stream(seed, 30 + attempt), 0, nh)[todo]
        s = np.sqrt(u)
        rr = a[owner[todo]] * s / np.maximum(1.0 - s, 1e-300)
        ok = rr <= 100.0 * a[owner[todo]]
        r[todo[ok]] = rr[ok]
        todo = todo[~ok]
        attempt += 1
    ux, uy, uz = _isotropic(nh, seed + 7)
    hx = cen[0][owner] + r * ux
    hy = cen[1][owner] + r * uy
    hz = cen[2][owner] + r * uz
    bx, by, bz = (-1.0 + 2.0 * uniform01(_stream(seed, 40 + d), 0, nb) for d in range(3))
    x = np.concatenate([hx, bx]).astype(np.float32)
    y = np.concatenate([hy, by]).astype(np.float32)
    z = np.concatenate([hz, bz]).astype(np.float32)
    ext = float(max(np.abs(x).max(), np.abs(y).max(), np.abs(z).max())) * 1.0001 + 1e-6
    m = np.full(n, 1.0 / n, np.float32)
    h = np.full(n, soft, np.float32)
    return System(x, y, z, m, h, [-ext, ext] * 3, name=f"hernquist{n}")
