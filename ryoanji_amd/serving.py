"""Served mixed-size stream: request generation and the continuous-batching server (ryo_server_* C ABI).

A stream is a list of Plummer systems with N log-uniform in [2^lo, 2^hi], theta drawn from {0.4, 0.5, 0.7}
and 1-20 leapfrog steps, arriving as a Poisson process. The same stream runs on S concurrent slots (continuous
batching under an HBM budget) and on one slot (one-at-a-time plain solves) for the throughput ratio.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import Request, Result, _P, lib
from . import workloads as W


def make_stream(num, rate_per_s, seed=0, lo_exp=12, hi_exp=22, steps=(1, 20), thetas=(0.4, 0.5, 0.7), dt=1e-3):
    """num requests; returns (list of System, dict of per-request parameters)."""
    rng = np.random.default_rng(seed)
    ns = np.floor(2.0 ** rng.uniform(lo_exp, hi_exp, num)).astype(np.int64)
    th = rng.choice(np.asarray(thetas, np.float32), num)
    st = rng.integers(steps[0], steps[1] + 1, num)
    gaps = rng.exponential(1000.0 / rate_per_s, num) if rate_per_s > 0 else np.zeros(num)
    arr = np.cumsum(gaps) - gaps[0]
    systems = [W.plummer(int(n), seed=seed * 1000 + i) for i, n in enumerate(ns)]
    return systems, dict(n=ns, theta=th, steps=st, arrival_ms=arr, dt=dt)


class Server:
    """slots resident-system workers and a batch lane for requests of at most batch_max_n particles (None: the
    library's default; 0: no lane, every request on a slot). Each slot and the lane run on a stream of their own:
    give the process at least slots + 1 hardware queues (the environment variable GPU_MAX_HW_QUEUES, read when HIP
    initialises; HIP's default is 4, shared round-robin between streams), or a stream's short kernels wait behind
    another's long ones in a shared queue. order "arrival" admits waiting requests first fit in arrival order;
    "least_work" admits the one with the fewest particle-solves first, except that requests that have waited
    max_wait_ms (when positive) go first in arrival order."""

    def __init__(self, slots, max_n, hbm_budget=0, batch_max_n=None, batch_capacity=1 << 22, batch_systems=1024,
                 order="arrival", max_wait_ms=0.0):
        self.h = _P()
        rc = lib().ryo_server_create(int(slots), int(hbm_budget), int(max_n), C.byref(self.h))
        if rc:
            raise RuntimeError(f"ryo_server_create failed ({rc})")
        if batch_max_n is not None:
            rc = lib().ryo_server_set_batching(self.h, int(batch_max_n), int(batch_capacity), int(batch_systems))
            if rc:
                raise RuntimeError(f"ryo_server_set_batching failed ({rc})")
        rc = lib().ryo_server_set_order(self.h, {"arrival": 0, "least_work": 1}[order], float(max_wait_ms))
        if rc:
            raise RuntimeError(f"ryo_server_set_order failed ({rc})")
        self.slots = slots

    def run(self, systems, params, keep_state=False, arrival_scale=1.0, accuracy_targets=0):
        num = len(systems)
        reqs = (Request * num)()
        keep = []
        for i, s in enumerate(systems):
            r = reqs[i]
            r.n = s.n
            arrs = [np.ascontiguousarray(a, np.float32) for a in s.arrays()]
            keep.append(arrs)
            r.x, r.y, r.z, r.m, r.h = (a.ctypes.data for a in arrs)
            for k in range(6):
                r.limits[k] = float(s.limits[k])
            r.periodic = int(s.periodic)
            r.theta = float(params["theta"][i])
            r.G = 1.0
            r.dt = float(params["dt"])
            r.steps = int(params["steps"][i])
            r.bucket_size = 64
            r.group_size = 64
            r.arrival_ms = float(params["arrival_ms"][i]) * arrival_scale
            r.accuracy_targets = int(accuracy_targets)
            # periodic requests: the reference's periodic set-up, 1 replica shell and the Ewald correction, unless
            # the stream's parameters say otherwise
            r.num_shells = int(np.broadcast_to(params.get("num_shells", 1), (num,))[i]) if s.periodic else 0
            r.ewald = int(np.broadcast_to(params.get("ewald", 1), (num,))[i]) if s.periodic else 0
        res = (Result * num)()
        states = None
        if keep_state:
            bufs = [np.zeros(7 * s.n, np.float32) for s in systems]
            states = (_P * num)(*[b.ctypes.data for b in bufs])
        rc = lib().ryo_server_run(self.h, num, reqs, res, states)
        if rc:
            raise RuntimeError(f"ryo_server_run failed ({rc})")
        out = [dict(arrival=r.arrival_ms, admit=r.admit_ms, finish=r.finish_ms, pot=r.total_pot, slot=r.slot,
                    status=r.status, err_median=r.err_median, err_p99=r.err_p99, err_max=r.err_max,
                    limits=np.array(list(r.limits), np.float32)) for r in res]
        if keep_state:
            for i, b in enumerate(bufs):
                n = systems[i].n
                out[i]["state"] = b.reshape(7, n)
        return out

    def close(self):
        if self.h:
            lib().ryo_server_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fit_box(x, y, z):
    """cubic box around all particles, padded by a relative 1e-4 (the resident leapfrog's fitBox)"""
    mn = np.array([x.min(), y.min(), z.min()], np.float64)
    mx = np.array([x.max(), y.max(), z.max()], np.float64)
    c = 0.5 * (mn + mx)
    half = 0.5 * float((mx - mn).max()) * (1 + 1e-4) + 1e-30
    lo = (c - half).astype(np.float32)
    L = np.float32(2 * half)
    return np.array([lo[0], lo[0] + L, lo[1], lo[1] + L, lo[2], lo[2] + L], np.float32)


def plain_solves(systems, params):
    """One-at-a-time plain solves of the same stream: each timestep of each request is one full pass from host
    memory (upload x,y,z,m,h, keys -> sort -> tree -> upsweep -> traversal, download the accelerations) with the
    kick-drift-kick on the host; nothing stays resident between steps and requests never overlap."""
    import time

    from . import Context, System, default_params

    ctx = Context()
    sysm = System(ctx, int(max(s.n for s in systems)))
    out = []
    t0 = time.perf_counter()
    try:
        for i, s in enumerate(systems):
            prm = default_params(theta=float(params["theta"][i]), group_size=64, bucket_size=64)
            dt = float(params["dt"])
            X = [np.ascontiguousarray(a, np.float32) for a in (s.x, s.y, s.z)]
            m, h = np.ascontiguousarray(s.m, np.float32), np.ascontiguousarray(s.h, np.float32)
            V = [np.zeros(s.n, np.float32) for _ in range(3)]

            def solve(lim):
                sysm.load(X[0], X[1], X[2], m, h, lim)
                sysm.build(prm)
                pot = sysm.gravity(prm)
                _, gx, gy, gz, o = (t.cpu().numpy() for t in sysm.results())
                A = [np.empty(s.n, np.float32) for _ in range(3)]
                for d, g in enumerate((gx, gy, gz)):
                    A[d][o] = g
                return A, pot

            A, pot = solve(s.limits)
            for _ in range(int(params["steps"][i])):
                for d in range(3):
                    V[d] += np.float32(0.5 * dt) * A[d]
                    X[d] += np.float32(dt) * V[d]
                A, pot = solve(_fit_box(*X))
                for d in range(3):
                    V[d] += np.float32(0.5 * dt) * A[d]
            out.append(dict(arrival=0.0, admit=0.0, finish=(time.perf_counter() - t0) * 1e3, pot=pot, slot=0,
                            status=0))
    finally:
        sysm.close()
        ctx.close()
    return out


def summarize(systems, params, results):
    lat = np.array([r["finish"] - r["arrival"] for r in results])
    makespan = max(r["finish"] for r in results) - min(r["arrival"] for r in results)
    work = float(sum(s.n * (int(k) + 1) for s, k in zip(systems, params["steps"])))  # gravity solves x particles
    return dict(requests=len(results), p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                makespan_ms=float(makespan), particle_solves_per_s=work / (makespan / 1e3),
                requests_per_s=len(results) / (makespan / 1e3))
