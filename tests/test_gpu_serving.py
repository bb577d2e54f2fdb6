"""Serving: continuous batching gives each request the same result alone and in the mix; leapfrog state stays
resident on the device; admission respects the HBM budget."""

import numpy as np
import pytest

import ryoanji_amd as R
from ryoanji_amd import serving as S
from ryoanji_amd import workloads as W

from gpu_util import dev, host

pytestmark = pytest.mark.gpu


def test_same_result_alone_and_in_the_mix(gpu):
    """the mix: 4 slots and the batch lane (requests of at most 2^14 particles packed together, the others on
    slots); alone: one slot and no lane, one request at a time"""
    systems, params = S.make_stream(12, rate_per_s=0, seed=3, lo_exp=10, hi_exp=17, steps=(1, 4))
    mixed = S.Server(slots=4, max_n=1 << 17, batch_max_n=1 << 14)
    out_mix = mixed.run(systems, params, keep_state=True)
    mixed.close()
    alone = S.Server(slots=1, max_n=1 << 17, batch_max_n=0)
    out_alone = alone.run(systems, params, keep_state=True)
    alone.close()
    assert all(r["status"] == 0 for r in out_mix + out_alone)
    lane = [r for r, sd in zip(out_mix, systems) if sd.n <= 1 << 14]
    assert lane and all(r["slot"] == 4 for r in lane)  # the small requests ran in the batch lane
    assert len({r["slot"] for r in out_mix}) > 2  # and the others on several slots
    # several lane members were resident at once (packed into one batched solve per step)
    assert any(a["admit"] < b["finish"] and b["admit"] < a["finish"] for a in lane for b in lane if a is not b)
    for a, b in zip(out_alone, out_mix):
        # identical kernels on identical data: bit-identical states (the stated tolerance is 1e-6 relative)
        sa, sb = a["state"], b["state"]
        assert np.array_equal(sa[6].view(np.int32), sb[6].view(np.int32))
        err = np.abs(sa[:6] - sb[:6]).max() / (np.abs(sa[:6]).max() + 1e-30)
        assert err <= 1e-6, err
        assert abs(a["pot"] - b["pot"]) <= 1e-9 * abs(a["pot"])


def _oracle_check(system, result, theta, group_list=None):
    """the served final state (positions in the final solve's SFC order, accelerations of that solve) against
    the oracle traversal of the same positions in the same box: the oracle's sort leaves the served order
    unchanged, and the accelerations agree at p99 <= 1e-5 (over all targets, or the 64-target groups listed)"""
    import oracle as O
    from gpu_util import oracle_tree, rel_err

    st = result["state"]
    order = st[6].view(np.int32)
    x, y, z = st[0], st[1], st[2]
    s = W.System(x, y, z, system.m[order], system.h[order], result["limits"])
    t = oracle_tree(s, inv_theta=1.0 / float(theta))
    assert np.array_equal(t["perm"], np.arange(s.n)), "served positions are not in the sort order of their box"
    if group_list is None:
        _, _, oax, oay, oaz, _ = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], t["x"], t["y"],
                                                   t["z"], t["h"], t["m"], s.limits, group_size=64)
        idx = np.arange(s.n)
    else:
        _, _, oax, oay, oaz, _ = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], t["x"], t["y"],
                                                   t["z"], t["h"], t["m"], s.limits, group_size=64,
                                                   group_list=group_list, want_pot=False)
        idx = np.concatenate([np.arange(64 * g, min(64 * g + 64, s.n)) for g in group_list])
    e = np.sort(rel_err(st[3][idx], st[4][idx], st[5][idx], oax[idx], oay[idx], oaz[idx]))
    assert e[int(0.99 * (e.size - 1))] <= 1e-5, (s.n, e[int(0.99 * (e.size - 1))])


def test_served_results_match_oracle(gpu):
    """every served request's final accelerations against the oracle on its final positions and box"""
    systems, params = S.make_stream(8, rate_per_s=0, seed=13, lo_exp=12, hi_exp=16, steps=(1, 3))
    srv = S.Server(slots=4, max_n=1 << 16, batch_max_n=1 << 14)  # the lane and the slots both serve
    out = srv.run(systems, params, keep_state=True)
    srv.close()
    assert len({r["slot"] for r in out}) > 1
    for i, (sysd, r) in enumerate(zip(systems, out)):
        assert r["status"] == 0
        print(f"served request {i}: n={sysd.n} theta={params['theta'][i]} steps={params['steps'][i]}", flush=True)
        _oracle_check(sysd, r, params["theta"][i])


def test_served_large_request_matches_oracle(gpu):
    """a stream with a request of 2^20 particles (the served stream's size range reaches 2^22) beside small ones;
    the large request is checked against the oracle on 48 seeded target groups, the small ones in full"""
    systems, params = S.make_stream(4, rate_per_s=0, seed=14, lo_exp=12, hi_exp=14, steps=(1, 2))
    systems[0] = W.plummer(1 << 20, seed=1401)
    params["steps"][0] = 2
    params["theta"][0] = np.float32(0.5)
    srv = S.Server(slots=2, max_n=1 << 20)
    out = srv.run(systems, params, keep_state=True)
    srv.close()
    assert all(r["status"] == 0 for r in out)
    groups = np.sort(np.random.default_rng(5).choice((1 << 20) // 64, 48, replace=False)).astype(np.int32)
    _oracle_check(systems[0], out[0], params["theta"][0], group_list=groups)
    for i in range(1, len(systems)):
        _oracle_check(systems[i], out[i], params["theta"][i])


def test_periodic_requests_alone_and_in_the_mix(gpu):
    """periodic requests (1 replica shell + the Ewald correction, wrapped into their boxes) served in the batch
    lane beside open ones: the same final states as one at a time on one slot, positions inside their boxes, and
    the final accelerations equal to the oracle traversal over the images plus the oracle's FP64 Ewald correction"""
    import oracle as O
    from gpu_util import oracle_tree

    systems = [W.uniform_cube(3000, seed=71, lo=0.0, hi=1.0, periodic=True), W.plummer(5000, seed=72),
               W.uniform_cube(9000, seed=73, lo=-1.0, hi=1.0, periodic=True), W.plummer(20000, seed=74),
               W.uniform_cube(1500, seed=75, lo=0.0, hi=2.0, periodic=True), W.plummer(40000, seed=76)]
    n = len(systems)
    params = dict(theta=np.array([0.5, 0.5, 0.6, 0.4, 0.5, 0.7], np.float32), steps=np.array([2, 3, 1, 2, 3, 1]),
                  arrival_ms=np.zeros(n), dt=2e-3, num_shells=1, ewald=1)
    mixed = S.Server(slots=2, max_n=1 << 16, batch_max_n=1 << 14)
    out_mix = mixed.run(systems, params, keep_state=True)
    mixed.close()
    alone = S.Server(slots=1, max_n=1 << 16, batch_max_n=0)
    out_alone = alone.run(systems, params, keep_state=True)
    alone.close()
    assert all(r["status"] == 0 for r in out_mix + out_alone)
    assert all(r["slot"] == 2 for r, sd in zip(out_mix, systems) if sd.n <= 1 << 14)  # lane members
    for i, (a, b) in enumerate(zip(out_alone, out_mix)):
        print(f"request {i}: n={systems[i].n} periodic={systems[i].periodic}", flush=True)
        sa, sb = a["state"], b["state"]
        assert np.array_equal(sa[6].view(np.int32), sb[6].view(np.int32)), i
        err = np.abs(sa[:6] - sb[:6]).max() / (np.abs(sa[:6]).max() + 1e-30)
        assert err <= 1e-6, (i, err)
        assert abs(a["pot"] - b["pot"]) <= 1e-9 * abs(a["pot"]), i
    for i, (sysd, r) in enumerate(zip(systems, out_mix)):
        if not sysd.periodic:
            continue
        st = r["state"]
        lim = np.asarray(sysd.limits, np.float32)
        assert np.array_equal(r["limits"], lim), i  # a periodic request keeps its box
        for d in range(3):
            assert np.all((st[d] >= lim[2 * d]) & (st[d] < lim[2 * d + 1])), (i, d)
        order = st[6].view(np.int32)
        s = W.System(st[0], st[1], st[2], sysd.m[order], sysd.h[order], lim, periodic=True)
        t = oracle_tree(s, inv_theta=1.0 / float(params["theta"][i]))
        assert np.array_equal(t["perm"], np.arange(s.n))
        _, _, oax, oay, oaz, _ = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], t["x"], t["y"],
                                                   t["z"], t["h"], t["m"], lim, group_size=64, num_shells=1)
        c64, q64 = O.upsweep(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"], lim.astype(np.float64),
                             1.0 / float(params["theta"][i]), precision="f64")
        _, _, eax, eay, eaz = O.ewald(q64[0], c64[0, :3], t["x"], t["y"], t["z"], t["m"], float(lim[1] - lim[0]))
        rms = np.sqrt(np.mean(oax ** 2 + oay ** 2 + oaz ** 2))
        d = np.sqrt((st[3] - oax - eax) ** 2 + (st[4] - oay - eay) ** 2 + (st[5] - oaz - eaz) ** 2) / rms
        assert np.quantile(d, 0.99) <= 1e-4, (i, np.quantile(d, 0.99))


def test_leapfrog_matches_one_solve_per_step(ctx):
    """the resident leapfrog equals kick-drift-kick done by hand with fresh solves"""
    s = W.plummer(20000, seed=5)
    dt = 1e-3
    sysm = R.System(ctx, s.n)
    sysm.load(*(dev(a) for a in s.arrays()), s.limits)
    prm = R.default_params()
    import ctypes

    pot = ctypes.c_double()
    ctx._check(R.lib().ryo_system_leapfrog(sysm.h, ctypes.byref(prm), dt, 2, ctypes.byref(pot)))
    p, ax, ay, az, order = (host(t) for t in sysm.results())
    x, y, z = (host(t) for t in sysm.arrays()[:3])
    # by hand: one solve per step from numpy state
    ref = R.System(ctx, s.n)
    X = [s.x.astype(np.float32).copy(), s.y.copy(), s.z.copy()]
    V = [np.zeros(s.n, np.float32) for _ in range(3)]
    idx = np.arange(s.n)

    def solve(X, lim):
        ref.load(dev(X[0]), dev(X[1]), dev(X[2]), dev(s.m[idx]), dev(s.h[idx]), lim)
        ref.build(prm)
        ref.gravity(prm)
        _, gx, gy, gz, o = (host(t) for t in ref.results())
        A = [np.empty(s.n, np.float32) for _ in range(3)]
        for d, g in enumerate((gx, gy, gz)):
            A[d][o] = g
        return A

    A = solve(X, s.limits)
    for _ in range(2):
        for d in range(3):
            V[d] = V[d] + np.float32(0.5 * dt) * A[d]
            X[d] = X[d] + np.float32(dt) * V[d]
        mn = np.array([X[d].min() for d in range(3)])
        mx = np.array([X[d].max() for d in range(3)])
        A = solve(X, _box(mn, mx))  # a tree on a slightly different box changes forces only at the BH-error level
        for d in range(3):
            V[d] = V[d] + np.float32(0.5 * dt) * A[d]
    for d, got in enumerate((x, y, z)):
        want = X[d][order]
        assert np.allclose(got, want, rtol=0, atol=1e-5 * np.abs(want).max())


def _box(mn, mx):
    c = 0.5 * (mn + mx)
    half = np.float32(0.5 * (mx - mn).max() * (1 + 1e-4) + 1e-30)
    lo = (c - half).astype(np.float32)
    L = np.float32(lo[0] + 2 * half) - lo[0]
    return np.array([lo[0], lo[0] + L, lo[1], lo[1] + L, lo[2], lo[2] + L], np.float32)


def test_served_results_report_their_accuracy(gpu):
    systems, params = S.make_stream(6, rate_per_s=0, seed=6, lo_exp=11, hi_exp=14, steps=(1, 3))
    params["theta"][:] = 0.5
    srv = S.Server(slots=3, max_n=1 << 14)
    out = srv.run(systems, params, accuracy_targets=256)
    srv.close()
    for r in out:
        assert r["status"] == 0
        assert 0 < r["err_median"] <= r["err_p99"] <= r["err_max"]
        assert r["err_p99"] < 2e-3 and r["err_max"] < 3e-2


def test_admission_budget_limits_concurrency(gpu):
    systems, params = S.make_stream(6, rate_per_s=0, seed=4, lo_exp=14, hi_exp=14.01, steps=(1, 1))
    fp = R.lib().ryo_server_footprint(systems[0].n)
    srv = S.Server(slots=4, max_n=1 << 15, hbm_budget=int(2.5 * fp))
    out = srv.run(systems, params)
    srv.close()
    # at most two requests can be resident at once under a budget of 2.5 footprints
    ev = sorted([(r["admit"], 1) for r in out] + [(r["finish"], -1) for r in out])
    live = peak = 0
    for _, d in ev:
        live += d
        peak = max(peak, live)
    assert peak <= 2


def test_least_work_order(gpu):
    """one slot, no lane, everything queued at t=0: least work first admits the requests in ascending
    n x (steps + 1) and computes what arrival order computes; a wait cap every request has already reached gives
    arrival order back"""
    systems, params = S.make_stream(8, rate_per_s=0, seed=6, lo_exp=10, hi_exp=14, steps=(1, 3))
    outs = {}
    for name, order, cap in (("arrival", "arrival", 0.0), ("least", "least_work", 0.0),
                             ("capped", "least_work", 1e-6)):
        srv = S.Server(slots=1, max_n=1 << 14, batch_max_n=0, order=order, max_wait_ms=cap)
        outs[name] = srv.run(systems, params, keep_state=True)
        srv.close()
        assert all(r["status"] == 0 for r in outs[name])
    work = [s.n * (int(k) + 1) for s, k in zip(systems, params["steps"])]
    by_admit = lambda out: [i for i, _ in sorted(enumerate(out), key=lambda e: e[1]["admit"])]
    assert by_admit(outs["least"]) == sorted(range(len(systems)), key=lambda i: (work[i], i))
    assert by_admit(outs["arrival"]) == list(range(len(systems)))
    assert by_admit(outs["capped"]) == list(range(len(systems)))
    assert by_admit(outs["least"]) != by_admit(outs["arrival"])  # the seed gives a stream out of work order
    for a, b in zip(outs["arrival"], outs["least"]):
        assert np.array_equal(a["state"], b["state"]) and a["pot"] == b["pot"]


def test_least_work_order_with_slots_and_lane(gpu):
    """several slots and the batch lane: the admission order changes which requests share the lane and the device,
    not their results (the final order bit for bit, states and potentials to the tolerances of the alone-vs-mix
    test, lane members included)"""
    systems, params = S.make_stream(12, rate_per_s=0, seed=8, lo_exp=10, hi_exp=16, steps=(1, 4))
    outs = []
    for order in ("arrival", "least_work"):
        srv = S.Server(slots=4, max_n=1 << 16, batch_max_n=1 << 13, order=order, max_wait_ms=5000.0)
        outs.append(srv.run(systems, params, keep_state=True))
        srv.close()
    a, b = outs
    assert all(r["status"] == 0 for r in a + b)
    assert any(r["slot"] == 4 for r in b) and any(r["slot"] < 4 for r in b)  # lane and slots both used
    for ra, rb in zip(a, b):
        sa, sb = ra["state"], rb["state"]
        assert np.array_equal(sa[6].view(np.int32), sb[6].view(np.int32))
        assert np.abs(sa[:6] - sb[:6]).max() <= 1e-6 * (np.abs(sa[:6]).max() + 1e-30)
        assert abs(ra["pot"] - rb["pot"]) <= 1e-9 * abs(ra["pot"])
