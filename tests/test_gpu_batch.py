"""Batches of systems (ryo_batch): every system of a batch gets exactly the keys, tree, multipoles and forces
it gets when solved on its own (ryo_system), including single-leaf, tiny and non-cubic systems."""

import numpy as np
import pytest

import ryoanji_amd as R
from ryoanji_amd import workloads as W

from gpu_util import dev, host

pytestmark = pytest.mark.gpu


def _systems():
    out = []
    kinds = [lambda n, s: W.uniform_cube(n, seed=s), lambda n, s: W.plummer(n, seed=s),
             lambda n, s: W.uniform_box(n, seed=s, extents=(2.0, 1.0, 0.25)),
             lambda n, s: W.plummer_disk(n, seed=s, flattening=0.3)]
    sizes = [1, 2, 63, 64, 65, 1000, 4095, 4096, 4097, 20000, 3, 777, 1025, 9000, 64, 12345]
    for i, n in enumerate(sizes):
        out.append(kinds[i % len(kinds)](n, 100 + i))
    return out


def _alone(ctx, s, prm):
    sysm = R.System(ctx, s.n)
    arrs = [dev(a) for a in s.arrays()]
    sysm.load(*arrs, s.limits)
    sysm.build(prm)
    pot = sysm.gravity(prm)
    res = [host(t) for t in sysm.results()]
    tr = {k: host(v) for k, v in sysm.tree().items()}
    info = sysm.info()
    sysm.close()
    return pot, res, tr, info


@pytest.mark.parametrize("group_size", [64, 32])
def test_batch_equals_systems_alone(ctx, group_size):
    systems = _systems()
    thetas = np.array([0.4, 0.5, 0.7, 0.6] * 4, np.float32)[: len(systems)]
    Gs = np.linspace(0.5, 2.0, len(systems)).astype(np.float32)
    b = R.Batch(ctx, sum(s.n for s in systems), len(systems))
    b.load(systems)
    prm = R.default_params(group_size=group_size)
    pots = b.solve(prm, theta=thetas, G=Gs)
    off = b.offsets()
    (x, y, z, m, h, p, ax, ay, az), order = b.results()
    arrs = [host(t) for t in (p, ax, ay, az)]
    order = host(order)
    tr = {k: host(v) for k, v in b.tree().items()}
    for s, sysd in enumerate(systems):
        prm_s = R.default_params(theta=float(thetas[s]), G=float(Gs[s]), group_size=group_size)
        pot, res, t1, info = _alone(ctx, sysd, prm_s)
        p0, p1 = int(off["P"][s]), int(off["P"][s + 1])
        l0, l1 = int(off["LO"][s]), int(off["LO"][s + 1])
        n0, n1 = int(off["NO"][s]), int(off["NO"][s + 1])
        q0 = int(off["PO"][s])
        L, nn = info["numLeaves"], info["numNodes"]
        assert l1 - l0 == L + 1 and n1 - n0 == nn, s
        assert np.array_equal(tr["keys"][p0:p1], t1["keys"]), s
        assert np.array_equal(tr["leaves"][l0:l1], t1["leaves"]), s
        assert np.array_equal(tr["layout"][l0:l1], t1["layout"]), s
        assert np.array_equal(tr["counts"][l0:l1 - 1], t1["counts"]), s
        for k in ("prefixes", "childOffsets", "internalToLeaf", "leafToInternal"):
            assert np.array_equal(tr[k][n0:n1], t1[k]), (s, k)
        assert np.array_equal(tr["parents"][q0:q0 + t1["parents"].size], t1["parents"]), s
        assert np.array_equal(tr["levelRange"][23 * s:23 * s + 23], t1["levelRange"]), s
        for k, w in (("centers", 4), ("multipoles", 8), ("centers64", 4), ("multipoles64", 8)):
            assert np.array_equal(tr[k][w * n0:w * n1], t1[k]), (s, k)
        assert np.array_equal(order[p0:p1] - p0, res[4]), s
        for k in range(4):
            assert np.array_equal(arrs[k][p0:p1], res[k]), (s, k)
        assert abs(pots[s] - pot) <= 1e-12 * abs(pot), (s, pots[s], pot)
    b.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_oracle(ctx, graph):
    """every system of a batch (eager and captured-graph solves) against the oracle pipeline on that system alone:
    sorted keys, permutation, leaves, counts, layout and the linked octree bit for bit; accelerations at p99 <= 1e-5
    of the oracle traversal with 64-target groups at the system's theta and G; the total potential to 1e-5"""
    from gpu_util import oracle_tree, rel_err

    import oracle as O

    systems = _systems()
    thetas = np.array([0.4, 0.5, 0.7, 0.6] * 4, np.float32)[: len(systems)]
    Gs = np.linspace(0.5, 2.0, len(systems)).astype(np.float32)
    b = R.Batch(ctx, sum(s.n for s in systems), len(systems))
    b.load(systems)
    pots = b.solve(R.default_params(group_size=64), theta=thetas, G=Gs, graph=graph)
    if graph:  # a replay of the captured graph
        pots = b.solve(R.default_params(group_size=64), theta=thetas, G=Gs, graph=True)
    off = b.offsets()
    (x, y, z, m, h, p, ax, ay, az), order = b.results()
    ax, ay, az, order = host(ax), host(ay), host(az), host(order)
    tr = {k: host(v) for k, v in b.tree().items()}
    for s, sysd in enumerate(systems):
        t = oracle_tree(sysd, inv_theta=1.0 / float(thetas[s]))
        p0, p1 = int(off["P"][s]), int(off["P"][s + 1])
        l0, l1 = int(off["LO"][s]), int(off["LO"][s + 1])
        n0, n1 = int(off["NO"][s]), int(off["NO"][s + 1])
        assert np.array_equal(tr["keys"][p0:p1].view(np.uint64), t["sorted_keys"]), s
        assert np.array_equal(order[p0:p1] - p0, t["perm"]), s
        assert np.array_equal(tr["leaves"][l0:l1].view(np.uint64), t["leaves"]), s
        assert np.array_equal(tr["layout"][l0:l1].view(np.uint32), t["layout"]), s
        assert np.array_equal(tr["counts"][l0:l1 - 1].view(np.uint32), t["counts"]), s
        nn = n1 - n0
        assert np.array_equal(tr["prefixes"][n0:n1].view(np.uint64), t["tree"]["prefixes"]), s
        for k in ("childOffsets", "internalToLeaf", "leafToInternal"):
            assert np.array_equal(tr[k][n0:n1], t["tree"][k][:nn]), (s, k)
        assert np.array_equal(tr["levelRange"][23 * s:23 * s + 23], t["tree"]["levelRange"]), s
        if sysd.n < 2:
            continue
        otot, _, oax, oay, oaz, _ = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], t["x"], t["y"],
                                                      t["z"], t["h"], t["m"], sysd.limits, G=float(Gs[s]),
                                                      group_size=64)
        e = np.sort(rel_err(ax[p0:p1], ay[p0:p1], az[p0:p1], oax, oay, oaz))
        assert e[int(0.99 * (e.size - 1))] <= 1e-5, (s, sysd.n, e[int(0.99 * (e.size - 1))])
        assert abs(pots[s] - otot) <= 1e-5 * abs(otot), (s, pots[s], otot)
    b.close()


def _mixed_periodic():
    """periodic cubes of several sizes and offsets with open systems between them"""
    return [W.uniform_cube(1000, seed=201, lo=0.0, hi=1.0, periodic=True), W.plummer(3000, seed=202),
            W.uniform_cube(4097, seed=203, lo=-0.5, hi=0.5, periodic=True),
            W.uniform_cube(63, seed=204, lo=2.0, hi=4.0, periodic=True), W.uniform_cube(9000, seed=205),
            W.uniform_cube(20000, seed=206, lo=0.0, hi=1.0, periodic=True)]


@pytest.mark.parametrize("ewald", [0, 1, 2])
def test_periodic_batch_equals_systems_alone(ctx, ewald):
    """periodic systems in a batch: images of each system's own box (1 shell) and, with ewald 1 or 2, the FP64
    Ewald correction from its root multipole; every system bit-identical to the same system solved alone as a
    resident system (periodic or open), the potentials to 1e-12"""
    systems = _mixed_periodic()
    thetas = np.array([0.5, 0.6, 0.4, 0.5, 0.7, 0.5], np.float32)
    Gs = np.array([1.0, 0.5, 2.0, 1.0, 1.5, 1.0], np.float32)
    b = R.Batch(ctx, sum(s.n for s in systems), len(systems))
    b.load(systems)
    prm = R.default_params(num_shells=1, ewald=ewald)
    pots = b.solve(prm, theta=thetas, G=Gs)
    off = b.offsets()
    (x, y, z, m, h, p, ax, ay, az), order = b.results()
    arrs = [host(t) for t in (p, ax, ay, az)]
    order = host(order)
    for s, sysd in enumerate(systems):
        print(f"system {s}: n={sysd.n} periodic={sysd.periodic} ewald={ewald}", flush=True)
        sysm = R.System(ctx, sysd.n)
        dv = [dev(a) for a in sysd.arrays()]
        sysm.load(*dv, sysd.limits, periodic=sysd.periodic)
        prm_s = R.default_params(theta=float(thetas[s]), G=float(Gs[s]), num_shells=1, ewald=ewald)
        sysm.build(prm_s)
        pot = sysm.gravity(prm_s)
        res = [host(t) for t in sysm.results()]
        sysm.close()
        p0, p1 = int(off["P"][s]), int(off["P"][s + 1])
        assert np.array_equal(order[p0:p1] - p0, res[4]), s
        for k in range(4):
            assert np.array_equal(arrs[k][p0:p1].view(np.int32), res[k].view(np.int32)), (s, k)
        assert abs(pots[s] - pot) <= 1e-12 * abs(pot), (s, pots[s], pot)
    with pytest.raises(R.RyoError):
        b.solve(prm, theta=thetas, G=Gs, graph=True)  # the captured graph takes open systems only
    # the second solve of the batch runs for capacities with one read at the end: the same arrays and potentials
    first = _batch_outputs(b)
    pots2 = b.solve(prm, theta=thetas, G=Gs)
    _assert_same(first, _batch_outputs(b))
    assert np.array_equal(pots, pots2)
    b.close()


def test_periodic_batch_matches_oracle(ctx):
    """the periodic systems of a mixed batch against the oracle traversal over the 27 images plus the oracle's
    FP64 Ewald correction from the oracle's FP64 root multipole (relative to the RMS acceleration: the periodic
    net force is a small residual of the image sum and the correction); open systems at p99 <= 1e-5"""
    from gpu_util import oracle_tree, rel_err

    import oracle as O

    systems = _mixed_periodic()
    b = R.Batch(ctx, sum(s.n for s in systems), len(systems))
    b.load(systems)
    pots = b.solve(R.default_params(num_shells=1, ewald=1))
    off = b.offsets()
    (x, y, z, m, h, p, ax, ay, az), order = b.results()
    ax, ay, az, order = host(ax), host(ay), host(az), host(order)
    for s, sysd in enumerate(systems):
        t = oracle_tree(sysd)
        p0, p1 = int(off["P"][s]), int(off["P"][s + 1])
        assert np.array_equal(order[p0:p1] - p0, t["perm"]), s
        args = (t["x"], t["y"], t["z"], t["h"], t["m"])
        otot, _, oax, oay, oaz, _ = O.compute_gravity(t["leaves"], t["layout"], t["centers"], t["mp"], *args,
                                                      sysd.limits, group_size=64, num_shells=1 if sysd.periodic else 0)
        if not sysd.periodic:
            e = np.sort(rel_err(ax[p0:p1], ay[p0:p1], az[p0:p1], oax, oay, oaz))
            assert e[int(0.99 * (e.size - 1))] <= 1e-5, (s, e[int(0.99 * (e.size - 1))])
            continue
        lim = np.asarray(sysd.limits, np.float64)
        c64, q64 = O.upsweep(t["x"], t["y"], t["z"], t["m"], t["leaves"], t["layout"], lim, 2.0, precision="f64")
        etot, _, eax, eay, eaz = O.ewald(q64[0], c64[0, :3], t["x"], t["y"], t["z"], t["m"], float(lim[1] - lim[0]))
        rx, ry, rz = oax + eax, oay + eay, oaz + eaz
        rms = np.sqrt(np.mean(oax ** 2 + oay ** 2 + oaz ** 2))
        d = np.sqrt((ax[p0:p1] - rx) ** 2 + (ay[p0:p1] - ry) ** 2 + (az[p0:p1] - rz) ** 2) / rms
        assert np.quantile(d, 0.99) <= 1e-4, (s, np.quantile(d, 0.99))
        assert abs(pots[s] - (otot + etot)) <= 1e-5 * (abs(otot) + abs(etot)), (s, pots[s], otot + etot)
    b.close()


def test_batch_reuse_and_capacity(ctx):
    """a batch object serves batches of different composition; loading beyond its capacity is an error"""
    b = R.Batch(ctx, 30000, 8)
    a = [W.plummer(5000, seed=1), W.uniform_cube(7000, seed=2)]
    c = [W.uniform_cube(100, seed=3), W.plummer(20000, seed=4), W.uniform_box(900, seed=5)]
    prm = R.default_params()
    b.load(a)
    pa = b.solve(prm)
    b.load(c)
    pc = b.solve(prm)
    b.load(a)
    pa2 = b.solve(prm)
    assert np.array_equal(pa, pa2) and pc.size == 3
    with pytest.raises(R.RyoError):
        b.load([W.plummer(40000, seed=6)])
    b.close()


def test_autotuner_picks_a_candidate(ctx):
    from ryoanji_amd import autotune as A

    systems = [W.plummer(2048, seed=40 + i) for i in range(16)]
    tuner = A.Autotuner(ctx, candidates=[(64, 6, 8), (32, 6, 8), (64, 4, 1), (64, 12, 8, 512, 2)], reps=1)
    best, timings = tuner.tune(systems)
    assert best in [c for c, _ in timings] and len(timings) == 4
    assert tuner.config_for(systems) == best
    # the tuned configuration changes the launch, not the results of its group size
    prm = tuner.apply(best)
    b = R.Batch(ctx, sum(s.n for s in systems), len(systems))
    b.load(systems)
    pots = b.solve(prm, graph=True)
    tuner.reset()
    pots2 = b.solve(prm)
    assert np.array_equal(pots, pots2)
    b.close()


def test_autotuner_coordinate_descent(ctx):
    """the default search visits every value of every dimension once and keeps a configuration it timed"""
    from ryoanji_amd import autotune as A

    systems = [W.plummer(1024, seed=60 + i) for i in range(8)]
    tuner = A.Autotuner(ctx, reps=1)
    best, timings = tuner.tune(systems)
    seen = {A._key(c) for c, _ in timings}
    for dim, values in A.DIMENSIONS.items():
        assert {c[dim] for c, _ in timings} == set(values), dim
    assert A._key(best) in seen and len(seen) <= sum(len(v) for v in A.DIMENSIONS.values())


@pytest.mark.parametrize("waves", [2, 8])
def test_block_size_keeps_results(ctx, waves):
    """2 or 8 waves per traversal block: the same results as the default 4, single system and batch"""
    s = W.plummer(30000, seed=12)
    prm = R.default_params()

    def solve_system():
        sysm = R.System(ctx, s.n)
        sysm.load(*(dev(a) for a in s.arrays()), s.limits)
        sysm.build(prm)
        pot = sysm.gravity(prm)
        out = [host(t) for t in sysm.results()]
        sysm.close()
        return pot, out

    systems = [W.plummer(5000, seed=13 + i) for i in range(5)]
    b = R.Batch(ctx, sum(q.n for q in systems), len(systems))
    b.load(systems)
    ref_sys = solve_system()
    ref_batch = b.solve(prm)
    ref_out = _batch_outputs(b)
    try:
        ctx.call("ryo_ctx_set_traversal_waves", waves)
        got = solve_system()
        assert got[0] == ref_sys[0]
        for u, v in zip(got[1], ref_sys[1]):
            assert np.array_equal(u, v)
        assert np.array_equal(b.solve(prm, graph=True), ref_batch)
        _assert_same(ref_out, _batch_outputs(b))
    finally:
        ctx.call("ryo_ctx_set_traversal_waves", 0)
    b.close()


def _batch_outputs(b):
    (x, y, z, m, h, p, ax, ay, az), order = b.results()
    out = {"arrays": [host(t) for t in (x, y, z, m, h, p, ax, ay, az)], "order": host(order)}
    out.update({k: host(v) for k, v in b.tree().items()})
    out.update({"off_" + k: v.copy() for k, v in b.offsets().items()})
    return out


def _assert_same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        if k == "arrays":
            for i, (u, v) in enumerate(zip(a[k], b[k])):
                assert np.array_equal(u.view(np.int32), v.view(np.int32)), (k, i)
        else:
            assert np.array_equal(a[k], b[k]), k


def _clumped(sizes, seed):
    """systems whose particles sit in two clumps 1e-5 wide at opposite corners: deep trees, many leaf slots"""
    out = []
    for i, n in enumerate(sizes):
        u = [W.uniform01(W._stream(seed + i, d), 0, n) for d in range(3)]
        side = np.where(np.arange(n) % 2 == 0, -0.9, 0.9)
        xyz = [(side + 1e-5 * c).astype(np.float32) for c in u]
        out.append(W.System(*xyz, np.full(n, 1.0 / n, np.float32), np.full(n, 1e-3, np.float32),
                            [-1, 1, -1, 1, -1, 1]))
    return out


def test_graph_solve_equals_eager(ctx):
    """the captured pipeline gives the eager solve's arrays bit for bit, on capture, on replay, and on a replay
    after new positions of the same counts were loaded"""
    systems = _systems()
    b = R.Batch(ctx, sum(s.n for s in systems), len(systems))
    prm = R.default_params()
    thetas = np.array([0.4, 0.5, 0.7, 0.6] * 4, np.float32)[: len(systems)]
    b.load(systems)
    pe = b.solve(prm, theta=thetas)
    eager = _batch_outputs(b)
    pg = b.solve(prm, theta=thetas, graph=True)  # exact solve, capture, replay
    _assert_same(eager, _batch_outputs(b))
    assert np.array_equal(pe, pg)
    pg2 = b.solve(prm, theta=thetas, graph=True)  # replay
    _assert_same(eager, _batch_outputs(b))
    assert np.array_equal(pe, pg2)
    st = b.graph_stats()
    assert st["captures"] == 1 and st["replays"] == 2 and st["fallbacks"] == 0, st
    # new positions, same counts: the graph is replayed with them (and another theta)
    moved = [W.System(s.x * np.float32(0.5), s.y, s.z, s.m, s.h, s.limits) for s in systems]
    b.load(moved)
    pg3 = b.solve(prm, theta=thetas[::-1].copy(), graph=True)
    graph3 = _batch_outputs(b)
    assert b.graph_stats()["captures"] == 1
    pe3 = b.solve(prm, theta=thetas[::-1].copy())
    _assert_same(_batch_outputs(b), graph3)
    assert np.array_equal(pe3, pg3)
    b.close()


def test_graph_solve_outgrowing_its_arrays(ctx):
    """trees that outgrow the captured arrays are detected on the device and solved exactly; the next call
    captures again for the grown arrays"""
    sizes = [4096] * 12
    flat = [W.uniform_cube(n, seed=70 + i) for i, n in enumerate(sizes)]
    deep = _clumped(sizes, 90)
    b = R.Batch(ctx, sum(sizes), len(sizes))
    prm = R.default_params()
    b.load(flat)
    b.solve(prm, graph=True)
    cap = b.graph_stats()["capacity"]
    b.load(deep)
    pg = b.solve(prm, graph=True)  # replayed, outgrows the arrays, solved exactly
    st = b.graph_stats()
    assert st["fallbacks"] == 1 and st["captures"] == 1 and st["replays"] == 2, st
    fallback = _batch_outputs(b)
    pe = b.solve(prm)
    eager = _batch_outputs(b)
    assert int(eager["off_LO"][-1]) > cap  # the deep trees need more leaf slots than the graph holds
    _assert_same(eager, fallback)
    assert np.array_equal(pe, pg)
    pg2 = b.solve(prm, graph=True)  # captured again for the grown arrays, then replayed
    _assert_same(eager, _batch_outputs(b))
    assert np.array_equal(pe, pg2)
    st = b.graph_stats()
    assert st["captures"] == 2 and st["fallbacks"] == 1, st
    b.close()


def _fresh_exact(ctx, systems, prm, thetas=None):
    """the first solve of a new batch object: node arrays sized exactly from the read-back leaf-slot offsets"""
    b = R.Batch(ctx, sum(s.n for s in systems), len(systems))
    b.load(systems)
    pot = b.solve(prm, theta=thetas)
    out = _batch_outputs(b)
    b.close()
    return pot, out


def test_sized_solve_equals_exact(ctx):
    """after its first solve a batch launches every stage for capacities and reads back once at the end: new
    positions, fewer systems, then trees that outgrow the capacities (solved again exactly, capacities grown), each
    bit for bit the exact solve of a fresh batch"""
    systems = _systems()
    prm = R.default_params()
    thetas = np.array([0.4, 0.5, 0.7, 0.6] * 4, np.float32)[: len(systems)]
    sizes = [4096] * 12
    flat = [W.uniform_cube(n, seed=70 + i) for i, n in enumerate(sizes)]
    deep = _clumped(sizes, 90)
    b = R.Batch(ctx, max(sum(s.n for s in systems), sum(sizes)), max(len(systems), len(sizes)))
    b.load(systems)
    b.solve(prm, theta=thetas)  # exact; sets the capacities
    cap0 = b.graph_stats()["capacity"]
    assert cap0 > 0
    moved = [W.System(s.x * np.float32(0.5), s.y, s.z, s.m, s.h, s.limits) for s in systems]
    for batch, th in ((moved, thetas[::-1].copy()), (systems[:3], thetas[:3].copy())):
        b.load(batch)
        pot = b.solve(prm, theta=th)  # for the capacities, one read at the end
        pe, ref = _fresh_exact(ctx, batch, prm, th)
        _assert_same(ref, _batch_outputs(b))
        assert np.array_equal(pot, pe)
    b.load(flat)
    b.solve(prm)
    cap1 = b.graph_stats()["capacity"]
    b.load(deep)
    pot = b.solve(prm)  # outgrows the capacities: solved again exactly
    pe, ref = _fresh_exact(ctx, deep, prm)
    assert int(ref["off_LO"][-1]) > cap1
    _assert_same(ref, _batch_outputs(b))
    assert np.array_equal(pot, pe)
    assert b.graph_stats()["capacity"] > int(ref["off_LO"][-1])  # grown for the next solves
    pot = b.solve(prm)  # within the grown capacities
    _assert_same(ref, _batch_outputs(b))
    assert np.array_equal(pot, pe)
    b.close()


def test_shallow_lifo_retries_at_full_depth(ctx):
    """with a LIFO too shallow for some groups, those groups are traversed again at full depth and every
    result equals the full-depth solve, in both ways of solving"""
    systems = [W.plummer(20000, seed=7), W.uniform_cube(9000, seed=8), W.plummer(3000, seed=9)]
    b = R.Batch(ctx, sum(s.n for s in systems), len(systems))
    prm = R.default_params()
    b.load(systems)
    pe = b.solve(prm)
    full = _batch_outputs(b)
    depth = ctx.stats()["maxStack"]  # the deepest LIFO any group needed
    assert depth > 18
    try:
        ctx.call("ryo_ctx_set_traversal_shape", int(depth // 2), 0, 0)
        ps = b.solve(prm)
        assert R.lib().ryo_traversal_retries(ctx.h) > 0
        _assert_same(full, _batch_outputs(b))
        assert np.array_equal(pe, ps)
        pg = b.solve(prm, graph=True)
        assert R.lib().ryo_traversal_retries(ctx.h) > 0
        _assert_same(full, _batch_outputs(b))
        assert np.array_equal(pe, pg)
    finally:
        ctx.call("ryo_ctx_set_traversal_shape", 0, 0, 0)
    b.close()
