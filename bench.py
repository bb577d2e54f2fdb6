"""Benchmark of the single-GPU Barnes-Hut pass on MI355X.

One step = one full pass over one particle system already resident in HBM:
SFC keys -> stable radix sort -> permute -> cornerstone leaves -> linked octree -> FP64 upsweep (centers, P2M,
M2M, MAC) -> fixed 64-target groups -> traversal (M2P + P2P), plus the FP64 Ewald correction for periodic
workloads.

Default workload (the headline configuration): N = 2^26 particles uniform in a unit periodic cube, equal masses
1/N, fp32, softening 1e-3, theta 0.5, Cartesian quadrupoles, bucket 64, traversal over 1 replica shell (27
images) and the Ewald correction (1 replica shell, real-space cut 2.6 L, k-space cut 2.8, alpha 2/L) in FP64.
Beside it, the same particles with open boundaries at theta 0.5 and 0.4 (`open_boundary`).

Prints ONE JSON line. value = N / (mean wall time of a pass) in particles/s. The roofline object is for the
dominant kernel (the traversal): algorithmic flops per launch (23 per P2P + 128 per M2P, the count of the
reference's README) over the kernel's mean duration from HIP events on the library's stream, against the FP32
vector peak. `ewald_roofline` prices the Ewald kernel against the FP64 vector peak, `build_roofline` the
key -> sort -> build -> upsweep phase against the HBM peak.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the serving leg runs 8 slot streams and the batch lane's stream in this process: one hardware queue per stream
# (HIP's default is 4 per process, shared round-robin) keeps a stream's small kernels from queueing behind another
# stream's long ones; read once, when HIP initialises, so it is set before anything touches the GPU
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (AMD published)
PEAK_FP32_UNPACKED_TFLOPS = 78.6  # the same without packed FP32 (256 CUs x 4 SIMDs x 16 lanes x 2 flops x 2.4 GHz)
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector peak (AMD published)
PEAK_HBM_GBS = 8000.0
EWALD_SETTINGS = (1, 2.6, 2.8, 2.0, 3.0e-3)  # numReplicaShells, lCut, hCut, alpha_scale, small_R_scale_factor
METRIC = "particles/s + P2P/M2P TFLOP/s at θ=0.5; served p50/p99 vs plain solves"

WORKLOADS = {
    "uniform26p": dict(kind="uniform_periodic", n=1 << 26, periodic=True),  # the headline (default)
    "uniform26": dict(kind="uniform", n=1 << 26, periodic=False),
    "plummer24": dict(kind="plummer", n=1 << 24, periodic=False),
    "hernquist24": dict(kind="hernquist", n=1 << 24, periodic=False),
}
DEFAULT_WORKLOAD = "uniform26p"


def log(msg):
    """a progress line on stderr (the JSON result is the only line on stdout)"""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def make_system(name, n=None, seed=42):
    from ryoanji_amd import workloads as W

    spec = WORKLOADS[name]
    n = n or spec["n"]
    if spec["kind"] == "plummer":
        return W.plummer(n, seed=seed)
    if spec["kind"] == "uniform":
        return W.uniform_cube(n, seed=seed, lo=-1.0, hi=1.0)
    if spec["kind"] == "uniform_periodic":
        return W.uniform_cube(n, seed=seed, lo=0.0, hi=1.0, periodic=True)
    return W.hernquist_mix(n, seed=seed)


def traffic_from_profiles(kernel="traverseKernel", workload="plummer24"):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if there is one."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


def cpu_model():
    """the host CPU's model name (first processor in /proc/cpuinfo)"""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(s, sample_frac=1 / 128, seed=0, num_shells=0, ewald_settings=None, budget_s=15.0):
    """Reference CPU path (oracle/_ref, the reference's OpenMP code) when it was built, else the restatement:
    keys, stable sort, permutation, leaves, linked tree and upsweep on the full system; the traversal (over the
    replica images when num_shells > 0) on a seeded sample of leaf ranges and, for a periodic system, the Ewald
    correction (ewald_settings) of the same sampled particles, both scaled to all particles."""
    import numpy as np

    import oracle as O

    kind = "reference" if O.have_ref() else "port"
    lib = "ref" if kind == "reference" else "orc"
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    keys = O.sfc_keys(s.x, s.y, s.z, s.limits, lib=lib)
    sk, perm = O.sort_by_key(keys, lib=lib)
    x, y, z, m, h = (a[perm] for a in s.arrays())
    leaves, counts = O.compute_octree(sk, 64, lib=lib)
    layout = O.layout_from_counts(counts)
    c, q = O.upsweep(x, y, z, m, leaves, layout, s.limits, 2.0, lib=lib)
    t_build = time.perf_counter() - t0
    L = leaves.size - 1
    rng = np.random.default_rng(seed)
    nchunks = 16
    span = max(1, int(L * sample_frac / nchunks))
    starts = np.sort(rng.choice(max(1, L - span), nchunks, replace=False))
    t1 = time.perf_counter()
    nsample = 0
    sample_idx = []
    used = 0
    for a in starts:  # chunk by chunk, stopping once the traversal sample has taken `budget_s`
        if kind == "reference":
            O.compute_gravity(leaves, layout, c, q, x, y, z, h, m, s.limits, lib="ref", num_shells=num_shells,
                              leaf_range=(int(a), int(a + span)), want_pot=False)
            nsample += int(layout[a + span] - layout[a])
            sample_idx.append(np.arange(layout[a], layout[a + span]))
        else:
            g0, g1 = int(layout[a]) // 16, (int(layout[a + span]) + 15) // 16
            groups = np.arange(g0, g1, dtype=np.int32)
            O.compute_gravity(leaves, layout, c, q, x, y, z, h, m, s.limits, group_list=groups, want_pot=False,
                              num_shells=num_shells)
            nsample += int(groups.size * 16)
            sample_idx.append(np.arange(g0 * 16, min(g1 * 16, s.n)))
        used += 1
        if time.perf_counter() - t1 > budget_s:
            break
    nchunks = used
    t_trav = time.perf_counter() - t1
    t_ewald, ewald_note = 0.0, ""
    if ewald_settings is not None:
        idx = np.concatenate(sample_idx)
        mroot, center = q[0].astype(np.float64), c[0, :3].astype(np.float64)
        t2 = time.perf_counter()
        if kind == "reference":
            O.ref_ewald(mroot, center, *(a[idx].astype(np.float64) for a in (x, y, z, m)), s.limits,
                        settings=ewald_settings, precision="f64")
        else:
            O.ewald(mroot, center, x[idx], y[idx], z[idx], m[idx], float(s.limits[1] - s.limits[0]),
                    settings=ewald_settings)
        t_ewald = time.perf_counter() - t2
        ewald_note = f", Ewald correction (FP64) of the same {idx.size} particles ({t_ewald:.2f} s)"
    t_full = t_build + (t_trav + t_ewald) * s.n / max(nsample, 1)
    images = f" over {(2 * num_shells + 1) ** 3} images" if num_shells else ""
    return {
        "value": s.n / t_full,
        "unit": "particles/s",
        "cores": threads,
        "cpu_model": cpu_model(),
        "kind": kind,
        "sample": (f"extrapolated: full keys/sort/permute/leaves/tree/upsweep of the same N={s.n} system "
                   f"({t_build:.2f} s), traversal (16-target groups{images}) timed on {nsample} particles in "
                   f"{nchunks} seeded leaf ranges ({t_trav:.2f} s){ewald_note}, scaled by N/{nsample}"),
    }


def ewald_flops(s, center, settings=EWALD_SETTINGS, sample=65536, seed=3):
    """Algorithmic FP64 flops of the Ewald correction per particle (ewald.hpp:201-327), counted from the terms the
    reference evaluates: every real-space image with |r - c + n L| <= lCut L or inside the replica shells costs
    77 flops (image offset 6, |R|^2 5, 1/R and 1/R^2 3, the Gaussian factor 4, erf/erfc 2, gamma_0..3 13,
    quadrupole evaluation and accumulation 44), every k-space vector 20 (h.r 5, cos and sin 2, the two
    combinations 6, accumulation 7), plus 10 per particle; exp, erf, erfc, sqrt, cos and sin count as one flop
    each. The image count is averaged over `sample` seeded particles."""
    import itertools
    import math

    import numpy as np

    nr, lcut, hcut = int(settings[0]), float(settings[1]), float(settings[2])
    L = float(s.limits[1] - s.limits[0])
    ns = max(int(math.ceil(lcut)), nr)
    idx = np.random.default_rng(seed).choice(s.n, min(sample, s.n), replace=False)
    r = np.stack([s.x[idx], s.y[idx], s.z[idx]], axis=1).astype(np.float64) - np.asarray(center, np.float64)
    cut2 = lcut * lcut * L * L
    count = np.zeros(idx.size)
    for ix, iy, iz in itertools.product(range(-ns, ns + 1), repeat=3):
        R = r + np.array([ix, iy, iz], np.float64) * L
        inner = max(abs(ix), abs(iy), abs(iz)) <= nr
        count += (np.einsum("ij,ij->i", R, R) <= cut2) | inner
    hr = int(math.ceil(hcut))
    nk = sum(1 for hx, hy, hz in itertools.product(range(-hr, hr + 1), repeat=3)
             if 0 < hx * hx + hy * hy + hz * hz <= hcut * hcut)
    images = float(count.mean())
    return 77.0 * images + 20.0 * nk + 10.0, images, nk


def build_roofline(phase, n, nodes):
    """the key -> sort -> build -> upsweep phase against the HBM peak: algorithmic bytes B_K = 112 N + 320 n (keys
    20, sort pairs 24, permute of x, y, z, m, h 44, leaf counting 8, centers + leaf P2M 16 per particle; about
    320 per tree node) over the phase time from keys through the upsweep (HIP events per phase)"""
    import numpy as np

    keys = ("keys", "sort", "permute", "leaves", "link", "upsweep")
    t = sum(float(np.mean([ph[k] for ph in phase])) for k in keys) / 1e3
    b = 112.0 * n + 320.0 * nodes
    gbs = b / t / 1e9
    return {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
            "algorithmic_bytes": b, "phase_ms": t * 1e3, "phases": list(keys)}


def accuracy(ctx, sysm, s, order, ax, ay, az, pot, num_targets=16384, seed=1, box_l=None, shells=0, ewald_l=None,
             p=None):
    """1st-percentile and max relative acceleration error vs an FP64 direct sum on sampled targets (with the same
    periodic images as the traversal when box_l is given; the FP64 Ewald pair sum over all images when ewald_l,
    the cubic box length, is given), plus the errors relative to the RMS acceleration and, with the per-particle
    potentials p, the relative error of the sampled potential sum."""
    import numpy as np
    import torch

    import ryoanji_amd as R

    arrs = [torch.from_numpy(a[order]).cuda() for a in (s.x, s.y, s.z, s.h, s.m)]
    tgt = np.sort(np.random.default_rng(seed).choice(s.n, min(num_targets, s.n), replace=False)).astype(np.int64)
    out = torch.zeros(tgt.size * 4, dtype=torch.float64, device="cuda")
    dtgt = torch.from_numpy(tgt).cuda()  # bound to a name while the pointer is in use
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if ewald_l is not None:
        ctx.call("ryo_ewald_pair_sum_f64", *(R.ptr(a) for a in arrs), s.n, R.ptr(dtgt), tgt.size, 1.0,
                 float(ewald_l), 0.0, R.ptr(out))
    else:
        bl = None if box_l is None else R.host_f64(box_l)
        ctx.call("ryo_direct_sum_f64", *(R.ptr(a) for a in arrs), s.n, R.ptr(dtgt), tgt.size, 1.0, bl, shells,
                 R.ptr(out))
    ref = out.cpu().numpy().reshape(-1, 4)
    t_ref = time.perf_counter() - t0
    d = np.sqrt((ax[tgt] - ref[:, 1]) ** 2 + (ay[tgt] - ref[:, 2]) ** 2 + (az[tgt] - ref[:, 3]) ** 2)
    norm = np.sqrt(ref[:, 1] ** 2 + ref[:, 2] ** 2 + ref[:, 3] ** 2)
    e = np.sort(d / norm)
    rms = float(np.sqrt(np.mean(norm ** 2)))
    er = np.sort(d / rms)
    res = {"targets": int(tgt.size), "images": "all (Ewald)" if ewald_l is not None else (2 * shells + 1) ** 3,
           "err_p1": float(e[int(0.99 * e.size)]), "err_max": float(e[-1]), "err_median": float(e[e.size // 2]),
           "err_rms_p1": float(er[int(0.99 * er.size)]), "err_rms_max": float(er[-1]), "reference_ms": t_ref * 1e3}
    if p is not None:
        res["pot_rel_err"] = float(abs(p[tgt].astype(np.float64).sum() - ref[:, 0].sum()) / abs(ref[:, 0].sum()))
    return res


def serving_bench(num=32, slots=8, seed=5):
    """Served mixed-size stream (Plummer, N log-uniform in [2^12, 2^22], theta in {0.4,0.5,0.7}, 1-20 leapfrog
    steps): continuous batching on `slots` slots plus the batch lane (requests of at most 2^16 particles packed
    into one batched solve per step; state resident on the device) against one-at-a-time plain solves of the same
    stream (every timestep a full pass from host memory, host kick-drift-kick), against the resident server
    restricted to one slot, and against the same slots without the lane. Throughput is measured with all requests
    queued at t=0; latency percentiles at Poisson arrival rates around the one-slot server's capacity."""
    import numpy as np

    from ryoanji_amd import serving as S

    systems, params = S.make_stream(num, rate_per_s=0, seed=seed)
    max_n = int(max(s.n for s in systems))
    out = {"requests": num, "slots": slots, "n_range": [int(min(s.n for s in systems)), max_n]}
    one = S.Server(1, max_n, batch_max_n=0)
    res1 = one.run(systems, params)
    one.close()
    slots_only = S.Server(slots, max_n, hbm_budget=64 << 30, batch_max_n=0)
    resN = slots_only.run(systems, params)
    slots_only.close()
    mixed = S.Server(slots, max_n, hbm_budget=64 << 30)
    resS = mixed.run(systems, params)
    s1, sN, sS = (S.summarize(systems, params, r) for r in (res1, resN, resS))
    sP = S.summarize(systems, params, S.plain_solves(systems, params))
    lane = [r for r, sd in zip(resS, systems) if r["slot"] == slots]
    out["plain_saturated"] = sP
    out["resident_one_slot_saturated"] = s1
    out["resident_slots_no_lane_saturated"] = sN
    out["batched_saturated"] = sS
    out["lane_requests"] = len(lane)
    small = [i for i, r in enumerate(resS) if r["slot"] == slots]
    if small:
        out["lane_request_p50_ms"] = float(np.percentile([resS[i]["finish"] - resS[i]["arrival"] for i in small], 50))
        out["no_lane_same_requests_p50_ms"] = float(np.percentile([resN[i]["finish"] - resN[i]["arrival"]
                                                                   for i in small], 50))
    out["throughput_ratio"] = sS["particle_solves_per_s"] / sP["particle_solves_per_s"]
    out["throughput_ratio_vs_resident_one_slot"] = sS["particle_solves_per_s"] / s1["particle_solves_per_s"]
    # precision report of the served results: every request's final accelerations against the FP64 direct sum
    # (256 sampled targets each; a separate run, so the timings above carry no checking work)
    racc = mixed.run(systems, params, accuracy_targets=256)
    worst = {th: max(r["err_p99"] for r, t in zip(racc, params["theta"]) if t == th) for th in set(params["theta"])}
    out["served_accuracy"] = {"targets_per_request": 256,
                              "max_err_p99_by_theta": {f"{k:.1f}": v for k, v in sorted(worst.items())},
                              "max_err_median": max(r["err_median"] for r in racc)}
    # the same server admitting least work first (a request that has waited 2 s goes first)
    least = S.Server(slots, max_n, hbm_budget=64 << 30, order="least_work", max_wait_ms=2000.0)
    out["batched_saturated_least_work"] = S.summarize(systems, params, least.run(systems, params))
    # latency under Poisson arrivals at fractions of the one-slot server's capacity
    rates = []
    cap_rps = s1["requests_per_s"]
    keys = ("p50_ms", "p99_ms", "requests_per_s")
    for frac in (0.5, 2.0):
        rate = frac * cap_rps
        _, p = S.make_stream(num, rate_per_s=rate, seed=seed)
        params["arrival_ms"] = p["arrival_ms"]
        r = mixed.run(systems, params)
        rates.append(dict(rate_per_s=rate, **{k: v for k, v in S.summarize(systems, params, r).items() if k in keys}))
        sl = S.summarize(systems, params, least.run(systems, params))
        rates[-1]["least_work_p50_ms"], rates[-1]["least_work_p99_ms"] = sl["p50_ms"], sl["p99_ms"]
        r1 = S.Server(1, max_n, batch_max_n=0)
        rr = r1.run(systems, params)
        r1.close()
        rates[-1]["one_slot_p50_ms"] = S.summarize(systems, params, rr)["p50_ms"]
        rates[-1]["one_slot_p99_ms"] = S.summarize(systems, params, rr)["p99_ms"]
    mixed.close()
    least.close()
    out["poisson"] = rates
    return out


def batching_bench(classes=(12, 14, 16), total_exp=22, seed=9):
    """Many small Plummer systems (2^total_exp particles in all, systems of 2^c particles for each size class c):
    one batched solve (one launch sequence for all systems) against solving them one at a time on a resident
    system. The batched solve runs eagerly (one host round trip sizes the node arrays) and as one captured HIP
    graph; the graph's launch configuration (group size, traversal block size, waves per CU, XCD ranges, LIFO
    depth) is tuned per size class by the autotuner. Inputs are resident on the device for all of them."""
    import numpy as np
    import torch

    import ryoanji_amd as R
    from ryoanji_amd import autotune as A
    from ryoanji_amd import workloads as W

    ctx = R.Context()
    tuner = A.Autotuner(ctx)
    rows = []
    for c in classes:
        n, K = 1 << c, 1 << (total_exp - c)
        systems = [W.plummer(n, seed=seed * 100000 + i) for i in range(K)]
        prm = R.default_params()
        sysm = R.System(ctx, n)
        devs = [[torch.from_numpy(a).cuda() for a in s.arrays()] for s in systems]
        for rep in range(2):  # the first round warms up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for s, d in zip(systems, devs):
                sysm.load(*d, s.limits)
                sysm.build(prm)
                sysm.gravity(prm)
            torch.cuda.synchronize()
            t_one = time.perf_counter() - t0
        sysm.close()
        del devs
        b = R.Batch(ctx, n * K, K)
        b.load(systems)
        t_eager = A.time_solve(b, prm, graph=False)
        t_default = A.time_solve(b, prm, graph=True)
        best, timings = tuner.tune(systems, b)
        t_best = min(t for _, t in timings)
        graph_stats = b.graph_stats()
        b.close()
        rows.append(dict(size_class=f"2^{c}", systems=K, particles=n * K,
                         one_at_a_time_ms=t_one * 1e3, one_system_pass_ms=t_one * 1e3 / K,
                         batched_eager_ms=t_eager * 1e3,
                         batched_graph_default_ms=t_default * 1e3, batched_graph_tuned_ms=t_best * 1e3,
                         tuned={**best, "lifo": best["lifo"] or 22 * 64}, configurations_timed=len(timings),
                         graph_stats=graph_stats,
                         particles_per_s_one_at_a_time=n * K / t_one, particles_per_s_batched=n * K / t_best,
                         speedup=t_one / t_best))
    ctx.close()
    return rows


def trav_summary(s, stats, t_trav_s):
    """traversal rates of one pass: TFLOP/s by the 23/128 count (north star, README) and by the 23/65 count (the
    quadrupole actually computed), particles/s of the traversal alone, interactions per particle"""
    flops_ns = 23.0 * stats["sumP2P"] + 128.0 * stats["sumM2P"]
    flops_code = 23.0 * stats["sumP2P"] + 65.0 * stats["sumM2P"]
    return {
        "traversal_ms": t_trav_s * 1e3,
        "tflops": flops_ns / t_trav_s / 1e12,
        "frac_of_fp32_peak": flops_ns / t_trav_s / 1e12 / PEAK_FP32_TFLOPS,
        "tflops_code_count": flops_code / t_trav_s / 1e12,
        "flops_per_launch": flops_ns,
        "particles_per_s_traversal": s.n / t_trav_s,
        "interactions_per_particle": {"p2p": stats["sumP2P"] / s.n, "m2p": stats["sumM2P"] / s.n,
                                      "max_stack": stats["maxStack"]},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=list(WORKLOADS))
    ap.add_argument("--n", type=int, default=None, help="override the particle count")
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--multipole", default="quadrupole", choices=["quadrupole", "mdq"],
                    help="CartesianQuadrupole or CartesianMDQpole expansions (MDQpole: open workloads only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-open", action="store_true", help="skip the open-boundary passes beside a periodic run")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--trav-stack", type=int, default=0, help="traversal LIFO entries per wave (0: default)")
    ap.add_argument("--trav-blocks", type=int, default=0, help="traversal blocks per CU (0: default)")
    ap.add_argument("--trav-ranges", type=int, default=0, help="traversal group ranges (0: one per XCD)")
    ap.add_argument("--build-only", action="store_true",
                    help="time only the key -> sort -> build -> upsweep phase (a profiling run of the build kernels); "
                         "prints the build phases and build_roofline")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run a plain device copy of this many MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU decomposition of the gravity pass; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import ryoanji_amd as R

    torch.cuda.set_device(0)
    if args.copy_calibration_mb:
        # an elementwise kernel that reads and writes exactly copy_calibration_mb MiB each
        src = torch.ones(args.copy_calibration_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    log(f"workload {args.workload}: generating particles")
    s = make_system(args.workload, args.n)
    spec = WORKLOADS[args.workload]
    ctx = R.Context()
    R.lib().ryo_ctx_set_timing(ctx.h, 1)
    ctx.call("ryo_ctx_set_traversal_shape", args.trav_stack, args.trav_blocks, args.trav_ranges)
    sysm = R.System(ctx, s.n)
    dev = [torch.from_numpy(a).cuda() for a in s.arrays()]

    if args.multipole == "mdq" and spec["periodic"]:
        print("the Ewald correction is for quadrupole expansions (as the reference's): use an open workload with "
              "--multipole mdq", file=sys.stderr)
        sys.exit(2)

    def params(theta, periodic):
        prm = R.default_params(theta=theta, group_size=64, bucket_size=64,
                               multipole=1 if args.multipole == "mdq" else 0)
        if periodic:
            prm.num_shells, prm.ewald = 1, 1
            for k, v in enumerate(EWALD_SETTINGS):
                prm.ewald_settings[k] = v
        return prm

    def one_pass(prm, periodic):
        # inputs resident in HBM: reset to the unsorted initial state (device-to-device), then a full pass
        sysm.load(*dev, s.limits, periodic=periodic)
        torch.cuda.synchronize()
        t = time.perf_counter()
        sysm.build(prm)
        pot = sysm.gravity(prm)
        torch.cuda.synchronize()
        return time.perf_counter() - t, pot

    def run(prm, periodic, steps, warmup):
        for _ in range(warmup):
            one_pass(prm, periodic)
        times, phase = [], []
        pot, stats = 0.0, None
        for _ in range(steps):
            dt, pot = one_pass(prm, periodic)
            times.append(dt)
            phase.append(sysm.phase_ms())
            stats = ctx.stats()
        return float(np.mean(times)), phase, stats, pot

    periodic = spec["periodic"]
    prm = params(args.theta, periodic)
    if args.build_only:
        # the build alone, from the unsorted initial state each time (the phase the build roofline prices)
        phase = []
        for i in range(args.warmup + args.steps):
            sysm.load(*dev, s.limits, periodic=periodic)
            sysm.build(prm)
            torch.cuda.synchronize()
            if i >= args.warmup:
                phase.append(sysm.phase_ms())
            # settles the device-sized build's counts as the gravity pass does (the next build then launches for
            # the leaf count it found, as in a full pass)
            sysm.info()
        info = sysm.info()
        ph = {k: float(np.mean([p[k] for p in phase])) for k in ("keys", "sort", "permute", "leaves", "link",
                                                                 "upsweep")}
        log(f"build {sum(ph.values()):.2f} ms: " + ", ".join(f"{k} {v:.3f}" for k, v in ph.items()))
        print(json.dumps({"workload": args.workload, "n": s.n, "steps": args.steps, "build_phases_ms": ph,
                          "build_roofline": build_roofline(phase, s.n, info["numNodes"]), "tree": info}))
        return
    log(f"N={s.n}: {args.warmup} warmup + {args.steps} timed passes")
    t_pass, phase, stats, pot = run(prm, periodic, args.steps, args.warmup)
    log(f"pass {t_pass * 1e3:.1f} ms ({s.n / t_pass / 1e6:.1f} M particles/s)")
    t_trav = float(np.mean([ph["traverse"] for ph in phase])) / 1e3
    info = sysm.info()
    trav = trav_summary(s, stats, t_trav)
    p, ax, ay, az, order = (t.cpu().numpy() for t in sysm.results())

    bc = "periodic, traversal over 1 replica shell (27 images) + FP64 Ewald correction (1 replica shell, lCut 2.6, " \
         "hCut 2.8, alpha 2/L)" if periodic else "open boundaries"
    out = {
        "metric": METRIC,
        "value": s.n / t_pass,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_pass * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: N={s.n} {spec['kind']} particles, theta={args.theta}, cartesian "
                        f"{'MDQpole' if args.multipole == 'mdq' else 'quadrupole'}, bucket 64, softening 1e-3, "
                        f"64-target groups, {bc}",
            "n": s.n, "theta": args.theta, "bucket": 64, "group_size": 64, "periodic": periodic,
            "traversal_shape": {"stack_per_wave": args.trav_stack or "default",
                                "blocks_per_cu": args.trav_blocks or "default",
                                "group_ranges": args.trav_ranges or "default"},
        },
        "roofline": {
            "bound": "valu",
            "achieved": trav["tflops"],
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": trav["frac_of_fp32_peak"],
            "traffic": traffic_from_profiles(workload=args.workload),
            "kernel": "traverseKernel",
            "flops_per_launch": trav["flops_per_launch"],
            "mean_launch_ms": t_trav * 1e3,
        },
        "build_roofline": build_roofline(phase, s.n, info["numNodes"]),
        "tflops_code_count": trav["tflops_code_count"],
        # the traversal by the flops its code executes (23 per P2P, 65 per quadrupole M2P) against the vector peak
        # without packed FP32, which this build compiles out
        "roofline_code_count": {"achieved": trav["tflops_code_count"], "peak": PEAK_FP32_UNPACKED_TFLOPS,
                                "unit": "TFLOP/s", "frac": trav["tflops_code_count"] / PEAK_FP32_UNPACKED_TFLOPS},
        "particles_per_s_traversal": trav["particles_per_s_traversal"],
        "interactions_per_particle": trav["interactions_per_particle"],
        "phases_ms": {k: float(np.mean([ph[k] for ph in phase])) for k in phase[0]},
        "tree": info,
        "total_potential": pot,
    }
    if periodic:
        t_ew = out["phases_ms"]["ewald"] / 1e3
        center = sysm.tree()["centers64"][:3].cpu().numpy()
        per, images, nk = ewald_flops(s, center)
        out["ewald_roofline"] = {"bound": "fp64 valu", "achieved": per * s.n / t_ew / 1e12, "peak": PEAK_FP64_TFLOPS,
                                 "unit": "TFLOP/s", "frac": per * s.n / t_ew / 1e12 / PEAK_FP64_TFLOPS,
                                 "kernel": "ewaldKernel", "mean_launch_ms": t_ew * 1e3,
                                 "flops_per_particle": per, "images_per_particle": images, "k_vectors": nk,
                                 "particles_per_s": s.n / t_ew}
    if not args.no_accuracy:
        log("accuracy against FP64 references")
        if periodic:
            L = float(s.limits[1] - s.limits[0])
            # the full periodic result (traversal over the replica images + Ewald correction) against the FP64
            # Ewald pair sum over all images; its potentials carry the reference's k-space potential sign, so
            # pot_rel_err there is that convention, not an accuracy figure
            out["accuracy_vs_fp64_ewald"] = accuracy(ctx, sysm, s, order, ax, ay, az, pot, num_targets=1024,
                                                     ewald_l=L, p=p)
            # the same solve with the k-space potential consistent with its acceleration (ewald = 2): the
            # accelerations are identical, the potentials follow the exact sum
            prm.ewald = 2
            sysm.gravity(prm)
            p2, ax2, ay2, az2, ord2 = (t.cpu().numpy() for t in sysm.results())
            out["accuracy_vs_fp64_ewald_consistent_potential"] = accuracy(ctx, sysm, s, ord2, ax2, ay2, az2, pot,
                                                                          num_targets=1024, ewald_l=L, p=p2)
            # the traversal over the replica images alone (no Ewald term) against the FP64 direct sum over the
            # same images; the Ewald term is FP64 in both and is reported separately in phases_ms
            prm.ewald = 0
            sysm.gravity(prm)
            p0, tax, tay, taz, tord = (t.cpu().numpy() for t in sysm.results())
            out["accuracy_vs_fp64_direct"] = accuracy(ctx, sysm, s, tord, tax, tay, taz, pot, num_targets=1024,
                                                      box_l=[L, L, L], shells=1, p=p0)
            prm.ewald = 1
        else:
            out["accuracy_vs_fp64_direct"] = accuracy(ctx, sysm, s, order, ax, ay, az, pot, p=p)
    if periodic and not args.no_open:
        # the same particles with open boundaries (the north star's 2^26 uniform throughput and accuracy target),
        # at theta 0.5 and 0.4
        rows = []
        for th in (args.theta, 0.4):
            po = params(th, False)
            log(f"open boundaries, theta {th}")
            t_o, ph_o, st_o, pot_o = run(po, False, max(3, args.steps // 4), 1)
            t_tr = float(np.mean([ph["traverse"] for ph in ph_o])) / 1e3
            row = {"theta": th, "ms_per_step": t_o * 1e3, "particles_per_s": s.n / t_o,
                   "phases_ms": {k: float(np.mean([ph[k] for ph in ph_o])) for k in ph_o[0]},
                   **trav_summary(s, st_o, t_tr), "total_potential": pot_o}
            if not args.no_accuracy:
                po_, oax, oay, oaz, oord = (t.cpu().numpy() for t in sysm.results())
                row["accuracy_vs_fp64_direct"] = accuracy(ctx, sysm, s, oord, oax, oay, oaz, pot_o, p=po_)
            rows.append(row)
        out["open_boundary"] = rows
    if spec["kind"] == "hernquist" and not args.no_accuracy:
        # the precision stress: the same pass with the softening-safe MAC (cells nearer than their radius plus
        # 2 h_max are opened, so no softened pair is approximated by a multipole)
        prm.soft_mac = 1
        sm_t = []
        for _ in range(2):
            dt, _pot = one_pass(prm, periodic)
            sm_t.append(dt)
        sst = ctx.stats()
        p3, ax3, ay3, az3, ord3 = (t.cpu().numpy() for t in sysm.results())
        out["soft_mac"] = {"ms_per_step": float(np.mean(sm_t)) * 1e3, "particles_per_s": s.n / float(np.mean(sm_t)),
                           "interactions_per_particle": {"p2p": sst["sumP2P"] / s.n, "m2p": sst["sumM2P"] / s.n},
                           "accuracy_vs_fp64_direct": accuracy(ctx, sysm, s, ord3, ax3, ay3, az3, _pot, p=p3)}
        prm.soft_mac = 0
    sysm.close()
    del dev
    if not args.no_serving:
        log("serving stream")
        out["serving"] = serving_bench()
    if not args.no_batching:
        log("batches of small systems")
        out["batching"] = batching_bench()
    if not args.no_cpu:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(s, num_shells=1 if periodic else 0,
                                           ewald_settings=EWALD_SETTINGS if periodic else None)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
