#!/usr/bin/env python3
"""Benchmark of the pattern engines on one MI355X.

Headline workload (value): the moldyn irregular reduction -- FCC lattice of
40^3 unit cells (256,000 molecules) at density 0.8442, positions jittered
+-0.05 sigma (seed 42), the fixed half list of all pairs within 2.5 sigma
(~7M pairs, i<j) in seeded random order, LJ eps = sigma = 1, forces scattered
to both endpoints, velocity-Verlet with dt = 0.005, FP32.  One step = one
force evaluation plus the velocity-Verlet update; value = pairs * steps / s
with the handle's default node-ordered rows (edge-ordered rows and the
atomic kernel are reported beside it).  Algorithmic bytes per step:
4 B per CSR entry (2 per pair) + 48 B per node (position read, velocity
read and write, position write, 12 B each; a kick-drift step's force only
feeds its own half kicks and is not stored).

Secondary engines in the same JSON line: the heat3D stencil (128^3 and 384^3
FP32, 100 steps) and, when built, kmeans.  roofline: the dominant kernel's
algorithmic bytes per launch over its average duration, measured with HIP
events around each launch on the stream it runs on.  cpu_baseline: the C
oracle (a port of the reference's CPU engine structure) on this host's cores.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 5]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ipdps15-pattern_amd"))

import numpy as np  # noqa: E402

import pattern_amd as pa  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "stencil Gcell/s, LJ edges/s, kmeans pt-iter/s vs roofline; p50/p99 job latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """The host cores this process may use: its CPU affinity set, capped by a
    cgroup CPU quota when one is set, and the CPU model name."""
    cpus = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, per = fh.read().split()[:2]
            if q != "max":
                quota = max(1, int(-(-int(q) // int(per))))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = min(cpus, quota) if quota else cpus
    return {"threads": threads, "affinity_cpus": cpus, "cgroup_cpu_quota": quota, "nproc": os.cpu_count(),
            "model": model}


def cpu_threads():
    return host_cpu()["threads"]


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    return orc


def rel_err(a, b):
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(float(np.max(np.abs(b))), 1e-300))


LANES_NAME = {1: "csr one lane per row", 4: "csr 4 lanes per row", 8: "csr 8 lanes per row",
              16: "csr 16 lanes per row", 32: "csr 32 lanes per row", 64: "sliced (SELL-64) one lane per row"}
FORM_NAME = {0: "global gathers, 4 ahead", 1: "global gathers, pipelined by 4", 2: "global gathers, pipelined by 8",
             3: "LDS-staged block neighbours, 4 ahead", 4: "LDS-staged block neighbours, 8 ahead"}


def kernel_name(md):
    lanes, form = md.variant()[1], md.sliced_form()
    name = LANES_NAME.get(lanes, "")
    return f"{name} ({FORM_NAME[form]})" if form >= 0 else name


def node_sorted(edges):
    lo, hi = np.minimum(edges[:, 0], edges[:, 1]), np.maximum(edges[:, 0], edges[:, 1])
    return edges[np.lexsort((hi, lo))]


def time_md(md, steps, warmup, n, E):
    """Graph-replayed steps: warm-up steps, then the step graph for `steps`
    captured and replayed once (a graph's first two launches carry one-time
    runtime costs, ~0.2 ms here), then the timed replay: value by the host
    clock around it, the force kernel's average launch time from HIP events
    recorded on the handle's stream around the same replay (one launch per
    step, so this counts the gaps between launches as kernel time and
    understates the kernel's bandwidth).  The same number of steps launched
    one by one with events around each kernel is reported beside it."""
    md.run(warmup)
    md.run(steps)  # captures the graph for this step count
    md.run(steps)
    pa.call("pat_device_sync")
    w0 = time.perf_counter()
    dev_ms = md.run_graph_timed(steps)
    wall = time.perf_counter() - w0
    _, kms = md.run_timed(steps)
    kavg = dev_ms / steps
    alg = 8 * E + 48 * n
    return {"pairs_per_s": E * steps / wall, "ms_per_step": wall * 1e3 / steps, "kernel_us": kavg * 1e3,
            "achieved_gbs": alg / (kavg * 1e-3) / 1e9, "kernel": kernel_name(md), "form": md.sliced_form(),
            "eager_kernel_us": float(np.mean(kms)) * 1e3}


def bench_lj(steps, warmup, do_cpu, forms=()):
    t0 = time.perf_counter()
    pos, box, edges = pa.moldyn_system(40, seed=42)
    n, E = len(pos), len(edges)
    log(f"[lj] moldyn 40^3: {n} molecules, {E} pairs, set-up {time.perf_counter() - t0:.1f}s")
    out = {"pairs": E, "nodes": n, "alg_bytes": 8 * E + 48 * n}
    for name, order in (("node_order", pa.PAT_IR_ORDER_NODE), ("edge_order", pa.PAT_IR_ORDER_EDGE)):
        md = pa.MD(pos, None, edges, box, 2.5, 0.005, order=order)
        r = time_md(md, steps, warmup, n, E)
        md.close()
        out[name] = r
        log(f"[lj] {name}: {r['ms_per_step'] * 1e3:.1f} us/step ({r['pairs_per_s'] / 1e9:.2f} G pairs/s), force "
            f"kernel {r['kernel_us']:.1f} us per launch in the step graph ({r['eager_kernel_us']:.1f} us launched "
            f"one by one) [{r['kernel']}], {r['achieved_gbs']:.0f} GB/s algorithmic")
    for f in forms:  # node order with each requested sliced form in place of the tuner's choice
        md = pa.MD(pos, None, edges, box, 2.5, 0.005, order=pa.PAT_IR_ORDER_NODE)
        md.set_sliced_form(f)
        r = time_md(md, steps, warmup, n, E)
        md.close()
        out.setdefault("forms", {})[str(f)] = r
        log(f"[lj] node_order, form {f}: {r['ms_per_step'] * 1e3:.1f} us/step ({r['pairs_per_s'] / 1e9:.2f} G pairs/s), "
            f"force kernel {r['kernel_us']:.1f} us [{r['kernel']}]")
    mda = pa.MD(pos, None, edges, box, 2.5, 0.005, variant=pa.PAT_IR_ATOMIC)
    mda.run_timed(warmup)
    pa.call("pat_device_sync")
    a0 = time.perf_counter()
    _, akms = mda.run_timed(steps)
    awall = time.perf_counter() - a0
    mda.close()
    out["atomic"] = {"pairs_per_s": E * steps / awall, "kernel_us": float(np.mean(akms)) * 1e3}
    log(f"[lj] atomic: {E * steps / awall / 1e9:.2f} G pairs/s, force kernel {np.mean(akms) * 1e3:.1f} us")

    out["cpu"] = None
    if do_cpu:
        orc = _oracle()
        hc = host_cpu()
        nt = hc["threads"]
        vel0 = np.zeros_like(pos)
        se = node_sorted(edges)  # the pair order the GPU headline runs in
        c0 = time.perf_counter()
        orc.md_run(pos, vel0, se, box, 2.5, 0.005, 2, threads=nt, refcpu=True)
        t2 = time.perf_counter() - c0
        sample = int(max(2, min(5000, 12.0 / max(t2 / 2, 1e-6))))  # about 12 s of CPU work
        c0 = time.perf_counter()
        orc.md_run(pos, vel0, se, box, 2.5, 0.005, sample, threads=nt, refcpu=True)
        ct = time.perf_counter() - c0
        out["cpu"] = {"value": E * sample / ct, "unit": "pairs/s", "cores": nt, "kind": "port",
                      "sample": f"moldyn 40^3 system, {sample} of its VV steps on the node-ordered pair list "
                                f"(threaded oracle: edge chunks into thread-private force arrays, the reference's "
                                f"CPU reduction-object structure)", "host": hc}
        log(f"[lj] cpu oracle: {sample} steps in {ct:.2f}s on {nt} threads ({out['cpu']['value'] / 1e9:.3f} G pairs/s)")
        # FP64 statement: 50 steps of the default handle from the workload's
        # start against the FP64 oracle over the same (node-ordered) list
        md = pa.MD(pos, None, edges, box, 2.5, 0.005)
        md.run(50)
        p, v, f = md.get()
        md.close()
        z = np.zeros((n, 3))
        rp, rv, rf, _ = orc.md_run(pos.astype(np.float64), z, se, box, 2.5, 0.005, 50, dtype=np.float64)
        dp = p - rp
        dp -= box * np.round(dp / box)  # positions compared through the minimum image
        out["fp64_rel_err"] = {"steps": 50, "pos_abs": float(np.max(np.abs(dp))), "vel": rel_err(v, rv),
                               "force": rel_err(f, rf)}
        log(f"[lj] FP64 oracle after 50 steps: force {out['fp64_rel_err']['force']:.2e}, velocity "
            f"{out['fp64_rel_err']['vel']:.2e} relative, positions {out['fp64_rel_err']['pos_abs']:.2e} absolute")
    return out


def minimd_fp64(pos, vel, box, nt):
    """miniMD 64^3, 100 steps with rebuilds every 20, from the workload's
    start: the default binned handle against the FP64 oracle (same rebuild
    schedule)."""
    orc = _oracle()
    md = pa.MD(pos, vel, None, box, 2.5, 0.005, skin=0.3, every=20)
    md.run(100)
    p, v, f = md.get()
    md.close()
    rp, rv, rf, _ = orc.md_run_rebuild_f64(pos, vel, box, 2.5, 2.8, 20, 0.005, 100, threads=nt)
    dp = p - rp
    dp -= box * np.round(dp / box)
    return {"steps": 100, "pos_abs": float(np.max(np.abs(dp))), "vel": rel_err(v, rv), "force": rel_err(f, rf)}


def bench_minimd(steps=100, fp64=False):
    """miniMD: FCC 64^3 cells (1,048,576 atoms), density 0.8442, Maxwell
    velocities at T = 1.44, cutoff 2.5, skin 0.3, binned neighbour rebuild
    every 20 steps on the device, dt 0.005; then 64 independent systems of
    8^3-32^3 cells (log-uniform) advanced as one batched launch per step.
    pairs/s counts the half pairs within the neighbour cutoff per step.
    The timed 100 steps follow 120 untimed ones (steps 120-220 of the run).
    The single system launches its steps one by one (a rebuild every 20
    steps reads the pair count back to the host); the batch replays captured
    graphs between rebuilds.  Rebuild times are latencies: HIP events around
    each rebuild's launches, its host synchronisations included."""
    pos, box = pa.gen_fcc(64, 64, 64, rho=0.8442, jitter=0.0, seed=7)
    vel = pa.gen_velocities(len(pos), 1.44, seed=8)
    t0 = time.perf_counter()
    md = pa.MD(pos, vel, None, box[0], 2.5, 0.005, skin=0.3, every=20)
    setup = time.perf_counter() - t0
    pos0, vel0 = pos, vel
    md.run(20)  # the first rebuild re-tunes on moved atoms
    md.run(steps)
    _, _, rb0, rbms0 = md.stats()
    pa.call("pat_device_sync")
    w0 = time.perf_counter()
    md.run(steps)
    wall = time.perf_counter() - w0
    npairs, _, rb1, rbms1 = md.stats()
    _, kms = md.run_timed(20)
    md.close()
    out = {"atoms": len(pos), "half_pairs": npairs, "steps": steps, "setup_s": setup,
           "pairs_per_s": npairs * steps / wall, "ms_per_step": wall * 1e3 / steps,
           "force_kernel_us": float(np.mean(kms)) * 1e3, "rebuilds": rb1 - rb0,
           "rebuild_ms_each": (rbms1 - rbms0) / max(rb1 - rb0, 1)}
    log(f"[minimd] 64^3: {len(pos)} atoms, {npairs} half pairs, {out['ms_per_step'] * 1e3:.1f} us/step with "
        f"rebuilds ({out['pairs_per_s'] / 1e9:.1f} G pairs/s), force kernel {out['force_kernel_us']:.1f} us, "
        f"{out['rebuilds']} rebuilds at {out['rebuild_ms_each']:.2f} ms each (latency)")
    systems = pa.batch_systems(64, 8, 32, seed=99)
    b = pa.MDBatch(systems)
    b.run(20)
    b.run(steps)  # captures the segment graphs and replays them: a graph's first two launches carry one-time costs
    _, _, brb0, brbms0 = b.stats()
    pa.call("pat_device_sync")
    w0 = time.perf_counter()
    b.run(steps)
    bwall = time.perf_counter() - w0
    bpairs, _, brb1, brbms1 = b.stats()
    b.close()
    natoms = sum(len(p) for p, _, _ in systems)
    brb_each = (brbms1 - brbms0) / max(brb1 - brb0, 1)
    out["batch"] = {"systems": 64, "atoms": natoms, "half_pairs": bpairs, "pairs_per_s": bpairs * steps / bwall,
                    "ms_per_step": bwall * 1e3 / steps, "rebuilds": brb1 - brb0, "rebuild_ms_each": brb_each}
    log(f"[minimd] batch of 64 systems (8^3-32^3 cells): {natoms} atoms, {bwall * 1e3 / steps * 1e3:.1f} us/step "
        f"({bpairs * steps / bwall / 1e9:.1f} G pairs/s), {brb1 - brb0} rebuilds at {brb_each:.2f} ms each (latency)")
    if fp64:
        out["fp64_rel_err"] = minimd_fp64(pos0, vel0, box[0], cpu_threads())
        e = out["fp64_rel_err"]
        log(f"[minimd] FP64 oracle after 100 steps: force {e['force']:.2e}, velocity {e['vel']:.2e} relative, "
            f"positions {e['pos_abs']:.2e} absolute")
    return out


def bench_stencil(n, steps, cpu=False):
    g = pa.gen_heat3d(n, n, n, seed=42)
    s = pa.Stencil(n, n, n, alpha=0.1)
    s.upload(g)
    s.run_timed(5)
    _, kms = s.run_timed(steps)  # sweeps launched one by one, events around each
    s.run(steps)  # captures the sweep graph
    s.run(steps)  # and replays it once: a graph's first two launches carry one-time costs
    pa.call("pat_device_sync")
    w0 = time.perf_counter()
    dev_ms = s.run_graph_timed(steps)
    wall = time.perf_counter() - w0
    s.close()
    cells = n ** 3
    kavg = dev_ms / steps  # device time per sweep inside the graph, gaps between launches included
    out = {
        "grid": f"{n}^3", "steps": steps, "gcells_per_s": cells * steps / wall / 1e9,
        "kernel_us": kavg * 1e3, "achieved_gbs": 8 * cells / (kavg * 1e-3) / 1e9,
        "frac": 8 * cells / (kavg * 1e-3) / 1e9 / HBM_PEAK_GBS, "eager_kernel_us": float(np.mean(kms)) * 1e3,
    }
    if cpu:
        orc = _oracle()
        # FP64 statement: the workload from its start against the FP64 oracle
        s = pa.Stencil(n, n, n, alpha=0.1)
        s.upload(g)
        s.run(steps)
        got = s.download()
        s.close()
        ref64 = orc.heat3d(g.astype(np.float64), steps, dtype=np.float64)
        ref32 = orc.heat3d(g, steps)
        out["oracle_bitexact"] = bool(np.array_equal(got, ref32))
        out["fp64_rel_err"] = rel_err(got, ref64)
        out["linf"] = float(np.max(np.abs(got)))
        # CPU baseline: the tiled threaded oracle (16^3 tiles statically split
        # over the threads, the reference's CPU stencil engine) on the whole
        # workload, repeated to about 3 s
        hc = host_cpu()
        reps, t = 0, 0.0
        while t < 3.0 and reps < 50:
            c0 = time.perf_counter()
            orc.heat3d(g, steps, threads=hc["threads"])
            t += time.perf_counter() - c0
            reps += 1
        out["cpu"] = {"value": cells * steps * reps / t / 1e9, "unit": "Gcell/s", "cores": hc["threads"],
                      "kind": "port", "sample": f"the whole {n}^3 x {steps} workload, {reps} run(s) (threaded oracle "
                                                f"over 16^3 tiles, the reference CPU stencil engine's structure)",
                      "host": hc}
    return out


def bench_kmeans(iters=20, cpu=False):
    """kmeans: N = 2^24, D = 32, K = 256 (256-component mixture, sigma 1,
    centres in [-10, 10]^32), K centres from the first K points, 20 Lloyd
    iterations; BF16-MFMA with FP32 rescoring vs the plain FP32 kernel."""
    n, d, k = 1 << 24, 32, 256
    t0 = time.perf_counter()
    x = pa.gen_gmm(n, d, 256, sigma=1.0, lo=-10, hi=10, seed=2024)
    xd = pa.DeviceArray.from_numpy(x)
    log(f"[kmeans] N={n} D={d} K={k}: set-up {time.perf_counter() - t0:.1f}s")
    out = {"n": n, "d": d, "k": k, "iters": iters}
    for mode, name in ((pa.PAT_KM_MFMA, "mfma"), (pa.PAT_KM_PLAIN, "plain")):
        km = pa.KMeans(xd, k, mode=mode)  # warm-up handle (first launches, LDS set-up)
        km.run_timed(2)
        km.close()
        km = pa.KMeans(xd, k, mode=mode)
        w0 = time.perf_counter()
        tot, ams = km.run_timed(iters)
        wall = time.perf_counter() - w0
        resc = km.stats()[0]
        km.close()
        kavg = float(np.mean(ams))
        out[name] = {"pt_iter_per_s": n * iters / wall, "ms_per_iter": wall * 1e3 / iters, "assign_kernel_ms": kavg,
                     "achieved_gbs": (4 * d + 4) * n / (kavg * 1e-3) / 1e9,
                     "rescored_fraction": resc / (n * iters) if mode == pa.PAT_KM_MFMA else None}
        log(f"[kmeans] {name}: {n * iters / wall / 1e9:.2f} G pt-iter/s, assign kernel {kavg:.3f} ms, "
            f"{out[name]['achieved_gbs']:.0f} GB/s algorithmic")
    out["mfma_over_plain"] = out["mfma"]["pt_iter_per_s"] / out["plain"]["pt_iter_per_s"]
    out["frac"] = out["mfma"]["achieved_gbs"] / HBM_PEAK_GBS
    if cpu:
        orc = _oracle()
        # FP64 statement on the MFMA run's 20th iteration: the centres against
        # the FP64 means of the same clusters; a 65,536-point sample's
        # assignments against the FP32 oracle (bit-exact) and the FP64 argmin
        km = pa.KMeans(xd, k, mode=pa.PAT_KM_MFMA)
        km.run(iters - 1)
        c_prev = km.centroids()
        km.run(1)
        c, a, _ = km.get()
        km.close()
        cnt = np.bincount(a, minlength=k)
        nz = cnt > 0
        mean64 = np.zeros((k, d))
        for j in range(d):
            mean64[:, j] = np.bincount(a, weights=x[:, j].astype(np.float64), minlength=k)
        mean64[nz] /= cnt[nz, None]
        idx = np.random.default_rng(0).choice(n, 65536, replace=False)
        ra, _ = orc.kmeans_assign(x[idx], c_prev)
        a64, _ = orc.kmeans_assign_f64(x[idx], c_prev.astype(np.float64))
        out["fp64_rel_err"] = {"centres_vs_fp64_cluster_means": rel_err(c[nz], mean64[nz]),
                               "sample_assign_equal_fp32_oracle": bool(np.array_equal(a[idx], ra)),
                               "sample_assign_differs_from_fp64_argmin": float(np.mean(a64 != a[idx]))}
        e = out["fp64_rel_err"]
        log(f"[kmeans] FP64: centres vs FP64 cluster means {e['centres_vs_fp64_cluster_means']:.2e} relative; sample "
            f"assignments equal the FP32 oracle: {e['sample_assign_equal_fp32_oracle']}, differ from the FP64 argmin "
            f"for {e['sample_assign_differs_from_fp64_argmin']:.2e} of points")
        # CPU baseline: the threaded oracle's Lloyd iterations (thread-private
        # reduction objects merged after the join, the reference's CPU
        # generalized-reduction engine) on the first m points, m sized for
        # about 10 s
        hc = host_cpu()
        sh = orc.kmeans_shift(n, float(np.max(np.abs(x))))
        m0 = 1 << 18
        c0 = time.perf_counter()
        orc.kmeans_lloyd(x[:m0], k, 1, sh, threads=hc["threads"])
        t1 = time.perf_counter() - c0
        m = int(min(n, max(m0, m0 * 5.0 / max(t1, 1e-6))))
        c0 = time.perf_counter()
        orc.kmeans_lloyd(x[:m], k, 2, sh, threads=hc["threads"])
        ct = time.perf_counter() - c0
        out["cpu"] = {"value": m * 2 / ct, "unit": "pt-iter/s", "cores": hc["threads"], "kind": "port",
                      "sample": f"2 Lloyd iterations over the first {m} of the {n} points (threaded oracle: "
                                f"map1 per point into thread-private reduction objects merged after the join)",
                      "host": hc}
        log(f"[kmeans] cpu oracle: {m} points x 2 iterations in {ct:.2f}s on {hc['threads']} threads "
            f"({m * 2 / ct / 1e9:.3f} G pt-iter/s)")
    xd.free()
    return out


def serving_mix(n, seed):
    """The serving mix: one third each of kmeans (N log-uniform 10k-4M,
    D in {8, 16, 32, 64, 128}, K log-uniform 16-1024, 10 iterations),
    heat3D (cube side uniform 32-384, 50 steps) and LJ (FCC cells per side
    uniform 4-48, 20 steps), in a seeded random order."""
    rng = np.random.default_rng(seed)
    kinds = rng.permutation(np.arange(n) % 3)
    jobs = []
    for i, kind in enumerate(kinds):
        if kind == 0:
            npts = int(np.exp(rng.uniform(np.log(1e4), np.log(4e6))))
            d = int(rng.choice([8, 16, 32, 64, 128]))
            k = int(np.exp(rng.uniform(np.log(16), np.log(1024))))
            jobs.append((pa.PAT_JOB_KMEANS, npts, 10, d, k))
        elif kind == 1:
            jobs.append((pa.PAT_JOB_HEAT3D, int(rng.integers(32, 385)), 50, 0, 0))
        else:
            jobs.append((pa.PAT_JOB_LJ, int(rng.integers(4, 49)), 20, 0, 0))
    return jobs


def run_serving(jobs, slots, plain, rate=None, seed=0, budget=200 << 30):
    """Submits the jobs (all at once, or as a Poisson stream at 'rate' jobs/s)
    and waits for them; returns latencies (ms, arrival to result on the host),
    the jobs' device times (start to end, ms) and the wall time."""
    s = pa.Serve(budget, slots=slots, plain=plain, fuse=not plain)  # the plain path runs one job at a time
    arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, len(jobs))) if rate else None
    ids = []
    t0 = time.perf_counter()
    for i, (kind, size, steps, d, k) in enumerate(jobs):
        if arrivals is not None:
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
        ids.append(s.submit(pa.job(kind, size, steps, 7000 + i, dim=d, k=k)))
        if i % 2000 == 1999:
            log(f"[serve] {i + 1} of {len(jobs)} jobs submitted ({time.perf_counter() - t0:.0f} s)")
    res = [s.wait(i) for i in ids]
    wall = time.perf_counter() - t0
    s.close()
    bad = [r for r in res if r["status"] != 0]
    if bad:
        raise RuntimeError(f"{len(bad)} serving jobs failed, first status {bad[0]['status']}")
    service = np.array([r["end_ms"] - r["start_ms"] for r in res])
    return np.array([r["latency_ms"] for r in res]), service, wall, np.array([r["batch"] for r in res])


def bench_serving(njobs, rate, slots=8):
    """The serving workload at one fixed Poisson arrival rate, beside the
    plain one-at-a-time path (one slot, plain FP32 kmeans) at the same
    arrivals.  The loop's capacity comes from a saturated run (all jobs
    queued at once); the plain path's from its jobs' summed service times."""
    jobs = serving_mix(njobs, seed=5)
    _, _, sat_wall, sat_batch = run_serving(jobs, slots, False)  # saturated: capacity of the loop
    capacity = njobs / sat_wall
    lat, serv, wall, batch = run_serving(jobs, slots, False, rate=rate, seed=6)
    plat, pserv, pwall, _ = run_serving(jobs, 1, True, rate=rate, seed=6)
    plain_capacity = njobs / (np.sum(pserv) * 1e-3)
    out = {"jobs": njobs, "slots": slots, "mix": "1/3 kmeans (N 10k-4M log, D 8-128, K 16-1024 log, 10 it), "
                                                 "1/3 heat3D (side 32-384, 50 steps), 1/3 LJ (4^3-48^3 cells, 20 steps)",
           "budget_bytes": 200 << 30, "arrival_rate_jobs_per_s": rate, "capacity_jobs_per_s": capacity,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "jobs_per_s": njobs / wall,
           "plain_one_at_a_time": {"p50_ms": float(np.percentile(plat, 50)), "p99_ms": float(np.percentile(plat, 99)),
                                   "jobs_per_s": njobs / pwall, "capacity_jobs_per_s": plain_capacity},
           "capacity_over_plain": capacity / plain_capacity}
    # LJ jobs fused into MD batches (jobs waiting with the same step count)
    lj = np.array([j[0] for j in jobs]) == pa.PAT_JOB_LJ
    out["lj_fused"] = {"saturated_share": float(np.mean(sat_batch[lj] > 1)),
                       "saturated_mean_batch": float(np.mean(sat_batch[lj])),
                       "rated_share": float(np.mean(batch[lj] > 1))}
    log(f"[serve] LJ jobs fused: {out['lj_fused']['saturated_share']:.0%} when saturated (mean batch "
        f"{out['lj_fused']['saturated_mean_batch']:.1f}), {out['lj_fused']['rated_share']:.0%} at {rate:.0f} jobs/s")
    # the rated run by job kind: what makes the tail (latency = queueing + service)
    kinds = np.array([j[0] for j in jobs])
    out["by_kind"] = {}
    for kind, name in ((pa.PAT_JOB_KMEANS, "kmeans"), (pa.PAT_JOB_HEAT3D, "heat3d"), (pa.PAT_JOB_LJ, "lj")):
        m = kinds == kind
        out["by_kind"][name] = {"jobs": int(m.sum()), "p50_ms": float(np.percentile(lat[m], 50)),
                                "p99_ms": float(np.percentile(lat[m], 99)),
                                "service_p50_ms": float(np.percentile(serv[m], 50)),
                                "service_p99_ms": float(np.percentile(serv[m], 99))}
    log("[serve] by kind at the fixed rate: " + "; ".join(
        f"{n} p50 {v['p50_ms']:.1f} p99 {v['p99_ms']:.1f} ms (service p99 {v['service_p99_ms']:.1f} ms)"
        for n, v in out["by_kind"].items()))
    log(f"[serve] {njobs} jobs, {slots} slots, Poisson {rate:.0f} jobs/s: p50 {out['p50_ms']:.1f} ms p99 "
        f"{out['p99_ms']:.1f} ms (capacity {capacity:.1f} jobs/s); plain one-at-a-time at the same arrivals: p50 "
        f"{out['plain_one_at_a_time']['p50_ms']:.1f} ms p99 {out['plain_one_at_a_time']['p99_ms']:.1f} ms "
        f"(capacity {plain_capacity:.1f} jobs/s)")
    return out


def bench_copy(mib=512, reps=5):
    """A plain device-to-device copy of known size: the reference point that
    scales the HBM byte counters (FETCH_SIZE, WRITE_SIZE) of the kernels."""
    n = mib << 20
    a = pa.DeviceArray((n,), np.uint8)
    b = pa.DeviceArray((n,), np.uint8)
    pa.call("pat_memset_dev", a.ptr, 1, n, None)
    pa.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        pa.call("pat_memcpy_d2d", b.ptr, a.ptr, n, None)
    pa.sync()
    dt = (time.perf_counter() - t0) / reps
    a.free()
    b.free()
    log(f"[copy] {mib} MiB device copy: {2 * n / dt / 1e9:.0f} GB/s (read + write)")
    return {"bytes": n, "gbs": 2 * n / dt / 1e9}


def profiled_traffic(key):
    """HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    passes over this benchmark (scaled by a plain copy), from the committed
    profile summary; None when absent."""
    f = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(f) as fh:
            return float(json.load(fh)[key])
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--engines", default="lj,stencil,kmeans,minimd,serve",
                    help="comma-separated subset of lj,stencil,kmeans,minimd,serve,copy (lj carries value)")
    ap.add_argument("--serve-jobs", type=int, default=10000, help="jobs in the serving mix")
    ap.add_argument("--serve-rate", type=float, default=60.0,
                    help="Poisson arrival rate (jobs/s) of the serving mix, for the loop and the plain path alike")
    ap.add_argument("--lj-forms", default="", help="comma-separated sliced LJ loop forms (0-4) to time beside the "
                                                    "tuner's choice")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every engine here runs its whole domain on a single MI355X, and no "
              f"multi-GPU path is built (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    pa.init(0)

    which = set(args.engines.split(","))
    engines = {}
    if "copy" in which:
        engines["copy"] = bench_copy()
    if "serve" in which:
        engines["serving"] = bench_serving(args.serve_jobs, args.serve_rate)
    cpu = not args.no_cpu_baseline
    if "kmeans" in which:
        engines["kmeans"] = bench_kmeans(cpu=cpu)
    if "minimd" in which:
        engines["minimd"] = bench_minimd(fp64=cpu)
    if "stencil" in which:
        for n in (128, 384):
            st = bench_stencil(n, 100, cpu=cpu and n == 128)
            engines[f"heat3d_{n}"] = st
            log(f"[stencil] {n}^3: {st['gcells_per_s']:.1f} Gcell/s, {st['kernel_us']:.1f} us per sweep in the graph "
                f"({st['eager_kernel_us']:.1f} us launched one by one), "
                f"{st['achieved_gbs']:.0f} GB/s ({100 * st['frac']:.1f}% HBM)")
    if "lj" not in which:
        print(json.dumps({"engines": engines}))
        return
    forms = [int(f) for f in args.lj_forms.split(",") if f != ""]
    lj = bench_lj(args.steps, args.warmup, not args.no_cpu_baseline, forms)
    engines["lj_moldyn"] = {k: v for k, v in lj.items() if k not in ("cpu", "fp64_rel_err")}
    head = lj["node_order"]
    # every engine's CPU baseline and FP64 statement, side by side
    cpu_all = {"lj_moldyn": lj["cpu"]}
    fp64 = {"lj_moldyn": lj.get("fp64_rel_err")}
    for key in ("heat3d_128", "kmeans", "minimd"):
        e = engines.get(key, {})
        if "cpu" in e:
            cpu_all[key] = e.pop("cpu")
        if "fp64_rel_err" in e:
            fp64[key] = e.pop("fp64_rel_err")

    out = {
        "metric": METRIC,
        "value": head["pairs_per_s"],
        "unit": "LJ pairs/s (edges/s)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded FCC lattice + jitter, random pair order)",
        "config": {"workload": "moldyn irregular reduction: FCC 40^3 cells (256,000 molecules), rho 0.8442, "
                               "jitter 0.05, half pair list within 2.5 (seeded random order), velocity-Verlet "
                               "dt 0.005, one force evaluation + update per step",
                   "pairs": lj["pairs"], "nodes": lj["nodes"],
                   "kernel": f"in-order CSR force sums in node order, {head['kernel']}, VV update fused"},
        "roofline": {"bound": "hbm", "achieved": head["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": head["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(f"lj_step_bytes_form{head['form']}")},
        "cpu_baseline": lj["cpu"],
        "cpu_baselines": cpu_all,
        "fp64_rel_err": fp64,
        "engines": engines,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
