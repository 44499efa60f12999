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


def cpu_threads():
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, min(n, 16))


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
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        nt = cpu_threads()
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
                                f"CPU reduction-object structure)"}
        log(f"[lj] cpu oracle: {sample} steps in {ct:.2f}s on {nt} threads ({out['cpu']['value'] / 1e9:.3f} G pairs/s)")
    return out


def bench_minimd(steps=100):
    """miniMD: FCC 64^3 cells (1,048,576 atoms), density 0.8442, Maxwell
    velocities at T = 1.44, cutoff 2.5, skin 0.3, binned neighbour rebuild
    every 20 steps on the device, dt 0.005; then 64 independent systems of
    8^3-32^3 cells (log-uniform) advanced as one batched launch per step.
    pairs/s counts the half pairs within the neighbour cutoff per step.
    The timed 100 steps follow 120 untimed ones (steps 120-220 of the run)."""
    pos, box = pa.gen_fcc(64, 64, 64, rho=0.8442, jitter=0.0, seed=7)
    vel = pa.gen_velocities(len(pos), 1.44, seed=8)
    t0 = time.perf_counter()
    md = pa.MD(pos, vel, None, box[0], 2.5, 0.005, skin=0.3, every=20)
    setup = time.perf_counter() - t0
    md.run(20)  # the first rebuild re-tunes on moved atoms
    md.run(steps)  # captures the segment graphs and replays them: a graph's first two launches carry one-time costs
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
        f"{out['rebuilds']} rebuilds at {out['rebuild_ms_each']:.2f} ms device time each")
    systems = pa.batch_systems(64, 8, 32, seed=99)
    b = pa.MDBatch(systems)
    b.run(20)
    b.run(steps)
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
        f"({bpairs * steps / bwall / 1e9:.1f} G pairs/s), {brb1 - brb0} rebuilds at {brb_each:.2f} ms device time each")
    return out


def bench_stencil(n, steps):
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
    return {
        "grid": f"{n}^3", "steps": steps, "gcells_per_s": cells * steps / wall / 1e9,
        "kernel_us": kavg * 1e3, "achieved_gbs": 8 * cells / (kavg * 1e-3) / 1e9,
        "frac": 8 * cells / (kavg * 1e-3) / 1e9 / HBM_PEAK_GBS, "eager_kernel_us": float(np.mean(kms)) * 1e3,
    }


def bench_kmeans(iters=20):
    """kmeans: N = 2^24, D = 32, K = 256 (256-component mixture, sigma 1,
    centres in [-10, 10]^32), K centres from the first K points, 20 Lloyd
    iterations; BF16-MFMA with FP32 rescoring vs the plain FP32 kernel."""
    n, d, k = 1 << 24, 32, 256
    t0 = time.perf_counter()
    x = pa.gen_gmm(n, d, 256, sigma=1.0, lo=-10, hi=10, seed=2024)
    xd = pa.DeviceArray.from_numpy(x)
    del x
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
    and waits for them; returns latencies (ms, arrival to result on the host)
    and the wall time."""
    s = pa.Serve(budget, slots=slots, plain=plain)
    arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, len(jobs))) if rate else None
    ids = []
    t0 = time.perf_counter()
    for i, (kind, size, steps, d, k) in enumerate(jobs):
        if arrivals is not None:
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
        ids.append(s.submit(pa.job(kind, size, steps, 7000 + i, dim=d, k=k)))
    res = [s.wait(i) for i in ids]
    wall = time.perf_counter() - t0
    s.close()
    bad = [r for r in res if r["status"] != 0]
    if bad:
        raise RuntimeError(f"{len(bad)} serving jobs failed, first status {bad[0]['status']}")
    return np.array([r["latency_ms"] for r in res]), wall


def bench_serving(njobs, slots=8):
    jobs = serving_mix(njobs, seed=5)
    _, sat_wall = run_serving(jobs, slots, False)  # saturated: capacity of the loop
    capacity = njobs / sat_wall
    rate = 0.7 * capacity
    lat, wall = run_serving(jobs, slots, False, rate=rate, seed=6)
    # one job at a time on the plain kernels, same arrivals
    plat, pwall = run_serving(jobs, 1, True, rate=rate, seed=6)
    _, seq_wall = run_serving(jobs, 1, True)
    out = {"jobs": njobs, "slots": slots, "mix": "1/3 kmeans (N 10k-4M log, D 8-128, K 16-1024 log, 10 it), "
                                                 "1/3 heat3D (side 32-384, 50 steps), 1/3 LJ (4^3-48^3 cells, 20 steps)",
           "capacity_jobs_per_s": capacity, "arrival_rate_jobs_per_s": rate,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "jobs_per_s": njobs / wall,
           "plain_one_at_a_time": {"p50_ms": float(np.percentile(plat, 50)), "p99_ms": float(np.percentile(plat, 99)),
                                   "jobs_per_s": njobs / pwall, "saturated_jobs_per_s": njobs / seq_wall},
           "speedup_vs_sequential_plain": capacity / (njobs / seq_wall)}
    log(f"[serve] {njobs} jobs, {slots} slots: capacity {capacity:.1f} jobs/s; at {rate:.1f} jobs/s p50 "
        f"{out['p50_ms']:.1f} ms p99 {out['p99_ms']:.1f} ms; plain one-at-a-time p50 "
        f"{out['plain_one_at_a_time']['p50_ms']:.1f} ms p99 {out['plain_one_at_a_time']['p99_ms']:.1f} ms, "
        f"sequential {njobs / seq_wall:.1f} jobs/s")
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
    ap.add_argument("--serve-jobs", type=int, default=600, help="jobs in the serving mix (the config's is 10000)")
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
        engines["serving"] = bench_serving(args.serve_jobs)
    if "kmeans" in which:
        engines["kmeans"] = bench_kmeans()
    if "minimd" in which:
        engines["minimd"] = bench_minimd()
    if "stencil" in which:
        for n in (128, 384):
            st = bench_stencil(n, 100)
            engines[f"heat3d_{n}"] = st
            log(f"[stencil] {n}^3: {st['gcells_per_s']:.1f} Gcell/s, {st['kernel_us']:.1f} us per sweep in the graph "
                f"({st['eager_kernel_us']:.1f} us launched one by one), "
                f"{st['achieved_gbs']:.0f} GB/s ({100 * st['frac']:.1f}% HBM)")
    if "lj" not in which:
        print(json.dumps({"engines": engines}))
        return
    forms = [int(f) for f in args.lj_forms.split(",") if f != ""]
    lj = bench_lj(args.steps, args.warmup, not args.no_cpu_baseline, forms)
    engines["lj_moldyn"] = {k: v for k, v in lj.items() if k != "cpu"}
    head = lj["node_order"]

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
        "engines": engines,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
