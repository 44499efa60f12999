"""The serving loop and the device generators it uses.

Bar: the device generators equal the host ones bit for bit (heat3D grid,
FCC lattice; the mixture points, which use the device's FP64 log/cos, within
1e-6 relative); every job's result is bit-identical alone and in a mix of
concurrent jobs, and equal to the standalone handle path on the same input;
admission never lets the jobs in flight exceed the memory budget.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MB = 1 << 20


def test_device_generators_match_host(pa, gpu):
    g = pa.gen_heat3d_dev(40, 33, 21, seed=7).numpy()
    assert np.array_equal(g, pa.gen_heat3d(40, 33, 21, seed=7))
    p4, box = pa.gen_fcc_dev(6, 5, 7, rho=0.8442, jitter=0.05, seed=3)
    pos, hbox = pa.gen_fcc(6, 5, 7, rho=0.8442, jitter=0.05, seed=3)
    assert np.array_equal(box, hbox)
    p4 = p4.numpy()
    assert np.array_equal(p4[:, :3], pos) and not p4[:, 3].any()
    x = pa.gen_gmm_dev(5000, 24, 17, sigma=1.0, lo=-10, hi=10, seed=9).numpy()
    hx = pa.gen_gmm(5000, 24, 17, sigma=1.0, lo=-10, hi=10, seed=9)
    same = np.mean(x == hx)
    print(f"gmm device vs host: {same:.4%} of values bit-identical")
    assert np.max(np.abs(x - hx) / np.maximum(np.abs(hx), 1.0)) < 1e-6


def mixed_jobs(pa):
    """Small jobs of every kind, each with its own output buffer."""
    specs = [(pa.PAT_JOB_HEAT3D, 20, 7, 0, 0), (pa.PAT_JOB_HEAT3D, 37, 12, 0, 0), (pa.PAT_JOB_HEAT3D, 64, 5, 0, 0),
             (pa.PAT_JOB_LJ, 4, 10, 0, 0), (pa.PAT_JOB_LJ, 6, 25, 0, 0), (pa.PAT_JOB_LJ, 9, 21, 0, 0),
             (pa.PAT_JOB_KMEANS, 3000, 4, 16, 20), (pa.PAT_JOB_KMEANS, 12000, 3, 32, 256),
             (pa.PAT_JOB_KMEANS, 5000, 3, 128, 40), (pa.PAT_JOB_KMEANS, 4000, 2, 64, 300),
             (pa.PAT_JOB_HEAT3D, 48, 9, 0, 0), (pa.PAT_JOB_LJ, 7, 3, 0, 0)]
    out = []
    for i, (kind, size, steps, dim, k) in enumerate(specs):
        n = {pa.PAT_JOB_HEAT3D: size ** 3, pa.PAT_JOB_LJ: 4 * size ** 3 * 3,
             pa.PAT_JOB_KMEANS: dim * k}[kind]
        out.append((kind, size, steps, dim, k, 1000 + i, n))
    return out


def run_jobs(pa, specs, slots, together, plain=False):
    s = pa.Serve(2 << 30, slots=slots, plain=plain)
    bufs, ids, res = [], [], []
    for kind, size, steps, dim, k, seed, n in specs:
        b = np.zeros(n, dtype=np.float32)
        bufs.append(b)
        i = s.submit(pa.job(kind, size, steps, seed, dim=dim, k=k, out=b), keep=b)
        if not together:
            res.append(s.wait(i))
        ids.append(i)
    if together:
        res = [s.wait(i) for i in ids]
    s.close()
    return bufs, res


def test_serve_alone_equals_mix(pa, gpu):
    specs = mixed_jobs(pa)
    alone, ra = run_jobs(pa, specs, slots=1, together=False)
    mix, rm = run_jobs(pa, specs, slots=8, together=True)
    for spec, a, m, x, y in zip(specs, alone, mix, ra, rm):
        assert x["status"] == 0 and y["status"] == 0, (spec, x, y)
        assert np.array_equal(a, m), spec
        assert x["checksum"] == y["checksum"] and x["kernel"] == y["kernel"]
        assert np.all(np.isfinite(a)) and np.any(a != 0)
    # the mix really overlapped on the device
    starts = np.array([r["start_ms"] for r in rm])
    ends = np.array([r["end_ms"] for r in rm])
    overlap = max(np.sum((starts <= t) & (ends > t)) for t in starts)
    print(f"mix: up to {overlap} jobs in flight; alone total {sum(r['end_ms'] - r['start_ms'] for r in ra):.1f} ms, "
          f"mix span {ends.max() - starts.min():.1f} ms")
    assert overlap > 1


def test_serve_equals_handles(pa, gpu):
    """Each job's result equals the standalone handle path on the same input."""
    s = pa.Serve(1 << 30, slots=4)
    # heat3D: 30^3, 9 steps
    h = np.zeros(30 ** 3, dtype=np.float32)
    ih = s.submit(pa.job(pa.PAT_JOB_HEAT3D, 30, 9, 77, out=h), keep=h)
    # LJ: 7^3 cells, 23 steps (a rebuild at step 20)
    p = np.zeros(4 * 7 ** 3 * 3, dtype=np.float32)
    il = s.submit(pa.job(pa.PAT_JOB_LJ, 7, 23, 78, out=p), keep=p)
    # kmeans: 20000 x 32, K = 64, 5 iterations
    c = np.zeros(64 * 32, dtype=np.float32)
    ik = s.submit(pa.job(pa.PAT_JOB_KMEANS, 20000, 5, 79, dim=32, k=64, out=c), keep=c)
    rh, rl, rk = s.wait(ih), s.wait(il), s.wait(ik)
    s.close()
    assert rh["status"] == rl["status"] == rk["status"] == 0

    st = pa.Stencil(30, 30, 30, alpha=0.1)
    st.upload(pa.gen_heat3d(30, 30, 30, seed=77))
    st.run(9)
    assert np.array_equal(h.reshape(30, 30, 30), st.download()[1:-1, 1:-1, 1:-1])
    st.close()

    pos, box = pa.gen_fcc(7, 7, 7, rho=0.8442, jitter=0.05, seed=78)
    md = pa.MD(pos, None, None, box[0], 2.5, 0.005, skin=0.3, every=20)
    md.run(23)
    assert np.array_equal(p.reshape(-1, 3), md.get()[0])
    md.close()

    x = pa.gen_gmm_dev(20000, 32, 64, seed=79)
    km = pa.KMeans(x, 64, mode=pa.PAT_KM_MFMA)
    km.run(5)
    assert np.array_equal(c.reshape(64, 32), km.centroids())
    km.close()


def test_serve_fuses_waiting_lj_jobs(pa, gpu):
    """LJ jobs waiting with the same step count are admitted together and run
    as one MD batch (one fused launch per step over every system).  Each
    result equals the same job with fusing off and the standalone handle,
    bit for bit."""
    specs = [(6, 31), (9, 32), (7, 33), (12, 34), (6, 35), (10, 36)]  # (cells per side, seed)
    steps = 21  # a neighbour rebuild at step 20

    def serve(fuse):
        s = pa.Serve(2 << 30, slots=1, fuse=fuse)
        # a heat3D job holds the one slot while the LJ jobs queue behind it
        first = s.submit(pa.job(pa.PAT_JOB_HEAT3D, 200, 50, 5))
        bufs, ids = [], []
        for c, seed in specs:
            b = np.zeros(4 * c ** 3 * 3, dtype=np.float32)
            bufs.append(b)
            ids.append(s.submit(pa.job(pa.PAT_JOB_LJ, c, steps, seed, out=b), keep=b))
        res = [s.wait(i) for i in ids]
        assert s.wait(first)["status"] == 0
        s.close()
        return bufs, res

    fb, fr = serve(True)
    ab, ar = serve(False)
    print("batch sizes fused:", [r["batch"] for r in fr], "unfused:", [r["batch"] for r in ar])
    assert all(r["status"] == 0 for r in fr + ar)
    assert all(r["batch"] == 1 for r in ar)
    assert all(r["batch"] == len(specs) for r in fr)
    assert len({(r["start_ms"], r["end_ms"]) for r in fr}) == 1  # one batch, one start and end
    for (c, seed), f, a, x, y in zip(specs, fb, ab, fr, ar):
        assert np.array_equal(f, a), c
        assert x["checksum"] == y["checksum"], c
        assert np.isclose(x["checksum"], f.astype(np.float64).sum(), rtol=1e-9)
    for (c, seed), f in list(zip(specs, fb))[:2]:
        pos, box = pa.gen_fcc(c, c, c, rho=0.8442, jitter=0.05, seed=seed)
        md = pa.MD(pos, None, None, box[0], 2.5, 0.005, skin=0.3, every=20)
        md.run(steps)
        assert np.array_equal(f.reshape(-1, 3), md.get()[0]), c
        md.close()


def test_serve_fuses_waiting_heat_jobs(pa, gpu):
    """Small heat3D jobs waiting with the same step count run as one batch
    (two sweeps per launch over every grid, an odd last sweep per grid).
    Each result equals the job with fusing off and the standalone handle,
    bit for bit."""
    specs = [(20, 41), (37, 42), (64, 43), (48, 44), (100, 45), (128, 46), (33, 47)]  # (side, seed)
    steps = 9

    def serve(fuse):
        s = pa.Serve(2 << 30, slots=1, fuse=fuse)
        # a grid too large to fuse holds the one slot while the others queue
        first = s.submit(pa.job(pa.PAT_JOB_HEAT3D, 300, 20, 6))
        bufs, ids = [], []
        for n, seed in specs:
            b = np.zeros(n ** 3, dtype=np.float32)
            bufs.append(b)
            ids.append(s.submit(pa.job(pa.PAT_JOB_HEAT3D, n, steps, seed, out=b), keep=b))
        res = [s.wait(i) for i in ids]
        r0 = s.wait(first)
        s.close()
        assert r0["status"] == 0 and r0["batch"] == 1
        return bufs, res

    fb, fr = serve(True)
    ab, ar = serve(False)
    print("batch sizes fused:", [r["batch"] for r in fr], "unfused:", [r["batch"] for r in ar])
    assert all(r["status"] == 0 for r in fr + ar)
    assert all(r["batch"] == 1 for r in ar)
    assert all(r["batch"] == len(specs) for r in fr)
    for (n, seed), f, a, x, y in zip(specs, fb, ab, fr, ar):
        assert np.array_equal(f, a), n
        assert x["checksum"] == y["checksum"], n
    for (n, seed), f in list(zip(specs, fb))[:3]:
        st = pa.Stencil(n, n, n, alpha=0.1)
        st.upload(pa.gen_heat3d(n, n, n, seed=seed))
        st.run(steps)
        assert np.array_equal(f.reshape(n, n, n), st.download()[1:-1, 1:-1, 1:-1]), n
        st.close()


def test_serve_fused_member_error_stays_its_own(pa, gpu):
    """A job whose output buffer is too small fails with PAT_ERR_INVALID
    inside a fused batch; the other jobs of the batch finish with the same
    bits as alone."""
    for kind, sizes, steps, per in ((pa.PAT_JOB_LJ, [6, 7, 8], 5, lambda n: 4 * n ** 3 * 3),
                                    (pa.PAT_JOB_HEAT3D, [24, 40, 31], 4, lambda n: n ** 3)):
        def serve(fuse):
            s = pa.Serve(1 << 30, slots=1, fuse=fuse)
            first = s.submit(pa.job(pa.PAT_JOB_HEAT3D, 250, 20, 8))
            bufs, ids = [], []
            for i, n in enumerate(sizes):
                b = np.zeros(per(n) - (1 if i == 1 else 0), dtype=np.float32)  # the second is one float short
                bufs.append(b)
                ids.append(s.submit(pa.job(kind, n, steps, 90 + i, out=b), keep=b))
            res = [s.wait(i) for i in ids]
            assert s.wait(first)["status"] == 0
            s.close()
            return bufs, res

        fb, fr = serve(True)
        ab, ar = serve(False)
        assert [r["batch"] for r in fr] == [3, 3, 3], kind
        assert [r["status"] for r in fr] == [r["status"] for r in ar] == [0, pa.PAT_ERR_INVALID, 0], kind
        for i in (0, 2):
            assert np.array_equal(fb[i], ab[i]) and fr[i]["checksum"] == ar[i]["checksum"], (kind, i)


def test_serve_fusion_respects_budget(pa, gpu):
    """A batch takes waiting jobs only while their estimates fit the budget:
    with room for five 8^3-cell LJ jobs (about 7 MB each) in 36 MB, eight
    waiting jobs run as a batch of five and then one of three, the bytes
    held never above the budget, every result as alone."""
    budget = 36 << 20

    def serve(fuse):
        s = pa.Serve(budget, slots=1, fuse=fuse)
        first = s.submit(pa.job(pa.PAT_JOB_HEAT3D, 100, 2000, 9))  # holds the slot while the others queue
        bufs, ids = [], []
        for i in range(8):
            b = np.zeros(4 * 8 ** 3 * 3, dtype=np.float32)
            bufs.append(b)
            ids.append(s.submit(pa.job(pa.PAT_JOB_LJ, 8, 6, 300 + i, out=b), keep=b))
        res = [s.wait(i) for i in ids]
        assert s.wait(first)["status"] == 0
        s.close()
        return bufs, res

    fb, fr = serve(True)
    ab, ar = serve(False)
    assert all(r["status"] == 0 for r in fr + ar)
    assert [r["batch"] for r in fr] == [5] * 5 + [3] * 3
    # each batch's jobs are admitted together and released together
    assert sum(r["device_bytes"] for r in fr[:5]) <= budget
    assert fr[4]["end_ms"] <= fr[5]["start_ms"]
    for f, a, x, y in zip(fb, ab, fr, ar):
        assert np.array_equal(f, a) and x["checksum"] == y["checksum"]


def test_serve_fused_zero_step_jobs(pa, gpu):
    """Jobs of zero steps fuse too and return their generated inputs: heat3D
    grids and LJ positions equal the host generators' values."""
    s = pa.Serve(1 << 30, slots=1)
    first = s.submit(pa.job(pa.PAT_JOB_HEAT3D, 250, 20, 8))
    jobs = [(pa.PAT_JOB_HEAT3D, 30, 61), (pa.PAT_JOB_LJ, 6, 62), (pa.PAT_JOB_HEAT3D, 45, 63), (pa.PAT_JOB_LJ, 7, 64)]
    bufs, ids = [], []
    for kind, n, seed in jobs:
        b = np.zeros(n ** 3 if kind == pa.PAT_JOB_HEAT3D else 4 * n ** 3 * 3, dtype=np.float32)
        bufs.append(b)
        ids.append(s.submit(pa.job(kind, n, 0, seed, out=b), keep=b))
    res = [s.wait(i) for i in ids]
    assert s.wait(first)["status"] == 0
    s.close()
    assert [r["status"] for r in res] == [0] * 4 and [r["batch"] for r in res] == [2] * 4
    for (kind, n, seed), b in zip(jobs, bufs):
        if kind == pa.PAT_JOB_HEAT3D:
            ref = pa.gen_heat3d(n, n, n, seed=seed)[1:-1, 1:-1, 1:-1]
            assert np.array_equal(b.reshape(n, n, n), ref), n
        else:
            pos, _ = pa.gen_fcc(n, n, n, rho=0.8442, jitter=0.05, seed=seed)
            assert np.array_equal(b.reshape(-1, 3), pos), n


def test_serve_admission_respects_budget(pa, gpu):
    # heat3D 200^3 holds about 2 x 32.5 MB: a 100 MB budget fits one at a time
    s = pa.Serve(100 * MB, slots=4)
    ids = [s.submit(pa.job(pa.PAT_JOB_HEAT3D, 200, 3, 500 + i)) for i in range(5)]
    small = [s.submit(pa.job(pa.PAT_JOB_HEAT3D, 16, 3, 600 + i)) for i in range(6)]
    res = [s.wait(i) for i in ids + small]
    with pytest.raises(RuntimeError, match="budget"):
        s.submit(pa.job(pa.PAT_JOB_HEAT3D, 300, 1, 1))
    s.close()
    assert all(r["status"] == 0 for r in res)
    ev = sorted([(r["start_ms"], r["device_bytes"]) for r in res] + [(r["end_ms"], -r["device_bytes"]) for r in res],
                key=lambda t: (t[0], t[1]))
    held = peak = 0
    for _, b in ev:
        held += b
        peak = max(peak, held)
    assert peak <= 100 * MB


def test_serving_mix_10000_jobs(pa, gpu):
    """The serving workload's job list: 10,000 jobs, one third each of
    kmeans (N 10k-4M, D in {8..128}, K 16-1024, 10 iterations), heat3D (side
    32-384, 50 steps) and LJ (4^3-48^3 cells, 20 steps), all submitted to 8
    slots under a 200 GB budget.  Every job succeeds; the jobs in flight
    never hold more than the budget; a seeded sample of 30 jobs -- the
    largest of each kind among them -- equals its standalone handle run bit
    for bit."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    jobs = bench.serving_mix(10000, seed=5)
    kinds = np.array([j[0] for j in jobs])
    size = np.array([j[1] for j in jobs], dtype=np.float64)
    work = np.array([j[1] * j[3] * j[4] if j[0] == pa.PAT_JOB_KMEANS else j[1] for j in jobs], dtype=np.float64)
    pick = set()
    for kind in (pa.PAT_JOB_HEAT3D, pa.PAT_JOB_LJ, pa.PAT_JOB_KMEANS):
        idx = np.flatnonzero(kinds == kind)
        pick.add(int(idx[np.argmax(work[idx] if kind == pa.PAT_JOB_KMEANS else size[idx])]))
    rng = np.random.default_rng(17)
    while len(pick) < 30:
        pick.add(int(rng.integers(len(jobs))))
    budget = 200 << 30
    s = pa.Serve(budget, slots=8)
    bufs, ids = {}, []
    for i, (kind, n, steps, d, k) in enumerate(jobs):
        out = None
        if i in pick:
            m = {pa.PAT_JOB_HEAT3D: n ** 3, pa.PAT_JOB_LJ: 4 * n ** 3 * 3, pa.PAT_JOB_KMEANS: d * k}[kind]
            out = bufs[i] = np.zeros(m, dtype=np.float32)
        ids.append(s.submit(pa.job(kind, n, steps, 7000 + i, dim=d, k=k, out=out), keep=out))
    res = [s.wait(i) for i in ids]
    s.close()
    bad = [(i, r["status"]) for i, r in enumerate(res) if r["status"] != 0]
    assert not bad, bad[:5]
    ev = sorted([(r["start_ms"], r["device_bytes"]) for r in res] + [(r["end_ms"], -r["device_bytes"]) for r in res],
                key=lambda t: (t[0], t[1]))
    held = peak = 0
    for _, b in ev:
        held += b
        peak = max(peak, held)
    span = max(r["end_ms"] for r in res) - min(r["start_ms"] for r in res)
    print(f"serving mix: 10,000 jobs in {span / 1e3:.1f} s ({1e4 / span * 1e3:.0f} jobs/s), peak admitted "
          f"{peak / 2 ** 30:.1f} GiB of {budget / 2 ** 30:.0f}")
    assert peak <= budget
    for i in sorted(pick):
        kind, n, steps, d, k = jobs[i]
        seed = 7000 + i
        if kind == pa.PAT_JOB_HEAT3D:
            st = pa.Stencil(n, n, n, alpha=0.1)
            st.upload(pa.gen_heat3d(n, n, n, seed=seed))
            st.run(steps)
            ref = st.download()[1:-1, 1:-1, 1:-1]
            st.close()
            assert np.array_equal(bufs[i].reshape(n, n, n), ref), (i, jobs[i])
        elif kind == pa.PAT_JOB_LJ:
            pos, box = pa.gen_fcc(n, n, n, rho=0.8442, jitter=0.05, seed=seed)
            md = pa.MD(pos, None, None, box[0], 2.5, 0.005, skin=0.3, every=20)
            md.run(steps)
            ref = md.get()[0]
            md.close()
            assert np.array_equal(bufs[i].reshape(-1, 3), ref), (i, jobs[i])
        else:
            x = pa.gen_gmm_dev(n, d, k, seed=seed)
            km = pa.KMeans(x, k, mode=pa.PAT_KM_MFMA)
            km.run(steps)
            ref = km.centroids()
            km.close()
            x.free()
            assert np.array_equal(bufs[i].reshape(k, d), ref), (i, jobs[i])
