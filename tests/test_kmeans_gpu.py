"""GPU parity of the generalized-reduction engine (kmeans) vs the oracle.

Bar: assignments, min_dist values, the reduction object (fixed-point sums,
counts, distance sums) and the updated centres are bit-exact against the
oracle, for the plain FP32 kernel and for the BF16-MFMA kernel with FP32
rescoring.  The error of the FP32 centres against an FP64 Lloyd run is
printed for the mid-size case.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oracle_iters(orc, pa, x, k, iters):
    maxabs = float(np.max(np.abs(x)))
    s, ds = pa.kmeans_shifts(len(x), x.shape[1], maxabs)
    c = x[:k].copy()
    for _ in range(iters):
        a, md = orc.kmeans_assign(x, c)
        sums, counts, dsum = orc.kmeans_accumulate(x, a, k, s, mindist=md, dshift=ds)
        c_next = orc.kmeans_update(sums, counts, s, c)
        yield c, a, md, sums, counts, dsum
        c = c_next


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n,d,k", [(1, 3, 1), (31, 3, 5), (1000, 3, 40), (4099, 8, 17), (20000, 16, 64),
                                   (5003, 32, 256), (3000, 33, 70), (2000, 64, 100), (1500, 128, 33),
                                   (3000, 32, 300),
                                   # rows too large for LDS: the MFMA path reads them from global memory
                                   (6000, 128, 1024), (4000, 64, 1000), (3000, 100, 700), (9000, 16, 8000)])
def test_single_pass_bitexact(pa, orc, gpu, mode, n, d, k):
    x = pa.gen_gmm(n, d, max(1, k // 2), sigma=1.0, seed=n + d + k)
    km = pa.KMeans(x, k, mode=mode)
    km.run(1)
    c1, a, md = km.get()
    sums, counts, dsums, s, ds = km.get_object()
    km.close()
    c0, ra, rmd, rsums, rcounts, rdsum = next(oracle_iters(orc, pa, x, k, 1))
    assert np.array_equal(a, ra)
    assert np.array_equal(md, rmd)
    assert np.array_equal(counts, rcounts) and np.array_equal(sums, rsums) and np.array_equal(dsums, rdsum)
    assert np.array_equal(c1, orc.kmeans_update(rsums, rcounts, s, c0))


@pytest.mark.parametrize("mode", [0, 1])
def test_points_at_unaligned_device_address(pa, orc, gpu, mode):
    """Points read in place from a device address that is 4-byte but not
    16-byte aligned: the set-up's max |x| pass and the assignment kernels
    take their unaligned forms.  Same bits as the oracle (the sums agree
    only if both found the same max |x|, hence the same fixed-point shift)."""
    n, d, k = 5001, 32, 64
    x = pa.gen_gmm(n, d, 32, sigma=1.0, seed=77)
    x[1234, 5] = -40.0  # the largest |x|, inside the array

    class View(pa.DeviceArray):  # points at an offset into base; base owns the memory
        def __init__(self, base, offset, shape):
            self.shape, self.dtype = shape, np.dtype(np.float32)
            self.nbytes = int(np.prod(shape)) * 4
            self.ptr = base.ptr + offset
            self._base = base

        def free(self):
            self.ptr = None

    base = pa.DeviceArray((n * d + 4,), np.float32)
    pa.call("pat_memcpy_h2d", base.ptr + 4, x.ctypes.data, x.nbytes, None)
    pa.call("pat_stream_sync", None)
    view = View(base, 4, (n, d))
    km = pa.KMeans(view, k, mode=mode)
    km.run(1)
    c1, a, md = km.get()
    sums, counts, dsums, s, ds = km.get_object()
    km.close()
    c0, ra, rmd, rsums, rcounts, rdsum = next(oracle_iters(orc, pa, x, k, 1))
    assert np.array_equal(a, ra)
    assert np.array_equal(md, rmd)
    assert np.array_equal(counts, rcounts) and np.array_equal(sums, rsums) and np.array_equal(dsums, rdsum)
    assert np.array_equal(c1, orc.kmeans_update(rsums, rcounts, s, c0))


def adversarial(kind, rng):
    if kind == "duplicates":  # repeated centres -> exact ties
        x = rng.normal(size=(4000, 32)).astype(np.float32)
        x[:64] = x[0]
        return x, 64
    if kind == "lattice":  # integer coordinates: many equal distances
        return rng.integers(-3, 4, size=(6000, 32)).astype(np.float32), 128
    if kind == "large":  # big norms make the BF16 error bound wide
        return (rng.normal(size=(5000, 32)) * 0.5 + 3000).astype(np.float32), 96
    if kind == "close":  # many centres within one tight cluster
        return (rng.normal(size=(8000, 32)) * 1e-3).astype(np.float32), 256
    raise ValueError(kind)


@pytest.mark.parametrize("kind,terms", [("duplicates", 3), ("lattice", 3), ("large", 3), ("close", 3),
                                        ("duplicates", 1), ("lattice", 1), ("large", 1), ("close", 1),
                                        ("close_wide", 3)])
def test_mfma_adversarial_bitexact(pa, orc, gpu, kind, terms):
    if kind == "close_wide":  # a tight cluster and K x D beyond LDS (rows from global memory)
        x, k = (np.random.default_rng(8).normal(size=(3000, 128)) * 1e-3).astype(np.float32), 1024
    else:
        x, k = adversarial(kind, np.random.default_rng(7))
    out = []
    for mode in (0, 1):
        km = pa.KMeans(x, k, mode=mode)
        if mode == 1 and kind != "close_wide":
            km.set_terms(terms)
        km.run(3)
        out.append(km.get() + km.get_object()[:3])
        km.close()
    for p, m in zip(*out):
        assert np.array_equal(p, m)
    ref = list(oracle_iters(orc, pa, x, k, 3))[-1]
    assert np.array_equal(out[1][1], ref[1])


@pytest.mark.parametrize("terms", [1, 3])
def test_mfma_terms_bitexact(pa, orc, gpu, terms):
    """Either number of BF16 terms (one: a wider margin, more rescoring)
    gives the oracle's assignments and centres bit for bit, also when the
    form changes between runs."""
    x = pa.gen_gmm(50000, 32, 64, sigma=1.0, lo=-10, hi=10, seed=13)
    km = pa.KMeans(x, 64, mode=1)
    km.set_terms(terms)
    assert km.terms() == terms
    km.run(3)
    km.set_terms(4 - terms)
    km.run(2)
    c, a, md = km.get()
    km.close()
    *_, last = oracle_iters(orc, pa, x, 64, 5)
    assert np.array_equal(a, last[1]) and np.array_equal(md, last[2])
    c4 = list(oracle_iters(orc, pa, x, 64, 6))[-1][0]
    assert np.array_equal(c, c4)


def test_lloyd_20_iterations_midsize(pa, orc, gpu):
    """N = 200,000, D = 32, K = 256 from a 256-component mixture, 20 iterations:
    plain and MFMA GPU runs and the oracle agree bit for bit."""
    x = pa.gen_gmm(200000, 32, 256, sigma=1.0, lo=-10, hi=10, seed=11)
    res = []
    for mode in (0, 1):
        km = pa.KMeans(x, 256, mode=mode)
        km.run(20)
        res.append(km.get())
        resc = km.stats()[0]
        km.close()
    assert all(np.array_equal(p, m) for p, m in zip(*res))
    *_, last = oracle_iters(orc, pa, x, 256, 20)
    c19, a20, md20 = last[0], last[1], last[2]
    assert np.array_equal(res[1][1], a20) and np.array_equal(res[1][2], md20)
    maxabs = float(np.max(np.abs(x)))
    s, _ = pa.kmeans_shifts(len(x), 32, maxabs)
    sums, counts, _ = orc.kmeans_accumulate(x, a20, 256, s)
    assert np.array_equal(res[1][0], orc.kmeans_update(sums, counts, s, c19))
    # FP64 statement: the centres against the FP64 means of the same final
    # clusters (a deterministic bound: fixed-point sums plus one FP32
    # rounding), and, printed, against an FP64 Lloyd run from the same start
    cnt = np.bincount(a20, minlength=256)
    means = np.zeros((256, 32))
    np.add.at(means, a20, x.astype(np.float64))
    full = cnt > 0
    means[full] /= cnt[full, None]
    merr = np.max(np.abs(res[1][0][full] - means[full])) / np.max(np.abs(means[full]))
    assert merr <= 1e-6
    c64, _ = orc.kmeans_lloyd_f64(x, 256, 20)
    err = np.max(np.abs(res[1][0] - c64)) / np.max(np.abs(c64))
    print(f"kmeans 200k x32 K256 x20: rescored {resc} point-passes ({resc / (20 * len(x)):.2%}); "
          f"centres vs FP64 means of their clusters {merr:.2e}, vs FP64 Lloyd {err:.2e}")


def test_kmeans_workload_full_size(pa, orc, gpu):
    """kmeans workload: N = 2^24, D = 32, K = 256 (256-component mixture,
    sigma 1, centres in [-10, 10]^32), 20 iterations.  Every iteration of the
    MFMA run is checked against the oracle: the assignment and min_dist of a
    65,536-point sample (fresh each iteration) from the previous centres, and
    the whole reduction object and centre update from the full assignment.
    The plain kernel ends bit-identical.  FP64 errors printed: the centres
    against the FP64 means of the same clusters, and the sample's FP32
    assignments against the FP64 argmin."""
    n, d, k = 1 << 24, 32, 256
    x = pa.gen_gmm(n, d, 256, sigma=1.0, lo=-10, hi=10, seed=2024)
    xd = pa.DeviceArray.from_numpy(x)
    s, ds = pa.kmeans_shifts(n, d, float(np.max(np.abs(x))))
    rng = np.random.default_rng(0)
    km = pa.KMeans(xd, k, mode=pa.PAT_KM_MFMA)
    c = x[:k].copy()
    worst_centre, worst_flip = 0.0, 0.0
    for it in range(20):
        km.run(1)
        c_next, a, md = km.get()
        idx = rng.choice(n, 65536, replace=False)
        ra, rmd = orc.kmeans_assign(x[idx], c)
        assert np.array_equal(a[idx], ra) and np.array_equal(md[idx], rmd), it
        sums, counts, dsums, gs, gds = km.get_object()
        rsums, rcounts, rdsum = orc.kmeans_accumulate(x, a, k, s, mindist=md, dshift=ds)
        assert (gs, gds) == (s, ds)
        assert np.array_equal(sums, rsums) and np.array_equal(counts, rcounts) and np.array_equal(dsums, rdsum), it
        assert counts.sum() == n
        assert np.array_equal(c_next, orc.kmeans_update(rsums, rcounts, s, c)), it
        # FP64 statements: cluster means of this assignment, argmin in FP64
        nz = counts > 0
        mean64 = np.zeros((k, d))
        for j in range(d):
            mean64[:, j] = np.bincount(a, weights=x[:, j].astype(np.float64), minlength=k)
        mean64[nz] /= counts[nz, None]
        worst_centre = max(worst_centre, float(np.max(np.abs(c_next[nz] - mean64[nz])) / np.max(np.abs(mean64[nz]))))
        a64, _ = orc.kmeans_assign_f64(x[idx], c.astype(np.float64))
        worst_flip = max(worst_flip, float(np.mean(a64 != ra)))
        c = c_next
    km.close()
    print(f"kmeans 2^24 x 32, K 256: all 20 iterations equal the oracle; centres vs FP64 cluster means "
          f"{worst_centre:.2e} relative (worst iteration); FP32 assignments differing from the FP64 argmin on "
          f"the samples: {worst_flip:.2e} (worst iteration)")
    assert worst_centre < 1e-6
    kp = pa.KMeans(xd, k, mode=pa.PAT_KM_PLAIN)
    kp.run(20)
    assert np.array_equal(kp.centroids(), c)
    kp.close()
