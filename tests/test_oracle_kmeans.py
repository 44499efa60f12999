"""Pins the kmeans oracle (oracle/oracle.c kmeans_*) without trusting it.

Checks: agreement with an FP64 brute-force argmin, hand-built ties (equal
distances and duplicated centres -> lowest index, as map1's strict < from
index 0 gives, kmeans/regular.h:89-103), a case where two different FP32
squared distances round to the same sqrt so the lower index must win
although it is farther, convergence to a fixed point on a well-separated mixture, SSE
monotone over Lloyd iterations, exact integer counts, fixed-point sums
within N * 2^-s of the FP64 sums and independent of thread count, and the
fixed-point shift formula shared with the product.
"""
import numpy as np
import pytest


def test_matches_fp64_bruteforce(pa, orc):
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2000, 7)).astype(np.float32) * 3
    c = rng.normal(size=(13, 7)).astype(np.float32) * 3
    a, md = orc.kmeans_assign(x, c)
    d64 = ((x.astype(np.float64)[:, None, :] - c.astype(np.float64)[None]) ** 2).sum(-1)
    best = d64.min(1)
    chosen = d64[np.arange(len(x)), a]
    # the FP32 choice is optimal up to FP32 rounding of the distances
    assert np.all(chosen <= best * (1 + 1e-5) + 1e-6)
    assert np.mean(a == d64.argmin(1)) > 0.999
    assert np.allclose(md, np.sqrt(chosen), rtol=1e-5)
    a64, md64 = orc.kmeans_assign_f64(x, c.astype(np.float64))
    assert np.array_equal(a64, d64.argmin(1))


def test_ties_lowest_index(orc):
    c = np.array([[1, 0], [-1, 0], [0, 1], [1, 0]], dtype=np.float32)  # c3 duplicates c0
    x = np.array([[0, 0], [1, 0], [0.5, 0.5], [0, -1]], dtype=np.float32)
    a, md = orc.kmeans_assign(x, c)
    # (0,0): c0, c1, c2, c3 all at distance 1 -> 0; (1,0): c0 == c3 -> 0;
    # (0.5,0.5): c0 and c2 equidistant -> 0; (0,-1): c0, c1 at sqrt 2 -> 0
    assert a.tolist() == [0, 0, 0, 0]
    assert md.tolist() == [1.0, 0.0, np.float32(np.sqrt(np.float32(0.5))), np.float32(np.sqrt(np.float32(2)))]


def find_sqrt_collision(seed=1):
    """Two centres whose FP32 squared distances to x differ but whose sqrt
    rounds to the same float, the nearer one listed second."""
    rng = np.random.default_rng(seed)
    for _ in range(200):
        x = rng.uniform(-2, 2, size=(1, 4)).astype(np.float32)
        cs = (x + rng.uniform(-1e-3, 1e-3, size=(20000, 4)).astype(np.float32) + np.float32(1.5)).astype(np.float32)
        t = cs - x
        d = np.zeros(len(cs), dtype=np.float32)
        for k in range(4):  # dimension order, FP32, no fused multiply-add
            d = (d + t[:, k] * t[:, k]).astype(np.float32)
        s = np.sqrt(d)
        order = np.argsort(d, kind="stable")
        ds, ss = d[order], s[order]
        hit = np.nonzero((ds[1:] > ds[:-1]) & (ss[1:] == ss[:-1]))[0]
        if len(hit):
            i = hit[0]
            near, far = order[i], order[i + 1]
            return x, np.stack([cs[far], cs[near]])
    raise AssertionError("no collision found")


def test_sqrt_collision_prefers_lower_index(orc):
    x, c = find_sqrt_collision()
    a, md = orc.kmeans_assign(x, c)
    # c[1] is strictly nearer in squared distance, but both distances round
    # to the same sqrt, and map1 compares sqrt values with a strict <
    assert a[0] == 0


def test_converges_and_sse_monotone(pa, orc):
    x = pa.gen_gmm(20000, 8, 6, sigma=0.3, lo=-10, hi=10, seed=3)
    maxabs = float(np.max(np.abs(x)))
    s = orc.kmeans_shift(len(x), maxabs)
    c, a, sse = orc.kmeans_lloyd(x, 6, 10, s, sse=True)
    assert np.all(np.diff(sse) <= 1e-9 * sse[0])
    # on a well-separated mixture Lloyd reaches a fixed point: the last
    # assignment is the nearest-centre map of the final centres
    comp = np.argmin(((x[:, None, :] - c[None]) ** 2).sum(-1), axis=1)
    assert np.array_equal(comp, a)
    sums, counts, _ = orc.kmeans_accumulate(x, a, 6, s)
    assert counts.sum() == len(x)


def test_fixed_point_sums(pa, orc):
    x = pa.gen_gmm(50000, 5, 4, seed=4)
    a = (np.arange(len(x)) % 4).astype(np.int32)
    maxabs = float(np.max(np.abs(x)))
    s = orc.kmeans_shift(len(x), maxabs)
    sums, counts, _ = orc.kmeans_accumulate(x, a, 4, s, threads=1)
    sums8, counts8, _ = orc.kmeans_accumulate(x, a, 4, s, threads=8)
    assert np.array_equal(sums, sums8) and np.array_equal(counts, counts8)
    exact = np.zeros((4, 5))
    np.add.at(exact, a, x.astype(np.float64))
    assert np.max(np.abs(sums * 2.0 ** -s - exact)) <= len(x) * 2.0 ** (-s - 1)
    assert counts.tolist() == [12500] * 4
    # every rounded value fits in 31 bits, the largest possible sum in int64
    assert maxabs * 2.0 ** s <= 2.0 ** 30 and len(x) * maxabs * 2.0 ** s < 2.0 ** 62


@pytest.mark.parametrize("n,m", [(1, 0.5), (1000, 1.0), (1 << 24, 14.7), ((1 << 24) + 1, 1e-3), (7, 3e5)])
def test_shift_formula_shared(pa, orc, n, m):
    assert pa.kmeans_fixed_shift(n, m) == orc.kmeans_shift(n, m)
    s = orc.kmeans_shift(n, m)
    assert n * m * 2.0 ** s < 2.0 ** 62 and m * 2.0 ** s <= 2.0 ** 30
    # the largest shift with both properties for a magnitude bound that is a
    # power of two (m < 2^e): one more bit breaks one of them
    e = int(np.frexp(np.float64(m))[1])
    assert n * 2.0 ** (e + s + 1) >= 2.0 ** 62 or 2.0 ** (e + s + 1) > 2.0 ** 30 or s == 60


def test_update_keeps_empty_cluster(orc):
    x = np.array([[0, 0], [2, 2]], dtype=np.float32)
    c = np.array([[0, 0], [2, 2], [50, 50]], dtype=np.float32)
    a, _ = orc.kmeans_assign(x, c)
    s = orc.kmeans_shift(2, 2.0)
    sums, counts, _ = orc.kmeans_accumulate(x, a, 3, s)
    c2 = orc.kmeans_update(sums, counts, s, c)
    assert counts.tolist() == [1, 1, 0]
    assert np.array_equal(c2, c)


def test_gmm_generator(pa):
    x = pa.gen_gmm(10000, 4, 3, sigma=1.0, lo=-10, hi=10, seed=5)
    y = pa.gen_gmm(10000, 4, 3, sigma=1.0, lo=-10, hi=10, seed=5)
    assert np.array_equal(x, y)
    assert x.dtype == np.float32 and x.shape == (10000, 4)
    assert np.max(np.abs(x)) < 16


def test_update_is_the_cluster_mean(pa, orc):
    # the update divides the fixed-point sums by the integer counts: within
    # the fixed-point and FP32 rounding of the FP64 cluster means (a count
    # off by one would miss by ~1/count)
    x = pa.gen_gmm(30000, 6, 5, sigma=1.0, seed=11)
    rng = np.random.default_rng(12)
    k = 7
    a = rng.integers(0, k - 1, len(x)).astype(np.int32)  # cluster k-1 stays empty
    maxabs = float(np.max(np.abs(x)))
    s = orc.kmeans_shift(len(x), maxabs)
    sums, counts, _ = orc.kmeans_accumulate(x, a, k, s)
    c0 = rng.normal(size=(k, 6)).astype(np.float32)
    c = orc.kmeans_update(sums, counts, s, c0)
    for j in range(k - 1):
        m = x[a == j].astype(np.float64).mean(axis=0)
        tol = np.abs(m) * 2.0 ** -24 + 2.0 ** (-s - 1) + 1e-300
        assert np.all(np.abs(c[j] - m) <= tol), j
        assert counts[j] == np.sum(a == j)
    assert np.array_equal(c[k - 1], c0[k - 1])


def test_lloyd_matches_fp64_lloyd(pa, orc):
    # a separated mixture: FP32 Lloyd (fixed-point sums) and FP64 Lloyd take
    # the same assignments and their centres agree to FP32 rounding
    x = pa.gen_gmm(40000, 8, 6, sigma=0.5, lo=-10, hi=10, seed=21)
    s = orc.kmeans_shift(len(x), float(np.max(np.abs(x))))
    c32, a32 = orc.kmeans_lloyd(x, 6, 8, s)[:2]
    c64, a64 = orc.kmeans_lloyd_f64(x, 6, 8)[:2]
    assert np.array_equal(a32, a64)
    rel = np.max(np.abs(c32 - c64)) / np.max(np.abs(c64))
    print(f"Lloyd FP32 vs FP64 after 8 iterations: {rel:.2e} relative")
    assert rel <= 1e-6
