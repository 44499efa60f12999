"""Generalized reduction with an application's own functors (apps/regular,
compiled for the device) and keyed reduction objects: wavefront LDS hash
tables whose full inserts make the wavefront merge into its device table and
run the refused maps again, folded in a fixed order into the final object.

Bar: integer values (counts, keys) exact against the oracle; FP32 sums within
1e-5 and FP64 sums within 1e-12 relative of the FP64 oracle sums (the fold
order differs from a serial run); the same bytes on every run, also with the
LDS objects forced to merge after a few keys.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tab(pa):
    return pa.app_table("regular", "gr")


def u32(b):
    return int(np.frombuffer(b, np.uint32)[0])


def test_velocity_statistics_any_key_buckets(pa, orc, gpu, tab):
    """moldyn's map/reduce (kmeans/regular.h:30-73): one logical key spread
    over ANY_KEY's 41 buckets; their sums are the oracle's."""
    v = pa.gen_velocities(150000, 1.44, seed=3).astype(np.float64)
    vaverh = np.float32(np.mean(np.sqrt((v ** 2).sum(1))))
    r = pa.Regular(v, len(v), 24, table=tab)
    out = r.run(0, param=np.array([vaverh], np.float32))
    assert sorted(u32(k) for k in out) == list(range(41))
    vals = np.array([np.frombuffer(b, np.float32) for b in out.values()], np.float64)
    s, c = orc.gr_vstats(v, vaverh)
    assert vals[:, 1].sum() == c
    assert abs(vals[:, 0].sum() - s) <= 1e-5 * s
    assert r.run(0, param=np.array([vaverh], np.float32)) == out  # the same bytes again
    r.close()


def test_kmeans_map1_reference_sizes(pa, orc, gpu, tab):
    """kmeans map1/reduce1 at the reference's D = 3, K = 40: every centre's
    count is the oracle assignment's, its coordinate and distance sums within
    FP32 summation error of the FP64 sums."""
    n = 300000
    x = pa.gen_gmm(n, 3, 40, sigma=1.0, seed=5)
    c = x[:40].copy()
    r = pa.Regular(x, n, 12, table=tab)
    out = r.run(1, param=c)
    r.close()
    a, md = orc.kmeans_assign(x, c)
    cnt = np.bincount(a, minlength=40)
    assert sorted(u32(k) for k in out) == [k for k in range(40) if cnt[k]]
    for kb, vb in out.items():
        k = u32(kb)
        v = np.frombuffer(vb, np.float32).astype(np.float64)
        assert v[3] == cnt[k]
        sel = a == k
        exact = np.concatenate([x[sel].astype(np.float64).sum(0), [md[sel].astype(np.float64).sum()]])
        scale = np.abs(x[sel]).astype(np.float64).sum(0).max()
        assert np.max(np.abs(v[[0, 1, 2, 4]] - exact)) <= 1e-5 * max(scale, exact[3]), k


def test_pair_energy(pa, orc, gpu, tab):
    """miniMD's map2/reduce2 over 41 buckets: the FP64 oracle sum."""
    pos, box = pa.gen_fcc(10, 10, 10, jitter=0.05, seed=6)
    e = pa.gen_pairs(pos, box[0], 2.8)
    prm = np.concatenate([[6.25], pos.astype(np.float64).ravel()])
    r = pa.Regular(e, len(e), 8, table=tab)
    out = r.run(2, param=prm)
    r.close()
    total = sum(np.frombuffer(b, np.float64)[0] for b in out.values())
    ref, _ = orc.gr_lj_energy(e, pos, 6.25)
    assert abs(total - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("lds_keys", [0, 4, 1])
def test_many_key_histogram_offsets_and_merges(pa, gpu, tab, lds_keys):
    """5,003 keys from 2,000,000 values placed by an Offset array with gaps;
    with the LDS objects limited to 4 keys or 1 the merge-and-run-again path
    runs thousands of times and the counts stay exact, the FP64 sums within
    1e-12."""
    n = 2000000
    rng = np.random.default_rng(7)
    vals = rng.integers(0, 2 ** 31, n, dtype=np.uint32)
    buf = np.full(3 * n, 0xdeadbeef, np.uint32)
    buf[1::3] = vals
    off = np.arange(n, dtype=np.int64) * 12 + 4
    r = pa.Regular(buf, n, 4, offsets=off, table=tab)
    r.set_capacity(8192, 32, lds_keys)
    out = r.run(3)
    merges, reruns, waves = r.stats()
    assert r.run(3) == out  # deterministic
    r.close()
    keys = vals % 5003
    cnt = np.bincount(keys, minlength=5003)
    sums = np.zeros(5003)
    np.add.at(sums, keys, vals.astype(np.float64) / 1000.0)
    assert len(out) == 5003
    for kb, vb in out.items():
        k = u32(kb)
        c = np.frombuffer(vb[:8], np.int64)[0]
        s = np.frombuffer(vb[8:], np.float64)[0]
        assert c == cnt[k]
        assert abs(s - sums[k]) <= 1e-12 * sums[k]
    print(f"histogram lds_keys={lds_keys}: {merges} LDS merges, {reruns} maps run again, {waves} wavefronts")
    if lds_keys:
        assert merges > waves and reruns > 0


@pytest.mark.parametrize("lds_keys", [0, 3])
def test_odd_sized_keys_and_values(pa, gpu, tab, lds_keys):
    """6-byte keys and 3-byte values (the objects' byte-wise copy and
    compare, keys padded to whole words): every key's three bytes equal the
    byte-wise sums modulo 256 of its units, also with merges forced."""
    n = 300000
    v = np.random.default_rng(11).integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    r = pa.Regular(v, n, 4, table=tab)
    r.set_capacity(512, 16, lds_keys)
    out = r.run(4, reduce_idx=4)
    assert r.run(4, reduce_idx=4) == out
    r.close()
    k = np.stack([v % 7, v // 7 % 11, v // 77 % 3], 1).astype(np.uint16)
    flat = k[:, 0].astype(np.int64) * 33 + k[:, 1].astype(np.int64) * 3 + k[:, 2]
    want = {}
    for col, vals in ((0, np.ones(n, np.int64)), (1, (v & 255).astype(np.int64)), (2, (v >> 8 & 255).astype(np.int64))):
        sums = np.bincount(flat, weights=vals, minlength=231).astype(np.int64) % 256
        for f in np.unique(flat):
            want.setdefault(int(f), [0, 0, 0])[col] = int(sums[f])
    assert len(out) == len(want) == 231
    for kb, vb in out.items():
        a, b, c = np.frombuffer(kb, np.uint16)
        assert len(kb) == 6 and len(vb) == 3
        assert list(vb) == want[int(a) * 33 + int(b) * 3 + int(c)]


def test_capacity_errors(pa, gpu, tab):
    vals = np.arange(100000, dtype=np.uint32)
    r = pa.Regular(vals, len(vals), 4, table=tab)
    r.set_capacity(1000, 32)  # 5,003 keys do not fit
    with pytest.raises(pa.PatternError, match="keys"):
        r.run(3)
    with pytest.raises(pa.PatternError, match="not in the table"):
        pa.call("pat_regular_set_map_idx", r.h, 5)
    r.close()
