"""Committed oracle fixtures (tests/golden/oracle_fixtures.npz, written by
tests/golden/make_fixtures.py): the oracle must still produce them (CPU), and
the HIP paths must produce them bit for bit through the C ABI (GPU).

  heat3D: 8^3 interior, 10 steps;
  LJ: FCC 3^3 cells (108 atoms), serial-scatter forces over the pair list in
      its given order (minimum image, cutoff 2.5);
  kmeans: 1,000 x 8, K 16, one Lloyd pass (assignments, minimum distances,
      updated centres).
"""
import os

import numpy as np
import pytest

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_fixtures.npz")


@pytest.fixture(scope="module")
def fx():
    with np.load(FIX) as d:
        return {k: d[k] for k in d.files}


def test_oracle_reproduces_the_fixtures(orc, pa, fx):
    assert np.array_equal(orc.heat3d(fx["heat_in"], 10), fx["heat_out"])
    f, _ = orc.lj_force(fx["lj_pos"], fx["lj_edges"], float(fx["lj_box"]), 2.5)
    assert np.array_equal(f, fx["lj_force"])
    x, k = fx["km_x"], int(fx["km_k"])
    s, ds = pa.kmeans_shifts(len(x), x.shape[1], float(np.max(np.abs(x))))
    a, md = orc.kmeans_assign(x, x[:k].copy())
    sums, counts, _ = orc.kmeans_accumulate(x, a, k, s, mindist=md, dshift=ds)
    assert np.array_equal(a, fx["km_assign"]) and np.array_equal(md, fx["km_mindist"])
    assert np.array_equal(orc.kmeans_update(sums, counts, s, x[:k].copy()), fx["km_centres"])


def test_fixture_invariants(fx):
    """What the fixtures must satisfy whatever produced them: the heat halo
    stays zero, the LJ forces sum to ~0 (Newton's third law, up to FP32
    rounding), every point is assigned to a centre at its minimum distance."""
    h = fx["heat_out"]
    assert not h[0].any() and not h[-1].any() and not h[:, 0].any() and not h[:, :, -1].any()
    f = fx["lj_force"].astype(np.float64)
    assert np.max(np.abs(f.sum(0))) <= 1e-4 * np.max(np.abs(f))
    x, k, a = fx["km_x"].astype(np.float64), int(fx["km_k"]), fx["km_assign"]
    d = np.sqrt(((x[:, None, :] - x[None, :k, :]) ** 2).sum(-1))
    assert np.all(d[np.arange(len(x)), a] <= d.min(1) * (1 + 1e-6))


@pytest.mark.gpu
def test_hip_paths_reproduce_the_fixtures(pa, gpu, fx):
    g = fx["heat_in"]
    nz, ny, nx = (s - 2 for s in g.shape)
    a = pa.DeviceArray.from_numpy(g)
    b = a.clone()
    out = pa.heat3d_run_dev(a, b, nx, ny, nz, 10)
    pa.sync()
    assert np.array_equal(out.numpy(), fx["heat_out"])

    pos, edges, box = fx["lj_pos"], fx["lj_edges"], float(fx["lj_box"])
    ir = pa.IrregularRuntime(edges, len(pos), pa.PAT_IR_LJ_MINIMAGE, np.array([box, 2.5], dtype=np.float64))
    ir.set_row_order(pa.PAT_IR_ORDER_EDGE)  # sums in the list's own order, as the fixture's scatter
    ir.reset_node_data(pos.astype(np.float64))
    ir.start()
    got = ir.get_reduction_result()
    ir.close()
    assert np.array_equal(got, fx["lj_force"])

    for mode in (pa.PAT_KM_PLAIN, pa.PAT_KM_MFMA):
        km = pa.KMeans(fx["km_x"], int(fx["km_k"]), mode=mode)
        km.run(1)
        c1, asg, md = km.get()
        km.close()
        assert np.array_equal(asg, fx["km_assign"]) and np.array_equal(md, fx["km_mindist"])
        assert np.array_equal(c1, fx["km_centres"])
