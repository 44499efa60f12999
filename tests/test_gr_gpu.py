"""GPU parity of the additive generalized reductions (pat_gr_*: velocity
statistics, kmeans/regular.h:30-73; pair energy, miniMD/regular.h:139-181)
against the oracle.

Bar: counts exact; FP64 sums within 1e-12 relative of the oracle's serial
sums (the device adds in a fixed tree order, so the bits differ from a
serial loop but are the same every run -- checked); per-unit values are FP32
(velocity) or FP64 (energy) values computed with the functor's arithmetic.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("n", [1, 63, 256, 1000, 100003, 1 << 20])
def test_vstats_vs_oracle(pa, orc, gpu, n):
    rng = np.random.default_rng(n)
    v = rng.normal(size=(n, 3)) * 1.3
    vaverh = float(np.float32(1.9))
    gr = pa.GR(pa.PAT_GR_VSTATS, v, vaverh)
    gr.run()
    vals, em = gr.output()
    gr.run()
    vals2, _ = gr.output()
    gr.close()
    s, c = orc.gr_vstats(v, vaverh)
    assert vals.shape == (1, 2) and em.tolist() == [n]
    assert vals[0, 1] == c  # the count is exact
    assert rel(vals[0, 0], s) < 1e-12
    assert np.array_equal(vals, vals2)  # same bits every run


def test_vstats_threshold_and_exact_lengths(pa, orc, gpu):
    v = np.array([[3, 4, 0], [2, 3, 6], [1, 4, 8], [2, 6, 9], [0, 0, 0], [-6, 2, -3]] * 100, dtype=np.float64)
    gr = pa.GR(pa.PAT_GR_VSTATS, v, 7.0)
    gr.run()
    vals, _ = gr.output()
    gr.set_param(8.9999)
    gr.run()
    vals2, _ = gr.output()
    gr.close()
    assert vals[0].tolist() == [3900.0, 200.0]
    assert vals2[0].tolist() == [3900.0, 200.0]


@pytest.mark.parametrize("cells", [3, 8])
def test_lj_energy_vs_oracle(pa, orc, gpu, cells):
    pos, box, edges = pa.moldyn_system(cells, seed=4)
    x = pos.astype(np.float64)
    cutsq = 2.5 * 2.5
    gr = pa.GR(pa.PAT_GR_LJ_ENERGY, edges, (cutsq, x))
    gr.run()
    vals, em = gr.output()
    # a new parameter (moved positions) on the same handle
    x2 = x + 0.01
    gr.set_param((cutsq * 0.9, x2))
    gr.run()
    vals2, em2 = gr.output()
    gr.close()
    s, n = orc.gr_lj_energy(edges, x, cutsq)
    s2, n2 = orc.gr_lj_energy(edges, x2, cutsq * 0.9)
    print(f"pair energy {cells}^3: {vals[0, 0]:.12e} vs oracle {s:.12e} ({em[0]} pairs in the cutoff)")
    assert em[0] == n and em2[0] == n2
    assert rel(vals[0, 0], s) < 1e-12 and rel(vals2[0, 0], s2) < 1e-12


def test_gr_rejects_bad_arguments(pa, gpu):
    v = np.zeros((10, 3))
    with pytest.raises(RuntimeError, match="map index 1"):
        pa.GR(1, v, (1.0, np.zeros((1, 3))))
    e = np.array([[0, 5]], np.int32)
    with pytest.raises(RuntimeError, match="outside"):
        pa.GR(pa.PAT_GR_LJ_ENERGY, e, (1.0, np.zeros((3, 3))))
    gr = pa.GR(pa.PAT_GR_VSTATS, v, 1.0)
    with pytest.raises(RuntimeError, match="reduce"):
        pa.call("pat_gr_set_reduce_idx", gr.h, 2)
    gr.close()
