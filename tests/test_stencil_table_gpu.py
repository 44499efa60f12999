"""Application stencil functors run on the device through their compiled
tables (apps/<app>/gpu_table.hip -> apps/bin/lib<app>_table.so), through the
C ABI's pat_stencil_set_table: the heat3D app's table (the workload rule and
heat_3D's own FP64 functor) and the kmeans directory's W = 3 star.

Bar: bit-exact against the oracle's restatement of each functor (grids of
the functor's element type and halo width, several steps, ragged shapes);
the workload rule through the table equals the tuned engine bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def seeded(pa, nx, ny, nz, w, dtype, seed):
    g = np.zeros((nz + 2 * w, ny + 2 * w, nx + 2 * w), dtype=dtype)
    g[w:-w, w:-w, w:-w] = np.random.default_rng(seed).random((nz, ny, nx))
    return g


@pytest.mark.parametrize("shape", [(40, 33, 21), (64, 64, 64), (7, 5, 3)])
def test_heat_rule_through_table_equals_tuned_and_oracle(pa, orc, gpu, shape):
    nx, ny, nz = shape
    g = pa.gen_heat3d(nx, ny, nz, seed=5)
    tab = pa.app_table("heat3d")
    st = pa.Stencil(nx, ny, nz, alpha=0.1, table=tab, idx=0)
    st.set_variant(pa.PAT_STENCIL_GENERIC)  # the application's own kernel, not the tuned engine
    st.upload(g)
    st.run(7)
    via_table = st.download()
    st.close()
    tuned = pa.Stencil(nx, ny, nz, alpha=0.1, table=tab, idx=0)  # same functor, tuned engine
    tuned.upload(g)
    tuned.run(7)
    assert np.array_equal(via_table, tuned.download())
    tuned.close()
    assert np.array_equal(via_table, orc.heat3d(g, 7))


@pytest.mark.parametrize("shape,steps", [((30, 28, 26), 5), ((96, 80, 72), 3), ((5, 4, 3), 2)])
def test_heat_ref_fp64_vs_oracle(pa, orc, gpu, shape, steps):
    """heat_3D's FP64 functor (StencilCreator<double>, table index 1)."""
    nx, ny, nz = shape
    g = seeded(pa, nx, ny, nz, 1, np.float64, 11)
    st = pa.Stencil(nx, ny, nz, width=1, dtype=np.float64, param=np.zeros(5), table=pa.app_table("heat3d"), idx=1)
    st.upload(g)
    st.run(steps)
    got = st.download()
    st.close()
    assert np.array_equal(got, orc.heat_ref(g, steps))


@pytest.mark.parametrize("shape,steps", [((40, 36, 33), 4), ((128, 128, 128), 3), ((3, 4, 5), 3)])
def test_star_w3_vs_oracle(pa, orc, gpu, shape, steps):
    """The W = 3 star on a 3-cell halo, run as the runtime's replayed sweep
    graph and launched one by one."""
    nx, ny, nz = shape
    g = seeded(pa, nx, ny, nz, 3, np.float32, 12)
    ref = orc.star(g, steps, w=3)
    st = pa.Stencil(nx, ny, nz, width=3, param=np.zeros(0, np.float32), table=pa.app_table("stencil2"), idx=0)
    st.upload(g)
    st.run(steps)
    assert np.array_equal(st.download(), ref)
    st.upload(g)
    _, kms = st.run_timed(steps)
    assert np.array_equal(st.download(), ref) and np.all(kms > 0)
    st.close()


@pytest.mark.parametrize("shape,steps", [((37, 23), 1), ((130, 61), 4), ((1, 1), 2)])
def test_cross2d_on_a_2d_grid(pa, gpu, shape, steps):
    """The runtime's 2-D path: a 2-D table functor on (ny + 2) x (nx + 2)
    grids (halo fixed at zero), several steps, against the same FP32
    operations in NumPy (no contraction: each sum and the product rounded as
    written)."""
    nx, ny = shape
    g = np.zeros((ny + 2, nx + 2), np.float32)
    g[1:-1, 1:-1] = np.random.default_rng(9).standard_normal((ny, nx)).astype(np.float32)
    ref = g.copy()
    for _ in range(steps):
        nxt = np.zeros_like(ref)
        a = ref[1:-1, 2:] + ref[1:-1, :-2]
        b = ref[2:, 1:-1] + ref[:-2, 1:-1]
        nxt[1:-1, 1:-1] = (a + b) * np.float32(0.25)
        ref = nxt
    st = pa.Stencil(nx, ny, None, width=1, param=np.zeros(0, np.float32), table=pa.app_table("stencil2"), idx=1)
    st.upload(g)
    st.run(steps)
    got = st.download()
    st.close()
    assert got.shape == ref.shape and np.array_equal(got, ref)


def test_table_errors(pa, gpu):
    # an index outside the table, a functor of another element type, and a
    # width the tuned engine cannot run without a table are refused
    tab = pa.app_table("stencil2")
    st = pa.Stencil(8, 8, 8, width=3, param=np.zeros(0, np.float32))
    st.upload(np.zeros(st.shape, np.float32))
    with pytest.raises(pa.PatternError, match="table"):
        st.run(1)
    call = pa.call
    call("pat_stencil_set_table", st.h, tab)
    with pytest.raises(pa.PatternError, match="not in the table"):
        call("pat_stencil_set_idx", st.h, 2)
    st.close()
    st = pa.Stencil(8, 8, 8, width=1, param=np.zeros(1), table=pa.app_table("heat3d"), idx=1)  # FP64 functor
    st.upload(np.zeros(st.shape, np.float32))
    with pytest.raises(pa.PatternError, match="byte elements"):
        st.run(1)
    st.close()
