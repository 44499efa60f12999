"""GPU parity of the stencil engine (through libpattern's C ABI) vs the oracle.

Bar: bit-exact against the FP32 oracle (same rule, same operation order, no
contraction on either side).  The FP64 error is reported for the heat3D
workload (128^3, seed 42, 100 steps).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_dev(pa, g, nsteps, **kw):
    nz, ny, nx = (s - 2 for s in g.shape)
    a = pa.DeviceArray.from_numpy(g)
    b = a.clone()
    out = pa.heat3d_run_dev(a, b, nx, ny, nz, nsteps, **kw)
    pa.sync()
    return out.numpy()


@pytest.mark.parametrize("shape", [(1, 1, 1), (63, 7, 5), (65, 9, 17), (130, 3, 2), (64, 16, 33), (200, 33, 12)])
def test_lds25d_bitexact_ragged(pa, orc, gpu, shape):
    nx, ny, nz = shape
    g = pa.gen_heat3d(nx, ny, nz, seed=11)
    ref = orc.heat3d(g, 3)
    got = run_dev(pa, g, 3, variant=pa.PAT_STENCIL_LDS25D)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("bx,by,zc", [(64, 4, 1), (64, 8, 3), (64, 16, 8), (64, 8, 64), (32, 16, 5), (128, 4, 2),
                                      (128, 4, 7)])
def test_lds25d_tiles(pa, orc, gpu, bx, by, zc):
    nx, ny, nz = 97, 41, 23
    g = pa.gen_heat3d(nx, ny, nz, seed=12)
    ref = orc.heat3d(g, 4)
    got = run_dev(pa, g, 4, variant=pa.PAT_STENCIL_LDS25D, bx=bx, by=by, zc=zc)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 4), (60, 12, 5), (61, 13, 17), (130, 29, 2), (64, 57, 33),
                                   (200, 33, 12)])
@pytest.mark.parametrize("steps", [1, 2, 5])
def test_temporal2_bitexact_ragged(pa, orc, gpu, shape, steps):
    """Two fused sweeps per pass: every window and chunk edge, odd step counts."""
    nx, ny, nz = shape
    g = pa.gen_heat3d(nx, ny, nz, seed=21)
    ref = orc.heat3d(g, steps)
    for by, zc in ((16, 1), (32, 3), (16, 64), (8, 2), (8, 5)):
        got = run_dev(pa, g, steps, variant=pa.PAT_STENCIL_TEMPORAL2, bx=64, by=by, zc=zc)
        assert np.array_equal(got, ref), (by, zc)


def test_temporal2_handle_replay(pa, orc, gpu):
    nx, ny, nz = 70, 50, 40
    g = pa.gen_heat3d(nx, ny, nz, seed=15)
    s = pa.Stencil(nx, ny, nz)
    s.set_variant(pa.PAT_STENCIL_TEMPORAL2, 64, 32, 8)
    s.upload(g)
    s.run(7)
    s.run(7)  # replays the captured graph from the other buffer
    s.run(6)
    tot, kms = s.run_timed(3)
    got = s.download()
    s.close()
    assert len(kms) == 3 and np.all(kms > 0)
    assert np.array_equal(got, orc.heat3d(g, 23))


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 4), (15, 17, 16), (16, 16, 16), (33, 9, 31), (61, 13, 17),
                                   (130, 29, 2)])
@pytest.mark.parametrize("steps", [1, 3, 4, 9])
def test_cubes_bitexact_ragged(pa, orc, gpu, shape, steps):
    """Several sweeps per launch on LDS blocks: every block edge, grids
    smaller than one block, step counts that leave sweeps over for the
    plain kernel."""
    nx, ny, nz = shape
    g = pa.gen_heat3d(nx, ny, nz, seed=23)
    ref = orc.heat3d(g, steps)
    for cube in range(5):
        got = run_dev(pa, g, steps, variant=pa.PAT_STENCIL_CUBES, bx=cube)
        assert np.array_equal(got, ref), cube


def test_cubes_handle_replay_and_errors(pa, orc, gpu):
    nx, ny, nz = 70, 50, 40
    g = pa.gen_heat3d(nx, ny, nz, seed=16)
    s = pa.Stencil(nx, ny, nz)
    with pytest.raises(RuntimeError):
        s.set_variant(pa.PAT_STENCIL_CUBES, 5, 0, 0)  # no such shape
    s.set_variant(pa.PAT_STENCIL_CUBES, 2, 0, 0)  # 16^3 blocks, 4 sweeps per launch
    s.upload(g)
    s.run(8)
    s.run(8)  # replays the captured graph from the other buffer
    s.run(5)
    tot, kms = s.run_timed(6)
    got = s.download()
    s.close()
    assert len(kms) == 6 and np.all(kms > 0)
    assert np.array_equal(got, orc.heat3d(g, 27))


def test_naive_bitexact(pa, orc, gpu):
    g = pa.gen_heat3d(45, 30, 20, seed=13)
    ref = orc.heat3d(g, 4)
    got = run_dev(pa, g, 4, variant=pa.PAT_STENCIL_GENERIC)
    assert np.array_equal(got, ref)


def test_heat3d_workload_128(pa, orc, gpu):
    """heat3D workload: 128^3, seed 42, 100 steps; bit-exact vs FP32 oracle."""
    n = 128
    g = pa.gen_heat3d(n, n, n, seed=42)
    ref = orc.heat3d(g, 100, threads=8)
    s = pa.Stencil(n, n, n, alpha=0.1)
    s.upload(g)
    s.run(100)
    got = s.download()
    s.close()
    assert np.array_equal(got, ref)
    ref64 = orc.heat3d(g.astype(np.float64), 100, dtype=np.float64)
    err = np.max(np.abs(got - ref64)) / np.max(np.abs(ref64))
    print(f"heat3d 128^3 x100: Linf={np.max(np.abs(got)):.6e} rel err vs FP64 {err:.3e}")
    assert err < 1e-5


def test_handle_graph_replay(pa, orc, gpu):
    nx, ny, nz = 70, 50, 40
    g = pa.gen_heat3d(nx, ny, nz, seed=14)
    s = pa.Stencil(nx, ny, nz)
    s.upload(g)
    s.run(7)
    s.run(7)  # replays the captured graph from the other buffer
    s.run(6)
    got = s.download()
    s.close()
    assert np.array_equal(got, orc.heat3d(g, 20))


def test_graph_timed_run_matches_run(pa, orc, gpu):
    """The timed replay advances the grid like run() (7 + 7 + 6 sweeps, so
    graphs start from both buffers) and reports a positive device time."""
    nx, ny, nz = 70, 50, 40
    g = pa.gen_heat3d(nx, ny, nz, seed=15)
    s = pa.Stencil(nx, ny, nz)
    s.upload(g)
    ms = [s.run_graph_timed(7), s.run_graph_timed(7), s.run_graph_timed(6)]
    got = s.download()
    s.close()
    assert all(t > 0 for t in ms)
    assert np.array_equal(got, orc.heat3d(g, 20))


def test_eigenmode_full_size_384(pa, gpu):
    """384^3 (over the 256 MB last-level cache): closed-form decay of a sine mode."""
    n, steps, alpha = 384, 10, 0.1
    k = np.sin(np.pi * np.arange(n + 2) / (n + 1))
    k[0] = k[-1] = 0.0
    g0 = (k[:, None, None] * k[None, :, None]) * k[None, None, :]
    g = pa.DeviceArray.from_numpy(g0.astype(np.float32))
    b = g.clone()
    out = pa.heat3d_run_dev(g, b, n, n, n, steps, alpha)
    pa.sync()
    lam = 1 + alpha * 3 * (2 * np.cos(np.pi / (n + 1)) - 2)
    err = np.max(np.abs(out.numpy() - g0 * lam ** steps))
    assert err < 1e-5
