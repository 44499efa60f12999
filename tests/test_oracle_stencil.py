"""Pins the stencil oracle (oracle/oracle.c heat3d_*) without trusting it.

Checks: the discrete sine eigenmode decays by its closed-form factor, the
7-point star addresses the neighbours GET_FLOAT3 (heat_3D/lib/macro.h:70)
defines (x fastest, then y, then z), linearity, axis-reflection symmetry,
Dirichlet halo, agreement with an independent FP64 slicing formulation, and
that the tiled threaded CPU path equals the serial path bit for bit.
"""
import numpy as np
import pytest


def eigenmode(nx, ny, nz, m=(1, 1, 1)):
    g = np.zeros((nz + 2, ny + 2, nx + 2))
    ix = np.sin(np.pi * m[0] * np.arange(nx + 2) / (nx + 1))
    iy = np.sin(np.pi * m[1] * np.arange(ny + 2) / (ny + 1))
    iz = np.sin(np.pi * m[2] * np.arange(nz + 2) / (nz + 1))
    g[:] = iz[:, None, None] * iy[None, :, None] * ix[None, None, :]
    g[0, :, :] = g[-1, :, :] = 0
    g[:, 0, :] = g[:, -1, :] = 0
    g[:, :, 0] = g[:, :, -1] = 0
    return g


def lam(nx, ny, nz, m, alpha):
    return 1 + alpha * sum(2 * np.cos(np.pi * mi / (n + 1)) - 2 for mi, n in zip(m, (nx, ny, nz)))


@pytest.mark.parametrize("m", [(1, 1, 1), (2, 3, 1), (5, 1, 4)])
def test_eigenmode_decay_f64(orc, m):
    nx, ny, nz, alpha, steps = 12, 9, 7, 0.1, 6
    g = eigenmode(nx, ny, nz, m)
    out = orc.heat3d(g, steps, alpha, dtype=np.float64)
    expect = g * lam(nx, ny, nz, m, alpha) ** steps
    assert np.max(np.abs(out - expect)) < 1e-13


def test_eigenmode_decay_f32(orc):
    nx, ny, nz, alpha, steps = 16, 12, 10, 0.1, 20
    g = eigenmode(nx, ny, nz, (1, 2, 1)).astype(np.float32)
    out = orc.heat3d(g, steps, alpha)
    expect = g.astype(np.float64) * lam(nx, ny, nz, (1, 2, 1), alpha) ** steps
    assert np.max(np.abs(out - expect)) < 2e-6


def test_neighbour_addressing(orc):
    # a unit source at interior (x, y, z) = (3, 2, 4) of an anisotropic grid:
    # after one step the six faces hold alpha and the centre 1 - 6 alpha
    nx, ny, nz, alpha = 7, 5, 6, 0.125
    g = np.zeros((nz + 2, ny + 2, nx + 2), dtype=np.float32)
    x, y, z = 3, 2, 4
    g[z, y, x] = 1.0
    out = orc.heat3d(g, 1, alpha)
    expect = np.zeros_like(g)
    expect[z, y, x] = 1 - 6 * alpha
    for dz, dy, dx in ((0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)):
        expect[z + dz, y + dy, x + dx] = alpha
    assert np.array_equal(out, expect)
    # flat-index view: x is +-1, y is +-(nx+2), z is +-(nx+2)(ny+2)
    flat = out.ravel()
    c = x + (nx + 2) * (y + (ny + 2) * z)
    for off in (1, nx + 2, (nx + 2) * (ny + 2)):
        assert flat[c + off] == np.float32(alpha) and flat[c - off] == np.float32(alpha)


def independent_step(u, alpha):
    # slicing formulation in FP64: a different indexing path from the C loop
    v = u.copy()
    c = u[1:-1, 1:-1, 1:-1]
    s = (u[1:-1, 1:-1, :-2] + u[1:-1, 1:-1, 2:] + u[1:-1, :-2, 1:-1] + u[1:-1, 2:, 1:-1] + u[:-2, 1:-1, 1:-1]
         + u[2:, 1:-1, 1:-1])
    v[1:-1, 1:-1, 1:-1] = c + alpha * (s - 6 * c)
    return v


def test_matches_independent_fp64(orc, pa):
    nx, ny, nz = 11, 6, 9
    g = pa.gen_heat3d(nx, ny, nz, seed=5).astype(np.float64)
    u = g.copy()
    for _ in range(4):
        u = independent_step(u, 0.1)
    out = orc.heat3d(g, 4, 0.1, dtype=np.float64)
    assert np.max(np.abs(out - u)) < 1e-14
    out32 = orc.heat3d(g.astype(np.float32), 4, 0.1)
    assert np.max(np.abs(out32 - u)) < 1e-6


def test_linearity_and_symmetry(orc, pa):
    nx, ny, nz = 10, 8, 6
    a = pa.gen_heat3d(nx, ny, nz, seed=1).astype(np.float64)
    b = pa.gen_heat3d(nx, ny, nz, seed=2).astype(np.float64)
    la = orc.heat3d(a, 3, 0.1, dtype=np.float64)
    lb = orc.heat3d(b, 3, 0.1, dtype=np.float64)
    lab = orc.heat3d(2 * a - 3 * b, 3, 0.1, dtype=np.float64)
    assert np.max(np.abs(lab - (2 * la - 3 * lb))) < 1e-12
    # reflecting x (and z) commutes with the sweep
    r = orc.heat3d(a[::-1, :, ::-1].copy(), 3, 0.1, dtype=np.float64)
    assert np.max(np.abs(r - la[::-1, :, ::-1])) < 1e-13


def test_halo_stays_zero_and_zero_field(orc, pa):
    g = pa.gen_heat3d(9, 7, 5, seed=3)
    out = orc.heat3d(g, 7)
    assert not out[0].any() and not out[-1].any()
    assert not out[:, 0].any() and not out[:, -1].any()
    assert not out[:, :, 0].any() and not out[:, :, -1].any()
    z = np.zeros_like(g)
    assert not orc.heat3d(z, 3).any()


def test_tiled_threads_equal_serial(orc, pa):
    g = pa.gen_heat3d(37, 20, 18, seed=9)
    s = orc.heat3d(g, 5)
    t = orc.heat3d(g, 5, threads=4)
    assert np.array_equal(s, t)


def test_generator_seeded(pa):
    a = pa.gen_heat3d(8, 8, 8, seed=42)
    b = pa.gen_heat3d(8, 8, 8, seed=42)
    c = pa.gen_heat3d(8, 8, 8, seed=43)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    inner = a[1:-1, 1:-1, 1:-1]
    assert inner.min() >= 0 and inner.max() < 1
    assert not a[0].any() and not a[:, :, -1].any()
