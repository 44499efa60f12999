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


# ---- the W = 3 star (kmeans/stencil.h stencil2) and heat_3D's FP64 functor


def test_star_point_source(orc):
    """One step from a unit impulse: exactly the 12 cells 1 and 2 away along
    the axes get 1/36 (the impulse itself and every other cell 0)."""
    w, n = 3, 9
    g = np.zeros((n + 2 * w,) * 3, dtype=np.float32)
    c = w + n // 2
    g[c, c, c] = 1.0
    out = orc.star(g, 1, w=w)
    expect = np.zeros_like(g)
    for m in (1, 2):
        for d in range(3):
            for sgn in (1, -1):
                idx = [c, c, c]
                idx[d] += sgn * m
                expect[tuple(idx)] = np.float32(1.0) / np.float32(36.0)
    assert np.array_equal(out, expect)


def test_star_constant_field(orc):
    """A constant interior of ones: cells at least 2 from every face see 12
    ones (12/36 rounded once); a cell on a face sees 10, an edge 8, a corner
    6 (the halo is 0)."""
    w, n = 3, 12
    g = np.zeros((n + 2 * w,) * 3, dtype=np.float32)
    g[w:-w, w:-w, w:-w] = 1.0
    out = orc.star(g, 1, w=w)[w:-w, w:-w, w:-w]
    assert np.all(out[2:-2, 2:-2, 2:-2] == np.float32(12.0) / np.float32(36.0))
    assert out[0, 5, 5] == np.float32(10.0) / np.float32(36.0)
    assert out[0, 0, 5] == np.float32(8.0) / np.float32(36.0)
    assert out[0, 0, 0] == np.float32(6.0) / np.float32(36.0)
    assert out[1, 1, 1] == np.float32(9.0) / np.float32(36.0)


def test_star_linear_and_symmetric(orc):
    rng = np.random.default_rng(3)
    w = 3
    a = np.zeros((14, 15, 16), dtype=np.float32)
    a[w:-w, w:-w, w:-w] = rng.random((8, 9, 10))
    b = np.zeros_like(a)
    b[w:-w, w:-w, w:-w] = rng.random((8, 9, 10))
    sa, sb, sab = orc.star(a, 2, w=w), orc.star(b, 2, w=w), orc.star(a + b, 2, w=w)
    assert np.max(np.abs(sab - (sa + sb))) < 1e-6
    # reflecting the grid along x reflects the result (the sum's order
    # differs, so within rounding)
    assert np.max(np.abs(orc.star(a[:, :, ::-1].copy(), 2, w=w) - sa[:, :, ::-1])) < 1e-6


def test_heat_ref_point_source(orc):
    """One step from a unit impulse: the cell itself gets 3 c^2 (c = 0.234,
    one c^2 per axis), its 6 face neighbours c, all else 0."""
    c = 0.234
    g = np.zeros((9, 9, 9))
    g[4, 4, 4] = 1.0
    out = orc.heat_ref(g, 1)
    assert out[4, 4, 4] == c * c + c * c + c * c
    for idx in ((3, 4, 4), (5, 4, 4), (4, 3, 4), (4, 5, 4), (4, 4, 3), (4, 4, 5)):
        assert abs(out[idx] - c) < 1e-16
    out[4, 4, 4] = 0
    for idx in ((3, 4, 4), (5, 4, 4), (4, 3, 4), (4, 5, 4), (4, 4, 3), (4, 4, 5)):
        out[idx] = 0
    assert not out.any()


def test_heat_ref_eigenmode(orc):
    """The discrete sine mode with the zero halo is an eigenvector: each step
    multiplies it by 2c sum_d cos(pi p_d / (M_d + 1)) + 3 c^2."""
    c = 0.234
    M = (10, 12, 9)  # z, y, x interior
    p = (2, 1, 3)
    axes = [np.sin(np.pi * pd * np.arange(Md + 2) / (Md + 1)) for pd, Md in zip(p, M)]
    g = axes[0][:, None, None] * axes[1][None, :, None] * axes[2][None, None, :]
    lam = 2 * c * sum(np.cos(np.pi * pd / (Md + 1)) for pd, Md in zip(p, M)) + 3 * c * c
    out = orc.heat_ref(g, 5)
    assert np.max(np.abs(out - lam ** 5 * g)) <= 1e-12 * np.max(np.abs(g))
