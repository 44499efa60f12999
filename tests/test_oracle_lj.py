"""Pins the LJ oracle (oracle/oracle.c lj_*, md_*, pairs_brute) and the
workload generators without trusting the oracle.

Checks: the two-particle closed form F(r) = 48(r^-13 - r^-7/2) and
phi(r) = 4(r^-12 - r^-6) (and F = -dphi/dr by finite differences), the
minimum image across the box boundary, zero net force on every atom of a
perfect FCC lattice, Newton's third law, invariance under edge order and node
renumbering, energy conservation of the velocity-Verlet loop, and the pair
generator against O(N^2) brute force.
"""
import numpy as np
import pytest


def two_atoms(r, axis=0, box=20.0):
    pos = np.zeros((2, 3))
    pos[0] = 5.0
    pos[1] = 5.0
    pos[0, axis] += r
    return pos, np.array([[0, 1]], dtype=np.int32)


@pytest.mark.parametrize("r", [0.95, 1.0, 2 ** (1 / 6), 1.5, 2.49])
def test_two_particle_closed_form(orc, r):
    pos, e = two_atoms(r)
    f, ep = orc.lj_force(pos, e, 20.0, 2.5, pbc=True, dtype=np.float64)
    F = 48 * (r ** -13 - 0.5 * r ** -7)
    assert abs(f[0, 0] - F) <= 1e-12 * max(1, abs(F))
    assert abs(f[1, 0] + F) <= 1e-12 * max(1, abs(F))
    assert f[0, 1] == 0 and f[0, 2] == 0
    assert abs(ep - 4 * (r ** -12 - r ** -6)) < 1e-12
    p32 = pos.astype(np.float32)
    f32, _ = orc.lj_force(p32, e, 20.0, 2.5, pbc=True)
    r32 = float(p32[0, 0]) - float(p32[1, 0])  # the distance the FP32 inputs encode
    F32 = 48 * (r32 ** -13 - 0.5 * r32 ** -7)
    assert abs(f32[0, 0] - F32) <= 2e-6 * max(1, abs(F32))


def test_force_is_minus_gradient(orc):
    r, h = 1.3, 1e-6
    _, ep_p = orc.lj_force(*two_atoms(r + h), 20.0, 2.5, dtype=np.float64)
    _, ep_m = orc.lj_force(*two_atoms(r - h), 20.0, 2.5, dtype=np.float64)
    f, _ = orc.lj_force(*two_atoms(r), 20.0, 2.5, dtype=np.float64)
    assert abs(f[0, 0] - (-(ep_p - ep_m) / (2 * h))) < 1e-6


def test_cutoff_excludes(orc):
    f, ep = orc.lj_force(*two_atoms(2.5), 20.0, 2.5, dtype=np.float64)
    assert not f.any() and ep == 0


def test_minimum_image(orc):
    L = 10.0
    pos = np.array([[0.5, 3, 3], [L - 0.6, 3, 3]])
    e = np.array([[0, 1]], dtype=np.int32)
    f, _ = orc.lj_force(pos, e, L, 2.5, pbc=True, dtype=np.float64)
    r = 1.1
    F = 48 * (r ** -13 - 0.5 * r ** -7)
    # atom 0 sees atom 1 at x = -0.6 through the boundary: d = +1.1
    assert abs(f[0, 0] - F) < 1e-12 and abs(f[1, 0] + F) < 1e-12
    fo, _ = orc.lj_force(pos, e, L, 2.5, pbc=False, dtype=np.float64)
    assert not fo.any()  # 8.9 apart without images


def test_pairs_generator_matches_brute_force(pa, orc):
    pos, box = pa.gen_fcc(3, 3, 3, rho=0.8442, jitter=0.05, seed=4)
    e = pa.gen_pairs(pos, box[0], 2.5)
    b = orc.pairs_brute(pos, box[0], 2.5)
    assert np.array_equal(e, b)  # both in (i, j) order
    es = pa.gen_pairs(pos, box[0], 2.5, shuffle=True, seed=9)
    assert not np.array_equal(es, e)
    assert np.array_equal(es[np.lexsort((es[:, 1], es[:, 0]))], e)


def test_pairs_generator_cell_list_path(pa, orc):
    # 4^3 cells: box 6.7 > 3*cutoff(2.2), so the cell-list path is used
    pos, box = pa.gen_fcc(4, 4, 4, rho=0.8442, jitter=0.05, seed=5)
    e = pa.gen_pairs(pos, box[0], 2.2)
    assert np.array_equal(e, orc.pairs_brute(pos, box[0], 2.2))


def test_fcc_lattice_structure(pa):
    pos, box = pa.gen_fcc(2, 3, 4, rho=0.8442, jitter=0.0)
    a = (4 / 0.8442) ** (1 / 3)
    assert pos.shape == (96, 3)
    assert np.allclose(box, [2 * a, 3 * a, 4 * a])
    # every coordinate is a multiple of a/2 with an even index sum
    idx = np.rint(pos / (a / 2)).astype(int)
    assert np.allclose(idx * (a / 2), pos, atol=1e-5)
    assert np.all(idx.sum(axis=1) % 2 == 0)
    assert len({tuple(r) for r in idx}) == 96
    # miniMD visit order: z outermost, then y, then x
    assert np.all(np.diff(idx[:, 2]) >= 0)


def test_perfect_fcc_zero_net_force(pa, orc):
    pos, box = pa.gen_fcc(4, 4, 4, rho=0.8442, jitter=0.0)
    e = pa.gen_pairs(pos, box[0], 2.5)
    # exact FP64 lattice sites (the generator's FP32 sites are off by ~1e-7)
    h = (4 / 0.8442) ** (1 / 3) / 2
    p64 = np.rint(pos.astype(np.float64) / h) * h
    f, _ = orc.lj_force(p64, e, 8 * h, 2.5, dtype=np.float64)
    assert np.max(np.abs(f)) < 1e-9
    # while a single displaced atom feels a restoring force
    p64[5, 0] += 0.01
    f, _ = orc.lj_force(p64, e, 8 * h, 2.5, dtype=np.float64)
    assert f[5, 0] < -0.1


def test_newton_third_law_and_invariances(pa, orc):
    pos, box = pa.gen_fcc(4, 4, 4, rho=0.8442, jitter=0.05, seed=6)
    e = pa.gen_pairs(pos, box[0], 2.5)
    p64 = pos.astype(np.float64)
    f, ep = orc.lj_force(p64, e, box[0], 2.5, dtype=np.float64)
    assert np.max(np.abs(f)) > 1.0
    assert np.max(np.abs(f.sum(axis=0))) < 1e-9
    rng = np.random.default_rng(0)
    # edge order and endpoint order
    e2 = e[rng.permutation(len(e))][:, ::-1].copy()
    f2, ep2 = orc.lj_force(p64, e2, box[0], 2.5, dtype=np.float64)
    assert np.max(np.abs(f2 - f)) < 1e-10 and abs(ep2 - ep) < 1e-9
    # node renumbering
    perm = rng.permutation(len(pos))
    inv = np.argsort(perm)
    f3, _ = orc.lj_force(p64[perm], inv[e].astype(np.int32), box[0], 2.5, dtype=np.float64)
    assert np.max(np.abs(f3 - f[perm])) < 1e-10


def test_f32_close_to_f64(pa, orc):
    pos, box = pa.gen_fcc(4, 4, 4, jitter=0.05, seed=7)
    e = pa.gen_pairs(pos, box[0], 2.5)
    f32, _ = orc.lj_force(pos, e, box[0], 2.5)
    f64, _ = orc.lj_force(pos.astype(np.float64), e, box[0], 2.5, dtype=np.float64)
    assert np.max(np.abs(f32 - f64)) / np.max(np.abs(f64)) < 1e-5
    ft = orc.lj_force_threads(pos, e, box[0], 2.5, threads=4)
    assert np.max(np.abs(ft - f64)) / np.max(np.abs(f64)) < 1e-5


def test_velocity_verlet_conserves_energy(pa, orc):
    pos, box = pa.gen_fcc(4, 4, 4, jitter=0.05, seed=8)
    vel = pa.gen_velocities(len(pos), 1.44, seed=3)
    e = pa.gen_pairs(pos, box[0], 2.8)
    p0 = pos.astype(np.float64)
    v0 = vel.astype(np.float64)
    _, _, _, ep0 = orc.md_run(p0, v0, e, box[0], 2.5, 0.005, 0, dtype=np.float64)
    e0 = ep0 + 0.5 * np.sum(v0 ** 2)
    p1, v1, _, ep1 = orc.md_run(p0, v0, e, box[0], 2.5, 0.005, 40, dtype=np.float64)
    e1 = ep1 + 0.5 * np.sum(v1 ** 2)
    # truncated (unshifted) LJ: small drift from pairs crossing the cutoff
    assert abs(e1 - e0) / abs(e0) < 5e-3
    # total momentum is conserved (it starts at ~1e-7 from FP32 rounding)
    assert np.max(np.abs(v1.sum(axis=0) - v0.sum(axis=0))) < 1e-9
    # time reversibility: run back with negated velocities
    p2, v2, _, _ = orc.md_run(p1, -v1, e, box[0], 2.5, 0.005, 40, dtype=np.float64)
    d = p2 - p0
    d -= box[0] * np.rint(d / box[0])
    assert np.max(np.abs(d)) < 1e-8


def test_velocities_generator(pa):
    v = pa.gen_velocities(4000, 1.44, seed=3)
    assert np.max(np.abs(v.astype(np.float64).sum(axis=0))) < 1e-3
    t = np.sum(v.astype(np.float64) ** 2) / (3 * (4000 - 1))
    assert abs(t - 1.44) < 1e-4


def test_gather_form_equals_serial_scatter(pa, orc):
    pos, box, edges = pa.moldyn_system(5, seed=3)  # shuffled edge order
    f_serial, _ = orc.lj_force(pos, edges, box, 2.5)
    for t in (1, 3, 8):
        assert np.array_equal(orc.lj_force_gather(pos, edges, box, 2.5, threads=t), f_serial)


def test_md_threads_keep_serial_order(pa, orc):
    pos, box, edges = pa.moldyn_system(4, seed=5)
    vel = pa.gen_velocities(len(pos), 1.0, seed=6)
    a = orc.md_run(pos, vel, edges, box, 2.5, 0.005, 6, threads=1)
    b = orc.md_run(pos, vel, edges, box, 2.5, 0.005, 6, threads=7)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("cells,cut", [(4, 2.2), (6, 2.8), (7, 2.5)])
def test_cell_list_pairs(pa, orc, cells, cut):
    pos, box = pa.gen_fcc(cells, cells, cells, jitter=0.2, seed=cells)
    e = orc.pairs_cells(pos, box[0], cut)
    assert np.array_equal(e, orc.pairs_brute(pos, box[0], cut))
    assert np.array_equal(e, pa.gen_pairs(pos, box[0], cut))


def test_md_rebuild_matches_fixed_list_between_rebuilds(pa, orc):
    # no rebuild inside the run: the same trajectory as a fixed list built at t=0
    pos, box = pa.gen_fcc(6, 6, 6, jitter=0.0, seed=9)
    vel = pa.gen_velocities(len(pos), 1.44, seed=10)
    e = orc.pairs_cells(pos, box[0], 2.8)
    a = orc.md_run(pos, vel, e, box[0], 2.5, 0.005, 15, threads=2)
    b = orc.md_run_rebuild(pos, vel, box[0], 2.5, 2.8, 100, 0.005, 15, threads=2)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


def test_f64_gather_equals_serial_scatter(pa, orc):
    pos, box, edges = pa.moldyn_system(5, seed=13)
    f_serial, _ = orc.lj_force(pos.astype(np.float64), edges, box, 2.5, dtype=np.float64)
    for t in (1, 5):
        assert np.array_equal(orc.lj_force_f64_gather(pos, edges, box, 2.5, threads=t), f_serial)


def test_md_rebuild_f64_matches_fixed_list(pa, orc):
    # one list build (every > steps): the FP64 loop with a fixed list of the
    # pairs within the neighbour cutoff at t=0, bit for bit
    pos, box = pa.gen_fcc(7, 7, 7, jitter=0.05, seed=19)
    vel = pa.gen_velocities(len(pos), 1.44, seed=20)
    e = orc.pairs_cells(pos, box[0], 2.8)
    a = orc.md_run(pos.astype(np.float64), vel.astype(np.float64), e, box[0], 2.5, 0.005, 15, dtype=np.float64)
    b = orc.md_run_rebuild_f64(pos, vel, box[0], 2.5, 2.8, 100, 0.005, 15, threads=3)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    # with rebuilds every 10 steps over 45 steps: the same trajectory as a
    # fixed list wide enough that no pair enters the cutoff from outside it
    e_wide = orc.pairs_cells(pos, box[0], 3.6)
    c = orc.md_run(pos.astype(np.float64), vel.astype(np.float64), e_wide, box[0], 2.5, 0.005, 45, dtype=np.float64)
    d = orc.md_run_rebuild_f64(pos, vel, box[0], 2.5, 2.8, 10, 0.005, 45, threads=2)
    for x, y in zip(c[:3], d[:3]):
        assert np.array_equal(x, y)
    # and the FP32 loop stays close to it
    f32 = orc.md_run_rebuild(pos, vel, box[0], 2.5, 2.8, 10, 0.005, 45, threads=2)
    assert np.max(np.abs(f32[0] - d[0])) < 1e-3


@pytest.mark.parametrize("form", [0, 1])
def test_reference_functor_forms(pa, orc, form):
    """The reference functors' own arithmetic (no 48 eps prefactor): 48 x it
    equals the FP64 force of the serial scatter to FP32 rounding; on two
    particles it is the closed form r148 d (moldyn) / rrd3 (rrd3 - 1/2) rrd d
    (miniMD), which are the same function."""
    pos, box, edges = pa.moldyn_system(5, seed=17)
    pos64 = pos.astype(np.float64)
    ref = orc.ir_lj_ref(pos64, edges, box, 2.5 ** 2, form)
    f64, _ = orc.lj_force(pos64, edges, box, 2.5, pbc=form == 0, dtype=np.float64)
    scale = np.max(np.abs(f64))
    assert np.max(np.abs(48.0 * ref.astype(np.float64) - f64)) < 1e-5 * scale
    p2 = np.array([[3.0, 3.0, 3.0], [4.3, 3.0, 3.0]])
    r = 1.3
    one = orc.ir_lj_ref(p2, np.array([[0, 1]], dtype=np.int32), 20.0, 6.25, form)
    expect = np.float32(-r * (r ** -14 - 0.5 * r ** -8))
    assert abs(one[0, 0] - expect) <= 1e-6 * abs(expect) and one[1, 0] == -one[0, 0]
    # a node with no pair inside the cutoff keeps zero
    p3 = np.array([[3.0, 3.0, 3.0], [5.6, 3.0, 3.0]])
    assert not orc.ir_lj_ref(p3, np.array([[0, 1]], dtype=np.int32), 20.0, 6.25, form).any()
