"""GPU parity of the device binned neighbour rebuild and the miniMD workload.

The device-built CSR (all pairs within cutoff + skin, periodic images, rows
sorted) must equal the CSR of the host pair generator exactly (the pair
generator is itself pinned to miniMD's Neighbor::build and to brute force).
Dynamics with rebuilds every 20 steps are compared with the FP32 oracle
rebuilding its pair list at the same steps: 1e-5 relative L-infinity.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(np.max(np.abs(b)), 1e-30))


def host_csr(n, edges):
    both = np.concatenate([edges, edges[:, ::-1]])
    both = both[np.lexsort((both[:, 1], both[:, 0]))]
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, both[:, 0] + 1, 1)
    return np.cumsum(rp), both[:, 1]


@pytest.mark.parametrize("cells,seed", [(6, 1), (9, 2), (13, 3)])
def test_binned_csr_equals_host_pairs(pa, gpu, cells, seed):
    pos, box = pa.gen_fcc(cells, cells, cells, rho=0.8442, jitter=0.2, seed=seed)
    md = pa.MD(pos, None, None, box[0], 2.5, 0.005, skin=0.3, every=20)
    rp, cl = md.csr()
    md.close()
    erp, ecl = host_csr(len(pos), pa.gen_pairs(pos, box[0], 2.8))
    assert np.array_equal(rp, erp) and np.array_equal(cl, ecl)


def test_binned_csr_dense_and_clustered(pa, gpu):
    """Bins of more than 64 atoms, bin neighbourhoods of more than 768
    candidates (several register slices per wavefront) and rows longer than
    128 (the rank-sort path): a jittered simple cubic lattice at density 2.5
    with a 5^3 cluster at spacing 0.5 inside one bin."""
    rng = np.random.default_rng(11)
    L, m = 17.0, 23
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3) * (L / m)
    g = g + rng.uniform(-0.1, 0.1, g.shape)
    g = g[~np.all((g > 2.6) & (g < 5.4), axis=1)]
    c = np.stack(np.meshgrid(*[np.arange(5)] * 3, indexing="ij"), -1).reshape(-1, 3) * 0.5 + 3.0
    pos = np.mod(np.concatenate([g, c]), L).astype(np.float32)
    md = pa.MD(pos, None, None, L, 2.5, 1e-5, skin=0.3, every=20)
    rp, cl = md.csr()
    erp, ecl = host_csr(len(pos), pa.gen_pairs(pos, L, 2.8))
    assert np.diff(erp).max() > 128
    assert np.array_equal(rp, erp) and np.array_equal(cl, ecl)
    # rows longer than their one-pass slot: later rebuilds count first.  The
    # rebuild before step 20's force uses the positions after its drift,
    # which are the positions the run ends with.
    md.run(20)
    p20 = md.get()[0]
    rp, cl = md.csr()
    assert md.stats()[2] == 2
    md.close()
    erp, ecl = host_csr(len(pos), pa.gen_pairs(p20, L, 2.8))
    assert np.array_equal(rp, erp) and np.array_equal(cl, ecl)


@pytest.mark.parametrize("lanes", [0, 1, 64])
def test_binned_md_vs_oracle(pa, orc, gpu, lanes):
    """Rebuilds before the force of steps 20 and 40: bit-identical to the
    oracle rebuilding its cell-list pair list at the same steps (tuned,
    one lane per row, sliced layout)."""
    pos, box = pa.gen_fcc(8, 8, 8, rho=0.8442, jitter=0.0, seed=4)
    vel = pa.gen_velocities(len(pos), 1.44, seed=5)
    md = pa.MD(pos, vel, None, box[0], 2.5, 0.005, skin=0.3, every=20, lanes_per_row=lanes)
    t0 = time.perf_counter()
    md.run(25)
    md.run(15)
    wall_ms = (time.perf_counter() - t0) * 1e3
    p, v, f = md.get()
    npairs, steps, rebuilds, rb_ms = md.stats()
    _, _, _, rb_ms_again = md.stats()
    md.close()
    assert steps == 40 and rebuilds == 3  # at creation, step 20, step 40
    # device time of the rebuilds (events around each, settled by stats);
    # the two inside the runs fit in the runs' wall time
    assert 0 < rb_ms and rb_ms == rb_ms_again
    assert rb_ms < wall_ms + 1e3
    rp, rv, rf, rn = orc.md_run_rebuild(pos, vel, box[0], 2.5, 2.8, 20, 0.005, 40)
    assert npairs == rn
    assert np.array_equal(p, rp) and np.array_equal(v, rv) and np.array_equal(f, rf)


@pytest.mark.parametrize("form", [0, 2, 3])
@pytest.mark.parametrize("order", [1, 2])
def test_binned_md_slice_orders(pa, orc, gpu, order, form):
    """Sliced layout with its rows in row-class or spatial (Morton) slice
    order, the latter recomputed at each rebuild: bit-identical to the
    oracle across the rebuilds at steps 20 and 40."""
    pos, box = pa.gen_fcc(9, 9, 9, rho=0.8442, jitter=0.05, seed=6)
    vel = pa.gen_velocities(len(pos), 1.44, seed=7)
    md = pa.MD(pos, vel, None, box[0], 2.5, 0.005, skin=0.3, every=20)
    md.set_sliced_form(form)
    md.set_slice_order(order)
    assert md.slice_order() == order
    md.run(23)
    md.run(22)
    assert md.slice_order() == order and md.sliced_form() == form
    p, v, f = md.get()
    md.close()
    rp, rv, rf, _ = orc.md_run_rebuild(pos, vel, box[0], 2.5, 2.8, 20, 0.005, 45)
    assert np.array_equal(p, rp) and np.array_equal(v, rv) and np.array_equal(f, rf)


def test_minimd_workload(pa, orc, gpu):
    """miniMD: FCC 64^3 cells (1,048,576 atoms), density 0.8442, Maxwell
    velocities at T = 1.44, cutoff 2.5, skin 0.3, binned rebuild every 20
    steps, dt 0.005, 100 steps."""
    pos, box = pa.gen_fcc(64, 64, 64, rho=0.8442, jitter=0.0, seed=7)
    vel = pa.gen_velocities(len(pos), 1.44, seed=8)
    md = pa.MD(pos, vel, None, box[0], 2.5, 0.005, skin=0.3, every=20)
    ep0, ek0 = md.energy()
    md.run(100)
    p, v, f = md.get()
    ep1, ek1 = md.energy()
    npairs, steps, rebuilds, rb_ms = md.stats()
    md.close()
    print(f"miniMD 64^3: {len(pos)} atoms, {npairs} half pairs within 2.8 after the last rebuild, {rebuilds} "
          f"rebuilds ({rb_ms:.1f} ms), E0 {ep0 + ek0:.6e} E100 {ep1 + ek1:.6e}")
    assert 3.8e7 < npairs < 4.2e7  # 39 per atom in the lattice, a little fewer once it melts
    assert abs((ep1 + ek1) - (ep0 + ek0)) / abs(ep0 + ek0) < 2e-3
    rp, rv, rf, rn = orc.md_run_rebuild(pos, vel, box[0], 2.5, 2.8, 20, 0.005, 100, threads=16)
    assert npairs == rn
    assert np.array_equal(p, rp) and np.array_equal(v, rv) and np.array_equal(f, rf)


@pytest.mark.parametrize("cells", [4, 5])
def test_binned_small_box_vs_oracle(pa, orc, gpu, cells):
    """Boxes of two neighbour bins per side (4^3 and 5^3 FCC cells: sides
    6.72 and 8.40 against cutoff + skin 2.8): every pair is tested on the
    device; across the rebuilds at steps 20 and 40 the run is bit-identical
    to the oracle rebuilding its O(N^2) pair list at the same steps."""
    pos, box = pa.gen_fcc(cells, cells, cells, rho=0.8442, jitter=0.05, seed=40 + cells)
    assert 2 * 2.8 < box[0] < 3 * 2.8
    vel = pa.gen_velocities(len(pos), 1.44, seed=50 + cells)
    md = pa.MD(pos, vel, None, box[0], 2.5, 0.005, skin=0.3, every=20)
    md.run(23)
    md.run(22)
    p, v, f = md.get()
    npairs, steps, rebuilds, _ = md.stats()
    rp_, cl_ = md.csr()
    md.close()
    assert steps == 45 and rebuilds == 3
    rp, rv, rf, rn = orc.md_run_rebuild(pos, vel, box[0], 2.5, 2.8, 20, 0.005, 45)
    assert npairs == rn
    assert np.array_equal(p, rp) and np.array_equal(v, rv) and np.array_equal(f, rf)
    # the last build's rows: every neighbour within 2.8 of the step-40 positions, sorted
    assert np.all(np.diff(rp_) >= 0) and rp_[-1] == 2 * npairs
    for r in range(0, len(pos), 37):
        row = cl_[rp_[r]:rp_[r + 1]]
        assert np.all(np.diff(row) > 0)
