"""Batched independent systems (one fused launch per step over every
system, per-node box in the w lane) give each system exactly the result it
gets when run alone, across neighbour rebuilds, whatever in-order kernel
(lanes per row, sliced layout, tuned) either side uses."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def alone(pa, pos, vel, box, steps, lanes):
    md = pa.MD(pos, vel, None, box, 2.5, 0.005, skin=0.3, every=20, lanes_per_row=lanes)
    md.run(steps)
    out = md.get()
    md.close()
    return out


@pytest.mark.parametrize("lanes", [8, 64, 0])
def test_batch_equals_alone(pa, gpu, lanes):
    systems = [pa.batch_systems(1, m, m, seed=10 + m)[0] for m in (8, 10, 9)]
    b = pa.MDBatch(systems, lanes_per_row=lanes)
    b.run(23)
    b.run(22)  # crosses the rebuilds at steps 20 and 40
    assert b.stats()[1] == 45
    for s, (pos, vel, box) in enumerate(systems):
        got = b.get(s)
        ref = alone(pa, pos, vel, box, 45, 1)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)
    b.close()


@pytest.mark.parametrize("order", [1, 2])
def test_batch_slice_orders_equal_alone(pa, gpu, order):
    """Row-class and spatial slice orders over the batch (the spatial sort
    keeps each system's rows together) change no bits."""
    systems = [pa.batch_systems(1, m, m, seed=20 + m)[0] for m in (9, 8, 11)]
    b = pa.MDBatch(systems)
    b.set_slice_order(order)
    b.run(45)
    for s, (pos, vel, box) in enumerate(systems):
        got = b.get(s)
        ref = alone(pa, pos, vel, box, 45, 1)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)
    b.close()


def test_batch_workload_64_systems(pa, orc, gpu):
    """64 systems of 8^3-32^3 FCC cells (log-uniform), T = 1.44, 100 steps
    with rebuilds every 20, as one batched launch per step."""
    systems = pa.batch_systems(64, 8, 32, seed=99)
    b = pa.MDBatch(systems)
    tot, kms = b.run_timed(100)
    npairs, steps, rebuilds, rb_ms = b.stats()
    natoms = sum(len(p) for p, _, _ in systems)
    print(f"batch 64 systems: {natoms} atoms, {npairs} half pairs, 100 steps in {tot:.1f} ms "
          f"(force kernels {np.sum(kms):.1f} ms, {rebuilds} rebuilds {rb_ms:.1f} ms)")
    for s in (0, 17, 63):
        pos, vel, box = systems[s]
        got = b.get(s)
        ref = alone(pa, pos, vel, box, 100, 1)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)
        # and the oracle, rebuilding its cell-list pair list at the same steps
        rp, rv, rf, _ = orc.md_run_rebuild(pos, vel, box, 2.5, 2.8, 20, 0.005, 100, threads=16)
        assert np.array_equal(got[0], rp) and np.array_equal(got[1], rv) and np.array_equal(got[2], rf), s
        # momentum of each system is conserved
        assert np.max(np.abs(got[1].astype(np.float64).sum(0) - vel.astype(np.float64).sum(0))) < 1e-2
    b.close()
