"""Pins the LJ oracle, the pair generator and the FCC generator against the
reference's own serial miniMD code (create_atoms, Neighbor::build,
Force::compute), built from the reference sources into oracle/_ref by
oracle/ref.mk.  Skipped where that library is not built."""
import numpy as np
import pytest

import oracle as _o

pytestmark = pytest.mark.skipif(not _o.ref_available(), reason="the reference's miniMD library oracle/_ref is not built")


@pytest.mark.parametrize("n,nbin", [(4, 4), (6, 5), (8, 6)])
def test_fcc_generator_matches_create_atoms(pa, n, nbin):
    rpos, _, _ = _o.ref_minimd_system(n, nbin=nbin)
    pos, box = pa.gen_fcc(n, n, n, rho=0.8442, jitter=0.0)
    assert pos.shape == rpos.shape
    # same atoms in the same order, up to FP32 rounding of the coordinates
    assert np.max(np.abs(pos - rpos)) < 1e-5 * box[0]


@pytest.mark.parametrize("n,nbin", [(4, 4), (6, 5), (8, 6)])
def test_pairs_match_neighbor_build(pa, orc, n, nbin):
    rpos, _, rpairs = _o.ref_minimd_system(n, nbin=nbin)
    # the reference list has no periodic images (miniMD without ghosts)
    mine = orc.pairs_brute(rpos.astype(np.float32), 1e9, 2.8, pbc=False)
    a = {tuple(sorted(p)) for p in rpairs.tolist()}
    b = {tuple(p) for p in mine.tolist()}
    assert len(a) == len(rpairs)  # every pair stored once
    assert a == b


@pytest.mark.parametrize("n,nbin", [(4, 4), (6, 5), (8, 6)])
def test_force_matches_force_compute(orc, n, nbin):
    rpos, rforce, rpairs = _o.ref_minimd_system(n, nbin=nbin)
    f, _ = orc.lj_force(rpos, rpairs, 1e9, 2.5, pbc=False, dtype=np.float64)
    assert np.max(np.abs(rforce)) > 0.05  # surface atoms feel a net force
    # the oracle writes out the 48*eps the reference folds elsewhere
    assert np.max(np.abs(f - 48.0 * rforce)) < 1e-10 * np.max(np.abs(f))
    f32, _ = orc.lj_force(rpos.astype(np.float32), rpairs, 1e9, 2.5, pbc=False)
    assert np.max(np.abs(f32 - 48.0 * rforce)) < 1e-5 * np.max(np.abs(f))
