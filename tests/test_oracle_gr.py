"""Pins the oracle's additive generalized reductions (velocity statistics,
kmeans/regular.h:30-73; pair energy, miniMD/regular.h:139-181) with values
fixed by the mathematics, not by restating the loops."""
import math

import numpy as np


def test_vstats_pythagorean_magnitudes(orc):
    # integer vectors with integer length: |v| and its FP32 rounding are exact
    v = np.array([[3, 4, 0], [2, 3, 6], [1, 4, 8], [2, 6, 9], [0, 0, 0], [-6, 2, -3]], dtype=np.float64)
    (s, c), per = orc.gr_vstats(v, 7.0, per_unit=True)
    assert per[:, 0].tolist() == [5.0, 7.0, 9.0, 11.0, 0.0, 7.0]
    assert s == 39.0
    # strictly above: the two lengths equal to 7 do not count
    assert c == 2.0 and per[:, 1].tolist() == [0, 0, 1, 1, 0, 0]


def test_vstats_sum_and_count_invariants(orc):
    rng = np.random.default_rng(5)
    v = rng.normal(size=(20000, 3)) * 1.2
    (s, c), per = orc.gr_vstats(v, 1.5, per_unit=True)
    # each value is |v| rounded once to FP32: within half an ulp of the FP64 norm
    norm = np.sqrt(np.einsum("ij,ij->i", v, v))
    assert np.all(np.abs(per[:, 0] - norm) <= np.spacing(per[:, 0].astype(np.float32)) / 2 * 1.0001)
    # the sum is order independent to FP64 rounding; the count is exact
    (s2, c2) = orc.gr_vstats(v[::-1].copy(), 1.5)
    assert c2 == c and abs(s2 - s) <= 1e-12 * s
    assert abs(s - math.fsum(per[:, 0].astype(np.float64))) <= 1e-12 * s


def test_lj_energy_closed_forms(orc):
    # phi = r^-12 - r^-6: minimum -1/4 at r = 2^(1/6); zero at r = 1
    x = np.array([[0.0, 0.0, 0.0], [2 ** (1 / 6), 0.0, 0.0], [5.0, 0.0, 0.0], [6.0, 0.0, 0.0]])
    s, n = orc.gr_lj_energy(np.array([[0, 1]], np.int32), x, 6.25)
    assert n == 1 and abs(s + 0.25) < 1e-15
    s, n = orc.gr_lj_energy(np.array([[2, 3]], np.int32), x, 6.25)
    assert n == 1 and abs(s) < 1e-15
    # the cutoff is strict and there is no minimum image
    s, n = orc.gr_lj_energy(np.array([[0, 2]], np.int32), x, 25.0)
    assert n == 0 and s == 0.0
    s, n = orc.gr_lj_energy(np.array([[0, 2]], np.int32), x, 25.0 + 1e-12)
    assert n == 1 and abs(s - (5.0 ** -12 - 5.0 ** -6)) < 1e-18


def test_lj_energy_pair_sum(orc):
    # a random cloud: the sum over pairs equals the sum of the closed form
    rng = np.random.default_rng(9)
    x = rng.uniform(0, 6, size=(60, 3))
    i, j = np.triu_indices(60, 1)
    e = np.stack([i, j], 1).astype(np.int32)
    r2 = np.sum((x[i] - x[j]) ** 2, axis=1)
    m = r2 < 6.25
    ref = math.fsum((r2[m] ** -6 - r2[m] ** -3).tolist())
    s, n = orc.gr_lj_energy(e, x, 6.25)
    assert n == int(m.sum())
    assert abs(s - ref) <= 1e-9 * max(1.0, abs(ref))
