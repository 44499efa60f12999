"""Writes tests/golden/oracle_fixtures.npz: small inputs and the oracle's
outputs for each engine, committed so that the oracle (CPU test) and the HIP
paths (GPU test) are checked against fixed vectors, not only against each
other at test time.

  heat3D: an 8^3 interior (10^3 with the zero halo), 10 steps of the heat rule;
  LJ:     FCC 3^3 cells (108 atoms, jitter), the pair list within 2.5 in its
          generated order, the serial-scatter forces (minimum image);
  kmeans: 1,000 points x 8, K 16 from the first 16 points, one Lloyd pass:
          assignments, minimum distances, the updated centres.

Run from the repository root: python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "ipdps15-pattern_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as orc  # noqa: E402
import pattern_amd as pa  # noqa: E402


def main():
    heat_in = pa.gen_heat3d(8, 8, 8, seed=3)
    heat_out = orc.heat3d(heat_in, 10)

    lj_pos, lj_box, lj_edges = pa.moldyn_system(3, seed=5)
    lj_force, _ = orc.lj_force(lj_pos, lj_edges, lj_box, 2.5)

    x = pa.gen_gmm(1000, 8, 16, sigma=1.0, seed=9)
    k = 16
    s, ds = pa.kmeans_shifts(len(x), x.shape[1], float(np.max(np.abs(x))))
    c0 = x[:k].copy()
    a, md = orc.kmeans_assign(x, c0)
    sums, counts, _ = orc.kmeans_accumulate(x, a, k, s, mindist=md, dshift=ds)
    c1 = orc.kmeans_update(sums, counts, s, c0)

    np.savez_compressed(os.path.join(HERE, "oracle_fixtures.npz"), heat_in=heat_in, heat_out=heat_out,
                        lj_pos=np.asarray(lj_pos, np.float32), lj_box=np.float32(lj_box),
                        lj_edges=np.asarray(lj_edges, np.int32), lj_force=lj_force, km_x=x, km_k=np.int32(k),
                        km_assign=a, km_mindist=md, km_centres=c1)


if __name__ == "__main__":
    main()
