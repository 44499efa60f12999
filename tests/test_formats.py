"""Readers of the reference's input files (host code in libpattern; no GPU).

Fixtures: tests/golden/reference_inputs/mold and lj.in are the reference's
own input files (heat_3D/mold, miniMD/lj.in), copied as data; their values
are read back and checked against what they say.  The graph and coordinate
files the mold file names (input/64-3.5c.mesh, .xyz) are not in the
reference, so .mesh/.xyz files are written here in ReadInputGraph's format
(kmeans/misc.c:63-75: header, then each node's 1-based neighbours).
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "reference_inputs")


def test_read_mold_reference_file(pa):
    m = pa.read_mold(os.path.join(REF, "mold"))
    assert (m["box_cells"], m["steps"], m["neigh_update"], m["cutoff"]) == (64, 100, 20, 3.5)
    assert (m["mesh"], m["xyz"]) == ("input/64-3.5c.mesh", "input/64-3.5c.xyz")
    assert (m["percent"], m["start_time"], m["update_freq"]) == (10.0, 0, 1)


def test_read_ljin_reference_file(pa):
    m = pa.read_ljin(os.path.join(REF, "lj.in"))
    assert (m["nx"], m["ny"], m["nz"], m["ntimes"]) == (50, 50, 50, 100)
    assert (m["nbinx"], m["nbiny"], m["nbinz"]) == (10, 10, 10)
    assert (m["dt"], m["t_request"], m["rho"]) == (0.00462, 1.444, 0.8442)
    assert (m["every"], m["cutforce"], m["cutneigh"], m["nstat"]) == (20, 2.5, 2.8, 0)


def write_mesh(path, n, edges, comments=("% written by the test",)):
    adj = [[] for _ in range(n)]
    for i, j in edges:
        adj[i].append(j + 1)
        adj[j].append(i + 1)
    with open(path, "w") as f:
        for c in comments:
            f.write(c + "\n")
        f.write(f"{n} {len(edges)}\n")
        for a in adj:
            f.write(" ".join(str(v) for v in a) + "\n")


def test_read_mesh_hand_file(pa, tmp_path):
    p = tmp_path / "g.mesh"
    p.write_text("% comment\n% another\n4 4\n2 3\n1 3 4\n1 2\n2\n")
    n, e = pa.read_mesh(str(p))
    assert n == 4 and e.tolist() == [[0, 1], [0, 2], [1, 2], [1, 3]]
    n, e1 = pa.read_mesh(str(p), zero_based=False)
    assert e1.tolist() == [[1, 2], [1, 3], [2, 3], [2, 4]]
    # a node listing itself is an edge to itself, as the reference counts it
    p.write_text("2 2\n1 2\n1\n")
    assert pa.read_mesh(str(p))[1].tolist() == [[0, 0], [0, 1]]


def test_read_mesh_errors(pa, tmp_path):
    p = tmp_path / "bad.mesh"
    p.write_text("3 5\n2\n1 3\n2\n")
    with pytest.raises(RuntimeError, match="header says 5"):
        pa.read_mesh(str(p))
    p.write_text("2 1\n3\n\n")
    with pytest.raises(RuntimeError, match="neighbour 3"):
        pa.read_mesh(str(p))
    with pytest.raises(RuntimeError, match="cannot open"):
        pa.read_mesh(str(tmp_path / "missing.mesh"))


def test_mesh_xyz_round_trip(pa, tmp_path):
    pos, box, edges = pa.moldyn_system(3, seed=8)
    n = len(pos)
    lo, hi = np.minimum(edges[:, 0], edges[:, 1]), np.maximum(edges[:, 0], edges[:, 1])
    write_mesh(tmp_path / "s.mesh", n, list(zip(lo.tolist(), hi.tolist())))
    nn, e = pa.read_mesh(str(tmp_path / "s.mesh"))
    assert nn == n and len(e) == len(edges)
    # the reader emits each edge once, from its smaller node, grouped by that node
    assert np.all(e[:, 0] < e[:, 1]) and np.all(np.diff(e[:, 0]) >= 0)
    assert sorted(map(tuple, e.tolist())) == sorted(zip(lo.tolist(), hi.tolist()))
    np.savetxt(tmp_path / "s.xyz", pos.astype(np.float64), fmt="%.17g")
    x = pa.read_xyz(str(tmp_path / "s.xyz"), n)
    assert np.array_equal(x, pos.astype(np.float64))
    with pytest.raises(RuntimeError, match="expected"):
        pa.read_xyz(str(tmp_path / "s.xyz"), n + 1)
