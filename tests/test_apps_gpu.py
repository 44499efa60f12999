"""The four application drivers written against the C++ pattern API
(ipdps15-pattern_amd/api: RuntimeInit, StencilCreator/StencilRuntime,
RegularRuntime, IrregularRuntime -- the reference's class and call names)
run on the GPU and agree with their own host functors."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ipdps15-pattern_amd", "apps", "bin")


def run(name, *args):
    p = subprocess.run([os.path.join(BIN, name), *map(str, args)], capture_output=True, text=True, timeout=240)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, p.stderr
    return p.stdout


def test_heat3d_app(gpu):
    out = run("heat3d", 64, 20)
    assert "one step, 0 of 262144 cells differ" in out and "Linf" in out
    assert "FP64 functor, one step, 0 of 262144 cells differ" in out


def test_stencil2_app(gpu):
    """The kmeans directory's W = 3 star through StencilCreator/StencilRuntime,
    its functor compiled for the device by the application."""
    out = run("stencil2", 96, 10)
    assert "one step, 0 of 884736 cells differ" in out and "Linf" in out


def test_kmeans_app(gpu):
    out = run("kmeans", 100000, 32, 64, 5)
    assert "0 of 1031 sampled keys differ" in out


def test_kmeans_app_units_not_packed(gpu):
    """Units placed 3 floats apart by their Offset array (the runtime
    gathers them before the engine runs): the same keys as the host functor,
    and the same result as the packed run."""
    packed = run("kmeans", 20000, 16, 32, 3)
    gathered = run("kmeans", 20000, 16, 32, 3, 3)
    assert "0 of 207 sampled keys differ" in gathered and "units gathered" in gathered
    tail = lambda s: s.strip().splitlines()[-1].split("sum of min distances")[1]
    assert tail(packed) == tail(gathered)


def test_moldyn_app(gpu):
    out = run("moldyn", 8, 20)
    assert "first pass" in out and "kinetic energy" in out and "velocity statistics" in out
    assert "reference functor (FP64, no 48)" in out and "0 nodes beyond 1e-5" in out


def test_minimd_app(gpu):
    out = run("minimd", 10)
    assert "5 passes match" in out and "pair energy sum" in out


def test_minimd_app_reads_lj_in(gpu, tmp_path):
    """The reference's own lj.in, with the problem size cut to 12^3 cells."""
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_inputs", "lj.in")
    lines = open(src).read().split("\n")
    lines[2] = "12 12 12        size of problem"
    (tmp_path / "lj.in").write_text("\n".join(lines))
    out = run("minimd", "-f", str(tmp_path / "lj.in"))
    assert "12 x 12 x 12 cells, density 0.8442, cutoffs 2.5 / 2.8" in out and "5 passes match" in out


def test_moldyn_app_replays_a_reference_dataset(pa, gpu, tmp_path):
    """A dataset in the reference's shape: the parameter file moldyn reads
    (kmeans/moldyn.cu:193-201) naming a .mesh graph and .xyz coordinates
    (kmeans/misc.c:63-75, 226-247), written here for a 4^3-cell FCC system
    at moldyn's density (box side (N/0.83134)^(1/3))."""
    n = 4 * 4 ** 3
    side = (n / 0.83134) ** 0.3333333
    pos, box = pa.gen_fcc(4, 4, 4, rho=n / side ** 3, jitter=0.02, seed=5)
    e = pa.gen_pairs(pos, box[0], 2.5)
    (tmp_path / "input").mkdir()
    adj = [[] for _ in range(n)]
    for i, j in e.tolist():
        adj[i].append(j + 1)
        adj[j].append(i + 1)
    with open(tmp_path / "input" / "g.mesh", "w") as f:
        f.write(f"% fcc 4^3\n{n} {len(e)}\n")
        for a in adj:
            f.write(" ".join(map(str, a)) + "\n")
    np.savetxt(tmp_path / "input" / "g.xyz", pos.astype(np.float64), fmt="%.9g")
    (tmp_path / "mold").write_text(f"4\n10\n20\n2.5\ninput/g.mesh\ninput/g.xyz\n10\n0\n1\n")
    out = run("moldyn", "-f", str(tmp_path / "mold"))
    assert f"{n} molecules, {len(e)} pairs" in out and "first pass" in out and "velocity statistics" in out


def test_regular_app(gpu):
    """RegularRuntime with the application's own functors on the device
    (keyed reduction objects): every pair agrees with the host functors."""
    out = run("regular")
    for line in ("map 0 velocity statistics", "map 1 kmeans D=3 K=40", "map 2 pair energy", "map 3 histogram",
                 "map 4 odd-sized records"):
        assert line in out
    assert "DIFFER" not in out and out.count("match") == 7
