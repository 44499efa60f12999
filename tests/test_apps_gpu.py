"""The four application drivers written against the C++ pattern API
(ipdps15-pattern_amd/api: RuntimeInit, StencilCreator/StencilRuntime,
RegularRuntime, IrregularRuntime -- the reference's class and call names)
run on the GPU and agree with their own host functors."""
import os
import subprocess

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
    assert "0 of 262144 cells differ" in out and "Linf" in out


def test_kmeans_app(gpu):
    out = run("kmeans", 100000, 32, 64, 5)
    assert "0 of 1031 sampled keys differ" in out


def test_moldyn_app(gpu):
    out = run("moldyn", 8, 20)
    assert "first pass" in out and "kinetic energy" in out and "velocity statistics" in out


def test_minimd_app(gpu):
    out = run("minimd", 10)
    assert "5 passes match" in out
