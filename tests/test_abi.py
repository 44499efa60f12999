"""The C ABI: libpattern.so loads, and exports every entry point
include/*.h declares (no compute calls; runs without a GPU)."""
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pat_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported(pa):
    names = declared()
    assert len(names) >= 10
    L = pa.lib()
    for n in sorted(names):
        assert hasattr(L, n), f"{n} declared in include/ but not exported"


def test_binding_covers_header(pa):
    assert declared() == set(pa.SIGNATURES), declared() ^ set(pa.SIGNATURES)


def test_error_path_without_gpu(pa):
    # argument validation runs before any device call
    import ctypes
    rc = pa.lib().pat_heat3d_run_dev(None, None, 4, 4, 4, 1, 0.1, 0, 0, 0, 0, None, None)
    assert rc == -1
    assert b"distinct" in pa.lib().pat_last_error()
    h = ctypes.c_void_p()
    interior = (ctypes.c_int64 * 3)(4, 4, 4)
    assert pa.lib().pat_stencil_create(1, interior, 1, 4, None, 0, ctypes.byref(h)) == -5  # 2-D or 3-D only
    assert pa.lib().pat_stencil_create(3, interior, 17, 4, None, 0, ctypes.byref(h)) == -5  # halo width 1..16
    assert pa.lib().pat_stencil_create(3, interior, 1, 2, None, 0, ctypes.byref(h)) == -5  # float or double


def test_regular_create_rejects_units_outside_the_input(pa):
    """The keyed generalized reduction checks every unit against the input
    before it allocates anything, without overflowing on extreme offsets."""
    import ctypes
    h = ctypes.c_void_p()
    buf = np.zeros(64, np.uint8)
    f = pa.lib().pat_regular_create
    off = (ctypes.c_int64 * 2)(0, 2 ** 63 - 4)
    assert f(buf.ctypes.data, 64, 2, 8, off, ctypes.byref(h)) == -1  # second unit far past the end
    assert b"outside" in pa.lib().pat_last_error()
    off = (ctypes.c_int64 * 2)(0, 57)
    assert f(buf.ctypes.data, 64, 2, 8, off, ctypes.byref(h)) == -1  # one byte past the end
    off = (ctypes.c_int64 * 2)(-8, 0)
    assert f(buf.ctypes.data, 64, 2, 8, off, ctypes.byref(h)) == -1
    assert f(buf.ctypes.data, 64, 2 ** 61, 8, None, ctypes.byref(h)) == -1  # count x size would overflow
    assert f(buf.ctypes.data, 64, 9, 8, None, ctypes.byref(h)) == -1  # 72 bytes in a 64-byte input
    assert f(buf.ctypes.data, 64, 1, 65, None, ctypes.byref(h)) == -1  # one unit larger than the input


def test_neighbour_build_entry_total_guard(pa):
    """The binned neighbour build refuses, before writing any row, a list
    whose 64-bit entry total does not fit the int32 row pointers -- also when
    the int32 scan has wrapped back to a positive number."""
    import ctypes
    f = pa.lib().pat_internal_nb_check_total
    f.argtypes = [ctypes.c_ulonglong, ctypes.c_int]
    f.restype = ctypes.c_int
    assert f(0, 0) == 0
    assert f(2 ** 31 - 1, 2 ** 31 - 1) == 0
    assert f(2 ** 31, -2 ** 31) == -5  # one past: the int32 total is negative
    assert b"int32" in pa.lib().pat_last_error()
    assert f(2 ** 32 + 5, 5) == -5  # wrapped to a small positive total
    assert f(3 * 2 ** 31, 2 ** 31 - 2 ** 31) == -5
    assert f(1000, 999) == -1  # the two totals disagree


def test_serving_structs_match_the_header(pa, tmp_path):
    """The ctypes mirrors of pat_job and pat_job_result have the header's
    sizes and field offsets (compiled with the host C compiler)."""
    import ctypes
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "layout.c"
    fields = {"pat_job": [f[0] for f in pa.Job._fields_], "pat_job_result": [f[0] for f in pa.JobResult._fields_]}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pattern.h"', "int main(void) {"]
    for t, fs in fields.items():
        lines.append(f'printf("{t} %zu\\n", sizeof({t}));')
        lines += [f'printf("{t}.{f} %zu\\n", offsetof({t}, {f}));' for f in fs]
    lines += ["return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for t, cls in (("pat_job", pa.Job), ("pat_job_result", pa.JobResult)):
        assert int(got[t]) == ctypes.sizeof(cls), t
        for f in fields[t]:
            assert int(got[f"{t}.{f}"]) == getattr(cls, f).offset, (t, f)


def test_python_constants_match_the_header(pa):
    """Every integer #define of include/pattern.h that the Python binding
    mirrors has the header's value; the status codes are all mirrored."""
    with open(os.path.join(ROOT, "include", "pattern.h")) as f:
        defs = dict(re.findall(r"#define (PAT_\w+) (-?\d+)\b", f.read()))
    assert all(hasattr(pa, n) for n in defs if n == "PAT_OK" or n.startswith("PAT_ERR_"))
    mirrored = [n for n in defs if hasattr(pa, n)]
    assert len(mirrored) >= 20
    for n in mirrored:
        assert getattr(pa, n) == int(defs[n]), n
