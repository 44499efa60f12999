"""ctypes binding of libpattern's C ABI (include/pattern.h).

This is the Python mirror of the reference's runtime-class calls used by the
tests and the benchmark.  Loading fails loudly when libpattern.so has not been
built: there is no Python or CPU fallback for any engine.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpattern.so")

PAT_OK = 0
PAT_STENCIL_HEAT = 0
PAT_STENCIL_AUTO = 0
PAT_STENCIL_LDS25D = 1
PAT_STENCIL_GENERIC = 2

_lib = None

c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_int = ctypes.c_int
c_float = ctypes.c_float
c_double = ctypes.c_double
c_vp = ctypes.c_void_p
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_fp = ctypes.POINTER(ctypes.c_float)
c_dp = ctypes.POINTER(ctypes.c_double)
c_intp = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes) for every symbol include/pattern.h declares
SIGNATURES = {
    "pat_last_error": (ctypes.c_char_p, []),
    "pat_init": (c_int, [c_int]),
    "pat_finalize": (c_int, []),
    "pat_info": (c_int, [ctypes.c_char_p, c_i64]),
    "pat_device_sync": (c_int, []),
    "pat_gen_heat3d": (c_int, [c_i64, c_i64, c_i64, c_u64, c_vp]),
    "pat_gen_fcc": (c_int, [c_i64, c_i64, c_i64, c_double, c_double, c_u64, c_vp, c_dp]),
    "pat_gen_velocities": (c_int, [c_i64, c_double, c_u64, c_vp]),
    "pat_gen_pairs": (c_int, [c_i64, c_vp, c_double, c_double, c_int, c_u64, c_vp, c_i64p]),
    "pat_gen_gmm": (c_int, [c_i64, c_int, c_int, c_double, c_double, c_double, c_u64, c_vp]),
    "pat_heat3d_run_dev": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_int, c_float, c_int, c_int, c_int, c_int,
                                   c_intp, c_vp]),
    "pat_stencil_create": (c_int, [c_int, c_i64p, c_int, c_int, c_vp, c_i64, ctypes.POINTER(c_vp)]),
    "pat_stencil_set_idx": (c_int, [c_vp, c_int]),
    "pat_stencil_set_variant": (c_int, [c_vp, c_int, c_int, c_int, c_int]),
    "pat_stencil_upload": (c_int, [c_vp, c_vp]),
    "pat_stencil_run": (c_int, [c_vp, c_int]),
    "pat_stencil_download": (c_int, [c_vp, c_vp]),
    "pat_stencil_device_ptr": (c_int, [c_vp, ctypes.POINTER(c_vp)]),
    "pat_stencil_destroy": (c_int, [c_vp]),
}


class PatternError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def lib():
    """Loads libpattern.so (raises if it is missing: no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libpattern.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(fn, rc):
    if rc != PAT_OK:
        raise PatternError(fn, rc, lib().pat_last_error().decode())
    return rc


def call(name, *args):
    return check(name, getattr(lib(), name)(*args))


def ptr(a):
    """Address of a numpy array or torch tensor (host or device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    assert a.is_contiguous()
    return a.data_ptr()


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


# ---------------------------------------------------------------- generators

def gen_heat3d(nx, ny, nz, seed=42):
    g = np.empty((nz + 2, ny + 2, nx + 2), dtype=np.float32)
    call("pat_gen_heat3d", nx, ny, nz, seed, ptr(g))
    return g


def gen_fcc(ncx, ncy, ncz, rho=0.8442, jitter=0.0, seed=1):
    n = 4 * ncx * ncy * ncz
    pos = np.empty((n, 3), dtype=np.float32)
    box = (c_double * 3)()
    call("pat_gen_fcc", ncx, ncy, ncz, rho, jitter, seed, ptr(pos), box)
    return pos, np.array(box[:])


def gen_velocities(n, t, seed=2):
    v = np.empty((n, 3), dtype=np.float32)
    call("pat_gen_velocities", n, t, seed, ptr(v))
    return v


def gen_pairs(pos, box, cut, shuffle=False, seed=3):
    n = pos.shape[0]
    cnt = c_i64(0)
    call("pat_gen_pairs", n, ptr(pos), float(box), float(cut), int(shuffle), seed, None, ctypes.byref(cnt))
    edges = np.empty((max(cnt.value, 1), 2), dtype=np.int32)
    call("pat_gen_pairs", n, ptr(pos), float(box), float(cut), int(shuffle), seed, ptr(edges), ctypes.byref(cnt))
    return edges[:cnt.value]


def gen_gmm(n, d, ncomp, sigma=1.0, lo=-10.0, hi=10.0, seed=7):
    x = np.empty((n, d), dtype=np.float32)
    call("pat_gen_gmm", n, d, ncomp, sigma, lo, hi, seed, ptr(x))
    return x


# ---------------------------------------------------------------- stencil

def heat3d_run_dev(a, b, nx, ny, nz, nsteps, alpha=0.1, variant=PAT_STENCIL_AUTO, bx=0, by=0, zc=0, stream=None):
    """Runs nsteps on device grids a, b; returns the tensor holding the result."""
    res = c_int(0)
    call("pat_heat3d_run_dev", ptr(a), ptr(b), nx, ny, nz, nsteps, alpha, variant, bx, by, zc, ctypes.byref(res),
         stream_ptr(stream))
    return b if res.value else a


class Stencil:
    """Handle mirror of StencilCreator/StencilRuntime (heat_3D/lib/stencil_creator.h)."""

    def __init__(self, nx, ny, nz, alpha=0.1, width=1):
        self.shape = (nz + 2, ny + 2, nx + 2)
        self._param = np.array([alpha], dtype=np.float32)
        h = c_vp()
        interior = (c_i64 * 3)(nx, ny, nz)
        call("pat_stencil_create", 3, interior, width, 4, ptr(self._param), 4, ctypes.byref(h))
        self.h = h

    def set_variant(self, variant, bx=0, by=0, zc=0):
        call("pat_stencil_set_variant", self.h, variant, bx, by, zc)

    def upload(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.float32)
        assert g.shape == self.shape
        call("pat_stencil_upload", self.h, ptr(g))

    def run(self, nsteps):
        call("pat_stencil_run", self.h, nsteps)

    def download(self):
        out = np.empty(self.shape, dtype=np.float32)
        call("pat_stencil_download", self.h, ptr(out))
        return out

    def close(self):
        if self.h:
            call("pat_stencil_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
