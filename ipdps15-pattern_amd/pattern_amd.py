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
PAT_STENCIL_TEMPORAL2 = 3
PAT_STENCIL_CUBES = 4
PAT_GR_VSTATS = 0
PAT_GR_LJ_ENERGY = 2
PAT_IR_LJ_MINIMAGE = 0
PAT_IR_LJ_OPEN = 1
PAT_IR_CSR = 0
PAT_IR_ATOMIC = 1
PAT_IR_CSR_TREE = 2
PAT_IR_TABLE_KERNEL = 3
PAT_IR_NO_FUNCTOR = -1
PAT_IR_SLICED = 64
PAT_IR_ORDER_EDGE = 0
PAT_IR_ORDER_NODE = 1
PAT_KM_PLAIN = 0
PAT_KM_MFMA = 1
PAT_JOB_HEAT3D = 0
PAT_JOB_LJ = 1
PAT_JOB_KMEANS = 2
PAT_SERVE_PLAIN = 1
PAT_SERVE_NO_FUSE = 2
# status codes (include/pattern.h)
PAT_OK, PAT_ERR_INVALID, PAT_ERR_HIP, PAT_ERR_NOMEM, PAT_ERR_BUDGET, PAT_ERR_UNSUPPORTED, PAT_ERR_STATE = \
    0, -1, -2, -3, -4, -5, -6

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


class Job(ctypes.Structure):
    """struct pat_job."""
    _fields_ = [("kind", ctypes.c_int), ("size", ctypes.c_int64), ("dim", ctypes.c_int), ("k", ctypes.c_int),
                ("steps", ctypes.c_int), ("seed", ctypes.c_uint64), ("out", ctypes.c_void_p),
                ("out_len", ctypes.c_int64)]


class JobResult(ctypes.Structure):
    """struct pat_job_result."""
    _fields_ = [("status", ctypes.c_int), ("kernel", ctypes.c_int), ("checksum", ctypes.c_double),
                ("t_submit_ms", ctypes.c_double), ("t_start_ms", ctypes.c_double), ("t_end_ms", ctypes.c_double),
                ("device_bytes", ctypes.c_int64), ("batch", ctypes.c_int)]


# name -> (restype, argtypes) for every symbol include/pattern.h declares
SIGNATURES = {
    "pat_last_error": (ctypes.c_char_p, []),
    "pat_init": (c_int, [c_int]),
    "pat_finalize": (c_int, []),
    "pat_info": (c_int, [ctypes.c_char_p, c_i64]),
    "pat_device_sync": (c_int, []),
    "pat_dev_malloc": (c_int, [ctypes.POINTER(c_vp), c_i64]),
    "pat_dev_free": (c_int, [c_vp]),
    "pat_memcpy_h2d": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "pat_memcpy_d2h": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "pat_memcpy_d2d": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "pat_memset_dev": (c_int, [c_vp, c_int, c_i64, c_vp]),
    "pat_stream_create": (c_int, [ctypes.POINTER(c_vp)]),
    "pat_stream_destroy": (c_int, [c_vp]),
    "pat_stream_sync": (c_int, [c_vp]),
    "pat_event_create": (c_int, [ctypes.POINTER(c_vp)]),
    "pat_event_destroy": (c_int, [c_vp]),
    "pat_event_record": (c_int, [c_vp, c_vp]),
    "pat_event_elapsed_ms": (c_int, [c_vp, c_vp, ctypes.POINTER(c_float)]),
    "pat_device_query": (c_int, [c_intp, c_i64p, c_i64p]),
    "pat_gen_heat3d": (c_int, [c_i64, c_i64, c_i64, c_u64, c_vp]),
    "pat_gen_fcc": (c_int, [c_i64, c_i64, c_i64, c_double, c_double, c_u64, c_vp, c_dp]),
    "pat_gen_velocities": (c_int, [c_i64, c_double, c_u64, c_vp]),
    "pat_gen_pairs": (c_int, [c_i64, c_vp, c_double, c_double, c_int, c_u64, c_vp, c_i64p]),
    "pat_gen_gmm": (c_int, [c_i64, c_int, c_int, c_double, c_double, c_double, c_u64, c_vp]),
    "pat_gen_heat3d_dev": (c_int, [c_i64, c_i64, c_i64, c_u64, c_vp, c_vp]),
    "pat_gen_fcc_dev": (c_int, [c_i64, c_i64, c_i64, c_double, c_double, c_u64, c_vp, c_dp, c_vp]),
    "pat_gen_gmm_dev": (c_int, [c_i64, c_int, c_int, c_double, c_double, c_double, c_u64, c_vp, c_vp]),
    "pat_lj_rcp_selftest": (c_int, [c_float, c_float, c_i64p, c_i64p]),
    "pat_ir_sell_build_dev": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_i64p, c_vp]),
    "pat_lj_force_sell_dev": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_float, c_int, c_float, c_int, c_vp, c_vp]),
    "pat_serve_start": (c_int, [c_i64, c_int, c_int, ctypes.POINTER(c_vp)]),
    "pat_serve_submit": (c_int, [c_vp, ctypes.POINTER(Job), c_i64p]),
    "pat_serve_poll": (c_int, [c_vp, c_i64, c_intp, ctypes.POINTER(JobResult)]),
    "pat_serve_wait": (c_int, [c_vp, c_i64, ctypes.POINTER(JobResult)]),
    "pat_serve_drain": (c_int, [c_vp]),
    "pat_serve_stats": (c_int, [c_vp, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p]),
    "pat_serve_stop": (c_int, [c_vp]),
    "pat_heat3d_run_dev": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_int, c_float, c_int, c_int, c_int, c_int,
                                   c_intp, c_vp]),
    "pat_stencil_create": (c_int, [c_int, c_i64p, c_int, c_int, c_vp, c_i64, ctypes.POINTER(c_vp)]),
    "pat_stencil_set_idx": (c_int, [c_vp, c_int]),
    "pat_stencil_set_table": (c_int, [c_vp, c_vp]),
    "pat_stencil_set_variant": (c_int, [c_vp, c_int, c_int, c_int, c_int]),
    "pat_stencil_upload": (c_int, [c_vp, c_vp]),
    "pat_stencil_run": (c_int, [c_vp, c_int]),
    "pat_stencil_run_timed": (c_int, [c_vp, c_int, c_fp, c_fp]),
    "pat_stencil_run_graph_timed": (c_int, [c_vp, c_int, c_fp]),
    "pat_stencil_download": (c_int, [c_vp, c_vp]),
    "pat_stencil_tile": (c_int, [c_vp, c_intp, c_intp, c_intp]),
    "pat_stencil_device_ptr": (c_int, [c_vp, ctypes.POINTER(c_vp)]),
    "pat_stencil_destroy": (c_int, [c_vp]),
    "pat_ir_csr_workspace": (c_int, [c_i64, c_i64, c_i64p]),
    "pat_ir_csr_build_dev_noalloc": (c_int, [c_vp, c_i64, c_i64, c_int, c_vp, c_vp, c_vp]),
    "pat_ir_csr_build_dev": (c_int, [c_vp, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "pat_lj_force_csr_dev": (c_int, [c_vp, c_vp, c_vp, c_i64, c_float, c_int, c_float, c_int, c_vp, c_vp]),
    "pat_lj_force_atomic_dev": (c_int, [c_vp, c_vp, c_i64, c_i64, c_float, c_int, c_float, c_vp, c_vp]),
    "pat_ir_create": (c_int, [c_i64, c_i64, c_vp, c_int, c_vp, c_i64, ctypes.POINTER(c_vp)]),
    "pat_ir_set_map_idx": (c_int, [c_vp, c_int]),
    "pat_ir_set_reduce_idx": (c_int, [c_vp, c_int]),
    "pat_ir_set_table": (c_int, [c_vp, c_vp]),
    "pat_ir_set_param": (c_int, [c_vp, c_vp, c_i64]),
    "pat_ir_set_variant": (c_int, [c_vp, c_int, c_int]),
    "pat_ir_set_row_order": (c_int, [c_vp, c_int]),
    "pat_ir_reset_node_data": (c_int, [c_vp, c_vp, c_int]),
    "pat_ir_start": (c_int, [c_vp]),
    "pat_ir_get_result": (c_int, [c_vp, c_vp]),
    "pat_ir_destroy": (c_int, [c_vp]),
    "pat_md_create": (c_int, [c_i64, c_vp, c_vp, c_i64, c_vp, c_double, c_int, c_double, c_double,
                              ctypes.POINTER(c_vp)]),
    "pat_md_create_binned": (c_int, [c_i64, c_vp, c_vp, c_double, c_double, c_double, c_int, c_double,
                                     ctypes.POINTER(c_vp)]),
    "pat_md_create_binned_dev": (c_int, [c_i64, c_vp, c_vp, c_double, c_double, c_double, c_int, c_double,
                                         c_vp]),
    "pat_md_stats": (c_int, [c_vp, c_i64p, c_i64p, c_i64p, c_fp]),
    "pat_md_csr": (c_int, [c_vp, c_vp, c_vp]),
    "pat_md_set_variant": (c_int, [c_vp, c_int, c_int]),
    "pat_md_set_row_order": (c_int, [c_vp, c_int]),
    "pat_md_get_variant": (c_int, [c_vp, c_intp, c_intp, c_intp]),
    "pat_md_set_sliced_form": (c_int, [c_vp, c_int]),
    "pat_read_mold": (c_int, [ctypes.c_char_p, c_vp]),
    "pat_read_ljin": (c_int, [ctypes.c_char_p, c_vp]),
    "pat_read_mesh": (c_int, [ctypes.c_char_p, c_int, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64), c_vp, c_i64]),
    "pat_read_xyz": (c_int, [ctypes.c_char_p, c_i64, c_vp]),
    "pat_gr_create": (c_int, [c_vp, c_i64, c_i64, c_int, c_vp, c_i64, ctypes.POINTER(c_vp)]),
    "pat_gr_set_map_idx": (c_int, [c_vp, c_int]),
    "pat_gr_set_reduce_idx": (c_int, [c_vp, c_int]),
    "pat_gr_set_param": (c_int, [c_vp, c_vp, c_i64]),
    "pat_gr_run": (c_int, [c_vp]),
    "pat_gr_get_output": (c_int, [c_vp, c_intp, c_intp, c_vp, c_vp]),
    "pat_gr_destroy": (c_int, [c_vp]),
    "pat_regular_create": (c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp]),
    "pat_regular_set_table": (c_int, [c_vp, c_vp]),
    "pat_regular_set_capacity": (c_int, [c_vp, c_i64, c_i64, c_int]),
    "pat_regular_set_map_idx": (c_int, [c_vp, c_int]),
    "pat_regular_set_reduce_idx": (c_int, [c_vp, c_int]),
    "pat_regular_set_param": (c_int, [c_vp, c_vp, c_i64]),
    "pat_regular_start": (c_int, [c_vp]),
    "pat_regular_get_output": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pat_regular_stats": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "pat_regular_destroy": (c_int, [c_vp]),
    "pat_md_get_sliced_form": (c_int, [c_vp, c_intp]),
    "pat_md_set_slice_order": (c_int, [c_vp, c_int]),
    "pat_md_get_slice_order": (c_int, [c_vp, c_intp]),
    "pat_mdb_set_slice_order": (c_int, [c_vp, c_int]),
    "pat_mdb_create": (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, c_double, c_double, c_int, c_double,
                               ctypes.POINTER(c_vp)]),
    "pat_mdb_set_lanes": (c_int, [c_vp, c_int]),
    "pat_mdb_run": (c_int, [c_vp, c_int, c_fp, c_fp]),
    "pat_mdb_get": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp]),
    "pat_mdb_stats": (c_int, [c_vp, c_i64p, c_i64p, c_i64p, c_fp]),
    "pat_mdb_destroy": (c_int, [c_vp]),
    "pat_md_run": (c_int, [c_vp, c_int]),
    "pat_md_run_async": (c_int, [c_vp, c_int, c_vp]),
    "pat_md_run_timed": (c_int, [c_vp, c_int, c_fp, c_fp]),
    "pat_md_run_graph_timed": (c_int, [c_vp, c_int, c_fp]),
    "pat_md_get": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "pat_md_energy": (c_int, [c_vp, c_dp, c_dp]),
    "pat_md_device_buffers": (c_int, [c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), ctypes.POINTER(c_vp)]),
    "pat_md_destroy": (c_int, [c_vp]),
    "pat_kmeans_fixed_shift": (c_int, [c_i64, c_float, c_intp]),
    "pat_kmeans_create": (c_int, [c_vp, c_i64, c_int, c_int, c_int, c_int, ctypes.POINTER(c_vp)]),
    "pat_kmeans_set_mode": (c_int, [c_vp, c_int]),
    "pat_kmeans_set_terms": (c_int, [c_vp, c_int]),
    "pat_kmeans_get_terms": (c_int, [c_vp, c_intp]),
    "pat_kmeans_set_centroids": (c_int, [c_vp, c_vp]),
    "pat_kmeans_run": (c_int, [c_vp, c_int]),
    "pat_kmeans_run_timed": (c_int, [c_vp, c_int, c_fp, c_fp]),
    "pat_kmeans_get": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "pat_kmeans_get_object": (c_int, [c_vp, c_vp, c_vp, c_vp, c_intp, c_intp]),
    "pat_kmeans_stats": (c_int, [c_vp, c_i64p, c_fp]),
    "pat_kmeans_destroy": (c_int, [c_vp]),
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
    """Address of a numpy array (host) or a DeviceArray (device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    if isinstance(a, DeviceArray):
        return a.ptr
    return int(a)


def stream_ptr(stream=None):
    return stream.h if isinstance(stream, Stream) else stream


def init(device=0):
    call("pat_init", device)


class DeviceArray:
    """A device allocation made through the C ABI (hipMalloc in libpattern)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        h = c_vp()
        call("pat_dev_malloc", ctypes.byref(h), self.nbytes)
        self.ptr = h.value

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        call("pat_memcpy_h2d", d.ptr, a.ctypes.data, d.nbytes, stream_ptr(stream))
        return d

    @classmethod
    def zeros(cls, shape, dtype, stream=None):
        d = cls(shape, dtype)
        call("pat_memset_dev", d.ptr, 0, d.nbytes, stream_ptr(stream))
        call("pat_stream_sync", stream_ptr(stream))
        return d

    def numpy(self, stream=None):
        out = np.empty(self.shape, dtype=self.dtype)
        call("pat_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes, stream_ptr(stream))
        return out

    def copy_from(self, other, stream=None):
        assert other.nbytes == self.nbytes
        call("pat_memcpy_d2d", self.ptr, other.ptr, self.nbytes, stream_ptr(stream))

    def clone(self, stream=None):
        d = DeviceArray(self.shape, self.dtype)
        d.copy_from(self, stream)
        call("pat_stream_sync", stream_ptr(stream))
        return d

    def free(self):
        if self.ptr:
            call("pat_dev_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        h = c_vp()
        call("pat_stream_create", ctypes.byref(h))
        self.h = h.value

    def sync(self):
        call("pat_stream_sync", self.h)

    def __del__(self):
        try:
            call("pat_stream_destroy", self.h)
        except Exception:
            pass


class Timer:
    """HIP events recorded on the stream the timed kernels run on."""

    def __init__(self, stream=None):
        self.stream = stream_ptr(stream)
        a, b = c_vp(), c_vp()
        call("pat_event_create", ctypes.byref(a))
        call("pat_event_create", ctypes.byref(b))
        self.a, self.b = a.value, b.value

    def start(self):
        call("pat_event_record", self.a, self.stream)

    def stop(self):
        call("pat_event_record", self.b, self.stream)
        ms = c_float(0)
        call("pat_event_elapsed_ms", self.a, self.b, ctypes.byref(ms))
        return ms.value

    def __del__(self):
        try:
            call("pat_event_destroy", self.a)
            call("pat_event_destroy", self.b)
        except Exception:
            pass


def sync(stream=None):
    call("pat_stream_sync", stream_ptr(stream))


def device_query():
    cus, tot, fr = c_int(0), c_i64(0), c_i64(0)
    call("pat_device_query", ctypes.byref(cus), ctypes.byref(tot), ctypes.byref(fr))
    return cus.value, tot.value, fr.value


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


_app_libs = {}


def app_table(app, kind="stencil"):
    """Pointer to the device functor table the application 'app' compiles
    (apps/<app>/gpu_table.hip, built as apps/bin/lib<app>_table.so); kind
    names its pat_app_<kind>_table() entry point."""
    lib_ = _app_libs.get(app)
    if lib_ is None:
        path = os.path.join(_HERE, "apps", "bin", f"lib{app}_table.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C ipdps15-pattern_amd)")
        lib_ = _app_libs[app] = ctypes.CDLL(path)
    f = getattr(lib_, f"pat_app_{kind}_table")
    f.restype = c_vp
    f.argtypes = []
    return f()


class Stencil:
    """Handle mirror of StencilCreator/StencilRuntime (heat_3D/lib/stencil_creator.h).
    width: halo width; dtype: element type (float32 or float64); param: the
    functor parameter bytes (default: alpha as float32); table: an
    application table (app_table(...)) whose functor idx runs."""

    def __init__(self, nx, ny, nz, alpha=0.1, width=1, dtype=np.float32, param=None, table=None, idx=0):
        """nz None: a 2-D grid (ny, nx)."""
        self.dtype = np.dtype(dtype)
        ndims = 2 if nz is None else 3
        self.shape = (ny + 2 * width, nx + 2 * width) if nz is None else (nz + 2 * width, ny + 2 * width,
                                                                          nx + 2 * width)
        self._param = np.array([alpha], dtype=np.float32) if param is None else np.ascontiguousarray(param)
        h = c_vp()
        interior = (c_i64 * 3)(nx, ny, 1 if nz is None else nz)
        call("pat_stencil_create", ndims, interior, width, self.dtype.itemsize, ptr(self._param),
             self._param.nbytes, ctypes.byref(h))
        self.h = h
        if table is not None:
            call("pat_stencil_set_table", self.h, table)
        if table is not None or idx:
            call("pat_stencil_set_idx", self.h, idx)

    def set_variant(self, variant, bx=0, by=0, zc=0):
        call("pat_stencil_set_variant", self.h, variant, bx, by, zc)

    def upload(self, grid):
        g = np.ascontiguousarray(grid, dtype=self.dtype)
        assert g.shape == self.shape
        call("pat_stencil_upload", self.h, ptr(g))

    def run(self, nsteps):
        call("pat_stencil_run", self.h, nsteps)

    def run_timed(self, nsteps):
        """Returns (total_ms, per-sweep kernel ms array)."""
        tot = c_float(0)
        k = np.zeros(nsteps, dtype=np.float32)
        call("pat_stencil_run_timed", self.h, nsteps, ctypes.byref(tot), k.ctypes.data_as(c_fp))
        return tot.value, k

    def run_graph_timed(self, nsteps):
        """run(); returns the device time (ms) between HIP events around the
        sweep graph's replay."""
        ms = c_float(0)
        call("pat_stencil_run_graph_timed", self.h, nsteps, ctypes.byref(ms))
        return ms.value

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        call("pat_stencil_download", self.h, ptr(out))
        return out

    def tile(self):
        a, b, c = c_int(0), c_int(0), c_int(0)
        call("pat_stencil_tile", self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def close(self):
        if self.h:
            call("pat_stencil_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- irregular

def csr_build_dev(edges_dev, num_edges, num_nodes, stream=None, order=PAT_IR_ORDER_EDGE):
    """Device CSR by destination; returns (rowptr, cols) DeviceArrays."""
    wb = c_i64(0)
    call("pat_ir_csr_workspace", num_nodes, num_edges, ctypes.byref(wb))
    work = DeviceArray((max(wb.value, 1),), np.uint8)
    rowptr = DeviceArray((num_nodes + 1,), np.int32)
    cols = DeviceArray((max(2 * num_edges, 1),), np.int32)
    call("pat_ir_csr_build_dev", ptr(edges_dev), num_edges, num_nodes, order, rowptr.ptr, cols.ptr, work.ptr,
         wb.value, stream_ptr(stream))
    sync(stream)
    work.free()
    return rowptr, cols


def sell_build_dev(rowptr, cols, num_nodes, stream=None):
    """Sliced (SELL-64) copy of a device CSR; returns (sptr, scol) DeviceArrays."""
    nsl = (num_nodes + 63) // 64
    sptr = DeviceArray((nsl + 1,), np.int32)
    total = c_i64(0)
    # sizing pass: capacity 0 returns PAT_ERR_INVALID (-1) with the size needed
    rc = lib().pat_ir_sell_build_dev(rowptr.ptr, cols.ptr, num_nodes, sptr.ptr, None, 0, ctypes.byref(total),
                                     stream_ptr(stream))
    if rc not in (PAT_OK, -1):
        check("pat_ir_sell_build_dev", rc)
    scol = DeviceArray((max(total.value, 1),), np.int32)
    call("pat_ir_sell_build_dev", rowptr.ptr, cols.ptr, num_nodes, sptr.ptr, scol.ptr, total.value,
         ctypes.byref(total), stream_ptr(stream))
    return sptr, scol


def to_float4(a3):
    a4 = np.zeros((a3.shape[0], 4), dtype=np.float32)
    a4[:, :3] = a3
    return a4


class IrregularRuntime:
    """Handle mirror of IrregularRuntime (kmeans/irregular/irregular_runtime.h).
    With table (app_table(app, "ir")) the application's map / reduce indices
    select its own functors, compiled for the device."""

    def __init__(self, edges, num_nodes, map_idx=PAT_IR_LJ_MINIMAGE, param=None, table=None, reduce_idx=0):
        e = np.ascontiguousarray(edges, dtype=np.int32)
        self.n = num_nodes
        self._param = np.ascontiguousarray(param)
        h = c_vp()
        if table is None:
            call("pat_ir_create", e.shape[0], num_nodes, ptr(e), map_idx, ptr(self._param), self._param.nbytes,
                 ctypes.byref(h))
        else:
            call("pat_ir_create", e.shape[0], num_nodes, ptr(e), PAT_IR_NO_FUNCTOR, None, 0, ctypes.byref(h))
        self.h = h
        if table is not None:
            call("pat_ir_set_table", self.h, table)
            call("pat_ir_set_map_idx", self.h, map_idx)
            call("pat_ir_set_reduce_idx", self.h, reduce_idx)
            call("pat_ir_set_param", self.h, ptr(self._param), self._param.nbytes)

    def set_map_idx(self, idx):
        call("pat_ir_set_map_idx", self.h, idx)
        call("pat_ir_set_param", self.h, ptr(self._param), self._param.nbytes)

    def set_variant(self, variant, lanes_per_row=0):
        call("pat_ir_set_variant", self.h, variant, lanes_per_row)

    def set_row_order(self, order):
        call("pat_ir_set_row_order", self.h, order)

    def reset_node_data(self, pos):
        p = np.ascontiguousarray(pos)
        call("pat_ir_reset_node_data", self.h, ptr(p), p.dtype.itemsize * 3)

    def start(self):
        call("pat_ir_start", self.h)

    def get_reduction_result(self):
        out = np.empty((self.n, 3), dtype=np.float32)
        call("pat_ir_get_result", self.h, ptr(out))
        return out

    def close(self):
        if self.h:
            call("pat_ir_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MD:
    """Device-resident velocity-Verlet LJ dynamics (pat_md_*).  With
    edges=None the neighbour CSR is built and rebuilt on the device
    (cutoff + skin, every 'every' steps)."""

    def __init__(self, pos, vel, edges, box, cutoff, dt, pbc=True, variant=PAT_IR_CSR, lanes_per_row=0, skin=0.3,
                 every=20, order=PAT_IR_ORDER_NODE):
        p = np.ascontiguousarray(pos, dtype=np.float32)
        v = np.ascontiguousarray(vel, dtype=np.float32) if vel is not None else None
        self.n = p.shape[0]
        h = c_vp()
        if edges is None:
            call("pat_md_create_binned", self.n, ptr(p), ptr(v), float(box), float(cutoff), float(skin), int(every),
                 float(dt), ctypes.byref(h))
        else:
            e = np.ascontiguousarray(edges, dtype=np.int32)
            call("pat_md_create", self.n, ptr(p), ptr(v), e.shape[0], ptr(e), float(box), int(pbc), float(cutoff),
                 float(dt), ctypes.byref(h))
        self.h = h
        if order != PAT_IR_ORDER_NODE:
            call("pat_md_set_row_order", self.h, order)
        if variant != PAT_IR_CSR or lanes_per_row:
            self.set_variant(variant, lanes_per_row)

    def set_variant(self, variant, lanes_per_row=0):
        call("pat_md_set_variant", self.h, variant, lanes_per_row)

    def set_sliced_form(self, form):
        """Sliced layout with loop form 0..4 (same bits for all)."""
        call("pat_md_set_sliced_form", self.h, form)

    def sliced_form(self):
        """The sliced loop form in use, or -1."""
        f = c_int(0)
        call("pat_md_get_sliced_form", self.h, ctypes.byref(f))
        return f.value

    def set_slice_order(self, order):
        """Rows of the sliced layout in node (0), row-class (1) or spatial
        (2) order (same bits for all)."""
        call("pat_md_set_slice_order", self.h, order)

    def slice_order(self):
        o = c_int(0)
        call("pat_md_get_slice_order", self.h, ctypes.byref(o))
        return o.value

    def variant(self):
        """(variant, lanes per row in use, row order)."""
        a, b, c = c_int(0), c_int(0), c_int(0)
        call("pat_md_get_variant", self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def run(self, nsteps):
        call("pat_md_run", self.h, nsteps)

    def run_timed(self, nsteps):
        """Returns (total_ms, per-step force-kernel ms array)."""
        tot = c_float(0)
        k = np.zeros(nsteps, dtype=np.float32)
        call("pat_md_run_timed", self.h, nsteps, ctypes.byref(tot), k.ctypes.data_as(c_fp))
        return tot.value, k

    def run_graph_timed(self, nsteps):
        """run() of a handle without rebuilds; returns the device time (ms)
        between HIP events around the step graph's replay."""
        ms = c_float(0)
        call("pat_md_run_graph_timed", self.h, nsteps, ctypes.byref(ms))
        return ms.value

    def run_async(self, nsteps, stream=None):
        call("pat_md_run_async", self.h, nsteps, stream_ptr(stream))

    def get(self):
        p = np.empty((self.n, 3), dtype=np.float32)
        v = np.empty((self.n, 3), dtype=np.float32)
        f = np.empty((self.n, 3), dtype=np.float32)
        call("pat_md_get", self.h, ptr(p), ptr(v), ptr(f))
        return p, v, f

    def stats(self):
        """(half pairs, steps, rebuilds, rebuild ms)."""
        a, b, c, d = c_i64(0), c_i64(0), c_i64(0), c_float(0)
        call("pat_md_stats", self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d))
        return a.value, b.value, c.value, d.value

    def csr(self):
        npairs = self.stats()[0]
        rp = np.empty(self.n + 1, dtype=np.int32)
        cl = np.empty(max(2 * npairs, 1), dtype=np.int32)
        call("pat_md_csr", self.h, ptr(rp), ptr(cl))
        return rp, cl[:2 * npairs]

    def energy(self):
        ep, ek = c_double(0), c_double(0)
        call("pat_md_energy", self.h, ctypes.byref(ep), ctypes.byref(ek))
        return ep.value, ek.value

    def close(self):
        if self.h:
            call("pat_md_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MDBatch:
    """Independent periodic LJ systems advanced as one (pat_mdb_*).
    systems: list of (pos (n, 3), vel (n, 3) or None, box)."""

    def __init__(self, systems, cutoff=2.5, skin=0.3, every=20, dt=0.005, lanes_per_row=0):
        self.sizes = [len(p) for p, _, _ in systems]
        off = np.zeros(len(systems) + 1, dtype=np.int64)
        off[1:] = np.cumsum(self.sizes)
        pos = np.ascontiguousarray(np.concatenate([p for p, _, _ in systems]), dtype=np.float32)
        vel = np.ascontiguousarray(np.concatenate([v if v is not None else np.zeros_like(p)
                                                   for p, v, _ in systems]), dtype=np.float32)
        box = np.array([b for _, _, b in systems], dtype=np.float64)
        h = c_vp()
        call("pat_mdb_create", len(systems), ptr(off), ptr(pos), ptr(vel), ptr(box), float(cutoff), float(skin),
             int(every), float(dt), ctypes.byref(h))
        self.h = h
        if lanes_per_row:
            call("pat_mdb_set_lanes", self.h, lanes_per_row)

    def run(self, nsteps):
        tot = c_float(0)
        call("pat_mdb_run", self.h, nsteps, ctypes.byref(tot), None)
        return tot.value

    def run_timed(self, nsteps):
        tot = c_float(0)
        k = np.zeros(nsteps, dtype=np.float32)
        call("pat_mdb_run", self.h, nsteps, ctypes.byref(tot), k.ctypes.data_as(c_fp))
        return tot.value, k

    def set_slice_order(self, order):
        call("pat_mdb_set_slice_order", self.h, order)

    def get(self, sys):
        n = self.sizes[sys]
        p, v, f = (np.empty((n, 3), dtype=np.float32) for _ in range(3))
        call("pat_mdb_get", self.h, sys, ptr(p), ptr(v), ptr(f))
        return p, v, f

    def stats(self):
        a, b, c, d = c_i64(0), c_i64(0), c_i64(0), c_float(0)
        call("pat_mdb_stats", self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d))
        return a.value, b.value, c.value, d.value

    def close(self):
        if self.h:
            call("pat_mdb_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_systems(nsys=64, lo=8, hi=32, seed=99, temperature=1.44):
    """nsys FCC systems with cells per side log-uniform in [lo, hi] (seeded),
    density 0.8442, Maxwell velocities at the given temperature."""
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=nsys))).astype(int)
    out = []
    for i, m in enumerate(sides):
        pos, box = gen_fcc(int(m), int(m), int(m), 0.8442, 0.0, seed + i)
        vel = gen_velocities(len(pos), temperature, seed + 1000 + i)
        out.append((pos, vel, float(box[0])))
    return out


def moldyn_system(ncells, rho=0.8442, jitter=0.05, cutoff=2.5, seed=42, shuffle=True):
    """FCC lattice of ncells^3 unit cells, jittered, with its half pair list
    within the cutoff (minimum image) in seeded random order."""
    pos, box = gen_fcc(ncells, ncells, ncells, rho, jitter, seed)
    edges = gen_pairs(pos, box[0], cutoff, shuffle=shuffle, seed=seed + 1)
    return pos, float(box[0]), edges


# ---------------------------------------------------------------- kmeans

def kmeans_fixed_shift(n, maxabs):
    v = c_int(0)
    call("pat_kmeans_fixed_shift", n, float(maxabs), ctypes.byref(v))
    return v.value


class KMeans:
    """Handle mirror of RegularRuntime running kmeans map1/reduce1 plus the
    Lloyd update (pat_kmeans_*).  x: host (N, D) float32 array or a
    DeviceArray (read in place)."""

    def __init__(self, x, k, mode=PAT_KM_MFMA):
        if isinstance(x, DeviceArray):
            self.n, self.d = x.shape
            src, on_dev = x.ptr, 1
            self._keep = x
        else:
            x = np.ascontiguousarray(x, dtype=np.float32)
            self.n, self.d = x.shape
            src, on_dev = x.ctypes.data, 0
        self.k = k
        h = c_vp()
        call("pat_kmeans_create", src, self.n, self.d, k, mode, on_dev, ctypes.byref(h))
        self.h = h

    def set_mode(self, mode):
        call("pat_kmeans_set_mode", self.h, mode)

    def set_terms(self, terms):
        """BF16 terms of the MFMA distance values (1 or 3; same bits)."""
        call("pat_kmeans_set_terms", self.h, terms)

    def terms(self):
        t = c_int(0)
        call("pat_kmeans_get_terms", self.h, ctypes.byref(t))
        return t.value

    def set_centroids(self, c):
        c = np.ascontiguousarray(c, dtype=np.float32)
        call("pat_kmeans_set_centroids", self.h, c.ctypes.data)

    def run(self, iters):
        call("pat_kmeans_run", self.h, iters)

    def run_timed(self, iters):
        tot = c_float(0)
        k = np.zeros(iters, dtype=np.float32)
        call("pat_kmeans_run_timed", self.h, iters, ctypes.byref(tot), k.ctypes.data_as(c_fp))
        return tot.value, k

    def get(self):
        c = np.empty((self.k, self.d), dtype=np.float32)
        a = np.empty(self.n, dtype=np.int32)
        md = np.empty(self.n, dtype=np.float32)
        call("pat_kmeans_get", self.h, c.ctypes.data, a.ctypes.data, md.ctypes.data)
        return c, a, md

    def centroids(self):
        c = np.empty((self.k, self.d), dtype=np.float32)
        call("pat_kmeans_get", self.h, c.ctypes.data, None, None)
        return c

    def get_object(self):
        sums = np.empty((self.k, self.d), dtype=np.int64)
        counts = np.empty(self.k, dtype=np.int64)
        dsums = np.empty(self.k, dtype=np.int64)
        s, ds = c_int(0), c_int(0)
        call("pat_kmeans_get_object", self.h, sums.ctypes.data, counts.ctypes.data, dsums.ctypes.data,
             ctypes.byref(s), ctypes.byref(ds))
        return sums, counts, dsums, s.value, ds.value

    def stats(self):
        r = c_i64(0)
        m = c_float(0)
        call("pat_kmeans_stats", self.h, ctypes.byref(r), ctypes.byref(m))
        return r.value, m.value

    def close(self):
        if self.h:
            call("pat_kmeans_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kmeans_shifts(n, d, maxabs):
    """(coordinate shift, distance shift) the handle derives from max |x|."""
    return (kmeans_fixed_shift(n, maxabs),
            kmeans_fixed_shift(n, float(np.float32(2.0 * float(maxabs) * np.sqrt(float(d))))))


# ---------------------------------------------------------------- device generators


def gen_heat3d_dev(nx, ny, nz, seed=42):
    """Halo-inclusive grid generated on the device (DeviceArray)."""
    g = DeviceArray((nz + 2, ny + 2, nx + 2), np.float32)
    call("pat_gen_heat3d_dev", nx, ny, nz, seed, g.ptr, None)
    sync()
    return g


def gen_fcc_dev(ncx, ncy, ncz, rho=0.8442, jitter=0.0, seed=1):
    """(float4 positions DeviceArray, box) generated on the device."""
    n = 4 * ncx * ncy * ncz
    p = DeviceArray((n, 4), np.float32)
    box = np.zeros(3)
    call("pat_gen_fcc_dev", ncx, ncy, ncz, rho, jitter, seed, p.ptr, box.ctypes.data_as(c_dp), None)
    sync()
    return p, box


def gen_gmm_dev(n, d, ncomp, sigma=1.0, lo=-10.0, hi=10.0, seed=1):
    x = DeviceArray((n, d), np.float32)
    call("pat_gen_gmm_dev", n, d, ncomp, sigma, lo, hi, seed, x.ptr, None)
    sync()
    return x


# ---------------------------------------------------------------- serving


def job(kind, size, steps, seed, dim=0, k=0, out=None):
    """A pat_job; out (float32 numpy array) receives the full result."""
    j = Job(kind=kind, size=size, dim=dim, k=k, steps=steps, seed=seed)
    if out is not None:
        j.out = out.ctypes.data
        j.out_len = out.size
    return j


def result_dict(r):
    return {"status": r.status, "kernel": r.kernel, "checksum": r.checksum, "submit_ms": r.t_submit_ms,
            "start_ms": r.t_start_ms, "end_ms": r.t_end_ms, "latency_ms": r.t_end_ms - r.t_submit_ms,
            "device_bytes": r.device_bytes, "batch": r.batch}


class Serve:
    """The serving loop (pat_serve_*): jobs admitted by device-memory budget
    and run concurrently on 'slots' streams."""

    def __init__(self, budget_bytes, slots=8, plain=False, fuse=True):
        h = c_vp()
        flags = (PAT_SERVE_PLAIN if plain else 0) | (0 if fuse else PAT_SERVE_NO_FUSE)
        call("pat_serve_start", int(budget_bytes), slots, flags, ctypes.byref(h))
        self.h = h
        self._keep = {}

    def submit(self, j, keep=None):
        i = c_i64(0)
        call("pat_serve_submit", self.h, ctypes.byref(j), ctypes.byref(i))
        if keep is not None:
            self._keep[i.value] = keep  # output buffer stays alive until the job is done
        return i.value

    def wait(self, job_id):
        r = JobResult()
        call("pat_serve_wait", self.h, job_id, ctypes.byref(r))
        self._keep.pop(job_id, None)
        return result_dict(r)

    def poll(self, job_id):
        d, r = c_int(0), JobResult()
        call("pat_serve_poll", self.h, job_id, ctypes.byref(d), ctypes.byref(r))
        return result_dict(r) if d.value else None

    def drain(self):
        call("pat_serve_drain", self.h)

    def stats(self):
        v = [c_i64(0) for _ in range(5)]
        call("pat_serve_stats", self.h, *[ctypes.byref(x) for x in v])
        return dict(zip(("submitted", "finished", "queued", "reserved", "arena_in_use"), [x.value for x in v]))

    def close(self):
        if self.h:
            call("pat_serve_stop", self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GR:
    """Generalized reduction with an additive map/reduce pair (pat_gr_*):
    PAT_GR_VSTATS over velocities (n x 3 FP64) with vaverh, or
    PAT_GR_LJ_ENERGY over edges (E x 2 int32) with (cutsq, positions)."""

    def __init__(self, map_idx, units, param):
        self.map_idx = map_idx
        u = np.ascontiguousarray(units, dtype=np.float64 if map_idx == PAT_GR_VSTATS else np.int32)
        self._units = u
        p = self._param_bytes(param)
        h = c_vp()
        call("pat_gr_create", ptr(u), u.shape[0], u.itemsize * u.shape[1], map_idx, ptr(p), p.nbytes,
             ctypes.byref(h))
        self.h = h

    def _param_bytes(self, param):
        if self.map_idx == PAT_GR_VSTATS:
            return np.array([param], dtype=np.float32)
        cutsq, x = param
        return np.concatenate([[float(cutsq)], np.ascontiguousarray(x, dtype=np.float64).ravel()])

    def set_param(self, param):
        p = self._param_bytes(param)
        call("pat_gr_set_param", self.h, ptr(p), p.nbytes)

    def run(self):
        call("pat_gr_run", self.h)

    def output(self):
        """(values: keys x values FP64, emits per key)."""
        k, v = c_int(0), c_int(0)
        call("pat_gr_get_output", self.h, ctypes.byref(k), ctypes.byref(v), None, None)
        vals = np.zeros((k.value, v.value), dtype=np.float64)
        em = np.zeros(k.value, dtype=np.int64)
        call("pat_gr_get_output", self.h, None, None, ptr(vals), ptr(em))
        return vals, em

    def close(self):
        if self.h:
            call("pat_gr_destroy", self.h)
            self.h = None


# ---------------------------------------------------------------- reference input files

class Regular:
    """Handle mirror of RegularRuntime with an application's device functor
    tables and keyed reduction objects (pat_regular_*).  data: the input
    (any numpy array, used as bytes); unit i is unit_bytes at offsets[i]
    (bytes), or at i * unit_bytes."""

    def __init__(self, data, num_units, unit_bytes, offsets=None, table=None):
        self._data = np.ascontiguousarray(data)
        self._off = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.int64)
        h = c_vp()
        call("pat_regular_create", ptr(self._data), self._data.nbytes, int(num_units), int(unit_bytes),
             ptr(self._off), ctypes.byref(h))
        self.h = h
        if table is not None:
            call("pat_regular_set_table", self.h, table)

    def set_capacity(self, max_keys, record_bytes, lds_keys=0):
        call("pat_regular_set_capacity", self.h, int(max_keys), int(record_bytes), int(lds_keys))

    def run(self, map_idx, reduce_idx=None, param=None):
        """One map + reduce pass; returns {key bytes: value bytes}."""
        call("pat_regular_set_map_idx", self.h, map_idx)
        call("pat_regular_set_reduce_idx", self.h, map_idx if reduce_idx is None else reduce_idx)
        self._param = None if param is None else np.ascontiguousarray(param)
        call("pat_regular_set_param", self.h, ptr(self._param), 0 if param is None else self._param.nbytes)
        call("pat_regular_start", self.h)
        return self.output()

    def output(self):
        nk, kb, vb = c_i64(0), c_i64(0), c_i64(0)
        call("pat_regular_get_output", self.h, ctypes.byref(nk), ctypes.byref(kb), ctypes.byref(vb), None, None,
             None, None, None, None)
        n = nk.value
        keys = np.zeros(max(kb.value, 1), np.uint8)
        vals = np.zeros(max(vb.value, 1), np.uint8)
        ki, ks, vi, vs = (np.zeros(max(n, 1), np.uint32) for _ in range(4))
        call("pat_regular_get_output", self.h, ctypes.byref(nk), ctypes.byref(kb), ctypes.byref(vb), ptr(keys),
             ptr(ki), ptr(ks), ptr(vals), ptr(vi), ptr(vs))
        return {keys[ki[i]:ki[i] + ks[i]].tobytes(): vals[vi[i]:vi[i] + vs[i]].tobytes() for i in range(n)}

    def stats(self):
        """(LDS merges, maps run again, wavefronts) of the last run."""
        m, r, w = c_i64(0), c_i64(0), c_int(0)
        call("pat_regular_stats", self.h, ctypes.byref(m), ctypes.byref(r), ctypes.byref(w))
        return m.value, r.value, w.value

    def close(self):
        if self.h:
            call("pat_regular_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mold(ctypes.Structure):
    _fields_ = [("box_cells", c_int), ("steps", c_int), ("neigh_update", c_int), ("cutoff", ctypes.c_double),
                ("mesh", ctypes.c_char * 512), ("xyz", ctypes.c_char * 512), ("percent", ctypes.c_double),
                ("start_time", c_int), ("update_freq", c_int)]


class LJIn(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("ntimes", c_int), ("nbinx", c_int),
                ("nbiny", c_int), ("nbinz", c_int), ("dt", ctypes.c_double), ("t_request", ctypes.c_double),
                ("rho", ctypes.c_double), ("every", c_int), ("cutforce", ctypes.c_double),
                ("cutneigh", ctypes.c_double), ("nstat", c_int)]


def _struct_dict(st):
    return {k: (getattr(st, k).decode() if isinstance(getattr(st, k), bytes) else getattr(st, k))
            for k, _ in st._fields_}


def read_mold(path):
    """moldyn's parameter file as a dict."""
    m = Mold()
    call("pat_read_mold", os.fsencode(path), ctypes.byref(m))
    return _struct_dict(m)


def read_ljin(path):
    """miniMD's lj.in as a dict."""
    m = LJIn()
    call("pat_read_ljin", os.fsencode(path), ctypes.byref(m))
    return _struct_dict(m)


def read_mesh(path, zero_based=True):
    """(num_nodes, edges (E x 2 int32)) from a moldyn .mesh graph file."""
    nn, ne = c_i64(0), c_i64(0)
    call("pat_read_mesh", os.fsencode(path), int(zero_based), ctypes.byref(nn), ctypes.byref(ne), None, 0)
    e = np.zeros((ne.value, 2), dtype=np.int32)
    call("pat_read_mesh", os.fsencode(path), int(zero_based), ctypes.byref(nn), ctypes.byref(ne), ptr(e), ne.value)
    return nn.value, e


def read_xyz(path, num_nodes):
    x = np.zeros((num_nodes, 3), dtype=np.float64)
    call("pat_read_xyz", os.fsencode(path), num_nodes, ptr(x))
    return x
