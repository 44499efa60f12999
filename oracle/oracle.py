"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, to compare against.  The product never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_float = ctypes.c_float
c_double = ctypes.c_double
c_vp = ctypes.c_void_p

_SIG = {
    "orc_heat3d_step_f32": (None, [c_vp, c_vp, c_int, c_int, c_int, c_float]),
    "orc_heat3d_step_f64": (None, [c_vp, c_vp, c_int, c_int, c_int, c_double]),
    "orc_heat3d_run_f32": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_float]),
    "orc_heat3d_run_f64": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_double]),
    "orc_heat3d_run_f32_tiled": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_float, c_int]),
    "orc_star_run_f32": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_int]),
    "orc_heat_ref_run_f64": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int]),
    "orc_ir_lj_ref": (None, [c_i64, c_i64, c_vp, c_vp, c_double, c_double, c_int, c_vp]),
    "orc_lj_force_f32": (c_double, [c_i64, c_i64, c_vp, c_vp, c_float, c_int, c_float, c_vp]),
    "orc_lj_force_f64": (c_double, [c_i64, c_i64, c_vp, c_vp, c_double, c_int, c_double, c_vp]),
    "orc_lj_force_f32_threads": (None, [c_i64, c_i64, c_vp, c_vp, c_float, c_int, c_float, c_vp, c_vp, c_int]),
    "orc_md_run_f32": (c_double, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_float, c_int, c_float, c_float, c_int,
                                  c_int, c_vp]),
    "orc_md_run_f64": (c_double, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_double, c_int, c_double, c_double, c_int]),
    "orc_pairs_brute": (c_i64, [c_i64, c_vp, c_float, c_int, c_float, c_vp, c_i64]),
    "orc_pairs_cells": (c_i64, [c_i64, c_vp, c_float, c_float, c_vp, c_i64]),
    "orc_lj_force_f32_gather": (None, [c_i64, c_i64, c_vp, c_vp, c_float, c_int, c_float, c_vp, c_int]),
    "orc_md_run_rebuild_f32": (c_double, [c_i64, c_vp, c_vp, c_vp, c_float, c_float, c_float, c_int, c_float, c_int,
                                          c_int, ctypes.POINTER(c_i64)]),
    "orc_lj_force_f64_gather": (None, [c_i64, c_i64, c_vp, c_vp, c_double, c_int, c_double, c_vp, c_int]),
    "orc_md_run_rebuild_f64": (c_double, [c_i64, c_vp, c_vp, c_vp, c_double, c_double, c_double, c_int, c_double,
                                          c_int, c_int, ctypes.POINTER(c_i64)]),
    "orc_kmeans_assign_f32": (None, [c_i64, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_int]),
    "orc_kmeans_assign_f64": (None, [c_i64, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_int]),
    "orc_kmeans_fixed_shift": (c_int, [c_i64, c_float]),
    "orc_kmeans_accumulate": (None, [c_i64, c_int, c_int, c_vp, c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_vp, c_int]),
    "orc_kmeans_update": (None, [c_int, c_int, c_vp, c_vp, c_int, c_vp]),
    "orc_kmeans_lloyd": (None, [c_i64, c_int, c_int, c_vp, c_int, c_int, c_vp, c_vp, c_vp, c_int]),
    "orc_kmeans_lloyd_f64": (None, [c_i64, c_int, c_int, c_vp, c_int, c_vp, c_vp, c_int]),
    "orc_gr_vstats": (None, [c_i64, c_vp, c_float, c_vp, c_vp]),
    "orc_gr_lj_energy": (c_double, [c_i64, c_vp, c_vp, c_double, ctypes.POINTER(c_i64)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built at {LIB_PATH}; run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


# ---------------------------------------------------------------- stencil

def heat3d(grid, nsteps, alpha=0.1, dtype=np.float32, threads=0):
    """Runs nsteps of the heat rule on a halo-inclusive (z, y, x) grid."""
    nz, ny, nx = (s - 2 for s in grid.shape)
    a = np.ascontiguousarray(grid, dtype=dtype).copy()
    b = a.copy()
    if dtype == np.float32:
        if threads:
            r = lib().orc_heat3d_run_f32_tiled(_p(a), _p(b), nx, ny, nz, nsteps, alpha, threads)
        else:
            r = lib().orc_heat3d_run_f32(_p(a), _p(b), nx, ny, nz, nsteps, alpha)
    else:
        r = lib().orc_heat3d_run_f64(_p(a), _p(b), nx, ny, nz, nsteps, alpha)
    return b if r else a


def star(grid, nsteps, w=3):
    """nsteps of the W = w star (the kmeans directory's stencil2) on a
    halo-inclusive (z, y, x) FP32 grid of halo width w."""
    nz, ny, nx = (d - 2 * w for d in grid.shape)
    a = np.ascontiguousarray(grid, dtype=np.float32).copy()
    b = a.copy()
    r = lib().orc_star_run_f32(_p(a), _p(b), nx, ny, nz, w, nsteps)
    return b if r else a


def heat_ref(grid, nsteps):
    """nsteps of heat_3D's own FP64 functor on a halo-inclusive (z, y, x)
    grid of halo width 1."""
    nz, ny, nx = (d - 2 for d in grid.shape)
    a = np.ascontiguousarray(grid, dtype=np.float64).copy()
    b = a.copy()
    r = lib().orc_heat_ref_run_f64(_p(a), _p(b), nx, ny, nz, nsteps)
    return b if r else a


# ---------------------------------------------------------------- LJ

def lj_force(pos, edges, box, rc, pbc=True, dtype=np.float32):
    """Serial-scatter LJ forces over a half edge list; returns (f, epot)."""
    n = pos.shape[0]
    e = np.ascontiguousarray(edges, dtype=np.int32)
    p = np.ascontiguousarray(pos, dtype=dtype)
    f = np.zeros((n, 3), dtype=dtype)
    fn = lib().orc_lj_force_f32 if dtype == np.float32 else lib().orc_lj_force_f64
    ep = fn(n, e.shape[0], _p(e), _p(p), box, int(pbc), rc * rc if dtype != np.float32 else np.float32(rc) ** 2,
            _p(f))
    return f, ep


def ir_lj_ref(pos, edges, box, cut2, form):
    """The reference's LJ functors in their own FP64 arithmetic (form 0:
    moldyn's map with the minimum image; 1: miniMD's map1, no images), serial
    over the edge list with the reduction object's copy-then-add semantics;
    returns N x 3 float32."""
    n = pos.shape[0]
    p = np.ascontiguousarray(pos, dtype=np.float64)
    e = np.ascontiguousarray(edges, dtype=np.int32)
    f = np.zeros((n, 3), dtype=np.float32)
    lib().orc_ir_lj_ref(n, e.shape[0], _p(e), _p(p), float(box), float(cut2), int(form), _p(f))
    return f


def lj_force_threads(pos, edges, box, rc, pbc=True, threads=0):
    n = pos.shape[0]
    e = np.ascontiguousarray(edges, dtype=np.int32)
    p = np.ascontiguousarray(pos, dtype=np.float32)
    f = np.zeros((n, 3), dtype=np.float32)
    nt = threads or os.cpu_count()
    scratch = np.empty((nt, n, 3), dtype=np.float32)
    lib().orc_lj_force_f32_threads(n, e.shape[0], _p(e), _p(p), box, int(pbc), np.float32(rc) ** 2, _p(f),
                                   _p(scratch), nt)
    return f


def md_run(pos, vel, edges, box, rc, dt, nsteps, pbc=True, dtype=np.float32, threads=1, refcpu=False):
    """Velocity-Verlet MD; returns (pos, vel, force, epot).  Forces sum in the
    serial scatter's order for any thread count, unless refcpu selects the
    reference CPU engine's chunked thread-private accumulation (baseline
    timing).  epot is only computed by the single-thread serial path."""
    n = pos.shape[0]
    e = np.ascontiguousarray(edges, dtype=np.int32)
    p = np.ascontiguousarray(pos, dtype=dtype).copy()
    v = np.ascontiguousarray(vel, dtype=dtype).copy()
    f = np.zeros((n, 3), dtype=dtype)
    if dtype == np.float32:
        scratch = np.empty((threads, n, 3), dtype=np.float32) if refcpu else None
        ep = lib().orc_md_run_f32(n, e.shape[0], _p(e), _p(p), _p(v), _p(f), box, int(pbc), np.float32(rc) ** 2,
                                  dt, nsteps, threads, _p(scratch))
    else:
        ep = lib().orc_md_run_f64(n, e.shape[0], _p(e), _p(p), _p(v), _p(f), box, int(pbc), rc * rc, dt, nsteps)
    return p, v, f, ep


def lj_force_gather(pos, edges, box, rc, pbc=True, threads=0):
    n = pos.shape[0]
    e = np.ascontiguousarray(edges, dtype=np.int32)
    p = np.ascontiguousarray(pos, dtype=np.float32)
    f = np.zeros((n, 3), dtype=np.float32)
    lib().orc_lj_force_f32_gather(n, e.shape[0], _p(e), _p(p), box, int(pbc), np.float32(rc) ** 2, _p(f),
                                  threads or os.cpu_count())
    return f


def pairs_cells(pos, box, cut):
    n = pos.shape[0]
    p = np.ascontiguousarray(pos, dtype=np.float32)
    cnt = lib().orc_pairs_cells(n, _p(p), box, cut, None, 0)
    if cnt < 0:
        raise ValueError("box too small for the cell list")
    e = np.empty((max(cnt, 1), 2), dtype=np.int32)
    lib().orc_pairs_cells(n, _p(p), box, cut, _p(e), cnt)
    return e[:cnt]


def md_run_rebuild(pos, vel, box, rc, cutneigh, every, dt, nsteps, threads=0):
    """MD with the pair list rebuilt (cell list within cutneigh) before the
    force of step 0 and of every step index multiple of 'every'; forces in
    serial order.  Returns (pos, vel, force, last pair count)."""
    n = pos.shape[0]
    p = np.ascontiguousarray(pos, dtype=np.float32).copy()
    v = np.ascontiguousarray(vel, dtype=np.float32).copy()
    f = np.zeros((n, 3), dtype=np.float32)
    npairs = c_i64(0)
    r = lib().orc_md_run_rebuild_f32(n, _p(p), _p(v), _p(f), box, np.float32(rc) ** 2, cutneigh, every, dt, nsteps,
                                     threads or os.cpu_count(), ctypes.byref(npairs))
    if r != r:
        raise ValueError("box too small for the cell list")
    return p, v, f, npairs.value


def md_run_rebuild_f64(pos, vel, box, rc, cutneigh, every, dt, nsteps, threads=0):
    """FP64 shadow of md_run_rebuild (same schedule; pair lists built from the
    positions rounded to FP32).  Returns (pos, vel, force, last pair count) in
    FP64."""
    n = pos.shape[0]
    p = np.ascontiguousarray(pos, dtype=np.float64).copy()
    v = np.ascontiguousarray(vel, dtype=np.float64).copy()
    f = np.zeros((n, 3), dtype=np.float64)
    npairs = c_i64(0)
    r = lib().orc_md_run_rebuild_f64(n, _p(p), _p(v), _p(f), float(box), float(rc) ** 2, float(cutneigh), every,
                                     float(dt), nsteps, threads or os.cpu_count(), ctypes.byref(npairs))
    if r != r:
        raise ValueError("box too small for the cell list")
    return p, v, f, npairs.value


def lj_force_f64_gather(pos, edges, box, rc, pbc=True, threads=0):
    """FP64 forces over a half edge list, threaded (same sums as lj_force with
    dtype float64)."""
    n = pos.shape[0]
    p = np.ascontiguousarray(pos, dtype=np.float64)
    e = np.ascontiguousarray(edges, dtype=np.int32)
    f = np.zeros((n, 3), dtype=np.float64)
    lib().orc_lj_force_f64_gather(n, e.shape[0], _p(e), _p(p), float(box), int(pbc), float(rc) ** 2, _p(f),
                                  threads or os.cpu_count())
    return f


def pairs_brute(pos, box, cut, pbc=True):
    n = pos.shape[0]
    p = np.ascontiguousarray(pos, dtype=np.float32)
    cnt = lib().orc_pairs_brute(n, _p(p), box, int(pbc), np.float32(cut) ** 2, None, 0)
    e = np.empty((max(cnt, 1), 2), dtype=np.int32)
    lib().orc_pairs_brute(n, _p(p), box, int(pbc), np.float32(cut) ** 2, _p(e), cnt)
    return e[:cnt]


# ---------------------------------------------------------------- kmeans

def kmeans_assign(x, c, threads=0):
    n, d = x.shape
    k = c.shape[0]
    a = np.empty(n, dtype=np.int32)
    md = np.empty(n, dtype=np.float32)
    lib().orc_kmeans_assign_f32(n, d, k, _p(np.ascontiguousarray(x, np.float32)),
                                _p(np.ascontiguousarray(c, np.float32)), _p(a), _p(md), threads)
    return a, md


def kmeans_assign_f64(x, c, threads=0):
    n, d = x.shape
    k = c.shape[0]
    a = np.empty(n, dtype=np.int32)
    md = np.empty(n, dtype=np.float64)
    lib().orc_kmeans_assign_f64(n, d, k, _p(np.ascontiguousarray(x, np.float32)),
                                _p(np.ascontiguousarray(c, np.float64)), _p(a), _p(md), threads)
    return a, md


def kmeans_shift(n, maxabs):
    return lib().orc_kmeans_fixed_shift(n, float(maxabs))


def kmeans_accumulate(x, assign, k, shift, mindist=None, dshift=0, threads=0):
    n, d = x.shape
    sums = np.empty((k, d), dtype=np.int64)
    counts = np.empty(k, dtype=np.int64)
    dsum = np.empty(k, dtype=np.int64) if mindist is not None else None
    lib().orc_kmeans_accumulate(n, d, k, _p(np.ascontiguousarray(x, np.float32)),
                                _p(np.ascontiguousarray(assign, np.int32)),
                                _p(mindist) if mindist is not None else None, shift, dshift, _p(sums), _p(counts),
                                _p(dsum), threads)
    return sums, counts, dsum


def kmeans_update(sums, counts, shift, c):
    k, d = sums.shape
    c2 = np.ascontiguousarray(c, np.float32).copy()
    lib().orc_kmeans_update(d, k, _p(sums), _p(counts), shift, _p(c2))
    return c2


def kmeans_lloyd(x, k, iters, shift, threads=0, sse=False):
    n, d = x.shape
    x = np.ascontiguousarray(x, np.float32)
    c = x[:k].copy()
    a = np.empty(n, dtype=np.int32)
    s = np.empty(iters, dtype=np.float64) if sse else None
    lib().orc_kmeans_lloyd(n, d, k, _p(x), iters, shift, _p(c), _p(a), _p(s), threads)
    return (c, a, s) if sse else (c, a)


def kmeans_lloyd_f64(x, k, iters, threads=0):
    n, d = x.shape
    x = np.ascontiguousarray(x, np.float32)
    c = x[:k].astype(np.float64)
    a = np.empty(n, dtype=np.int32)
    lib().orc_kmeans_lloyd_f64(n, d, k, _p(x), iters, _p(c), _p(a), threads)
    return c, a


# ---------------------------------------------------------------- the reference's own miniMD code

REF_PATH = os.path.join(_HERE, "_ref", "libref_minimd.so")


def ref_available():
    return os.path.exists(REF_PATH)


def ref_minimd_system(n, rho=0.8442, nbin=None, cutneigh=2.8, cutforce=2.5):
    """miniMD's own create_atoms + Neighbor::build + Force::compute (built
    from the reference sources into oracle/_ref).  Returns (pos, force,
    pairs) in FP64; the reference's force omits the 48*eps prefactor."""
    L = ctypes.CDLL(REF_PATH)
    f = L.ref_minimd_system
    f.restype = c_int
    f.argtypes = [c_int, c_double, c_int, c_double, c_double, c_vp, c_vp, ctypes.POINTER(c_int), c_vp,
                  ctypes.POINTER(ctypes.c_long)]
    na = 4 * n ** 3
    pos = np.empty((na, 3))
    force = np.empty((na, 3))
    cap = na * 80
    pairs = np.empty((cap, 2), dtype=np.int32)
    nat = c_int(na)
    npairs = ctypes.c_long(cap)
    rc = f(n, rho, nbin or n, cutneigh, cutforce, _p(pos), _p(force), ctypes.byref(nat), _p(pairs),
           ctypes.byref(npairs))
    if rc != 0:
        raise RuntimeError(f"ref_minimd_system failed ({rc})")
    return pos[:nat.value], force[:nat.value], pairs[:npairs.value]


# ---------------------------------------------------------------- additive generalized reductions

def gr_vstats(vel, vaverh, per_unit=False):
    """moldyn velocity statistics (sum of |v| as FP32 values, count above
    vaverh); returns (sum, count) or ((sum, count), per-unit values)."""
    v = np.ascontiguousarray(vel, dtype=np.float64)
    out = np.zeros(2, dtype=np.float64)
    per = np.zeros((len(v), 2), dtype=np.float32) if per_unit else None
    lib().orc_gr_vstats(len(v), _p(v), np.float32(vaverh), _p(out), _p(per))
    return ((out[0], out[1]), per) if per_unit else (out[0], out[1])


def gr_lj_energy(edges, x, cutsq):
    """miniMD map2 pair energy sum over the edges within cutsq (FP64, no
    minimum image); returns (sum, pairs within the cutoff)."""
    e = np.ascontiguousarray(edges, dtype=np.int32)
    p = np.ascontiguousarray(x, dtype=np.float64)
    n = c_i64(0)
    s = lib().orc_gr_lj_energy(len(e), _p(e), _p(p), float(cutsq), ctypes.byref(n))
    return s, n.value
