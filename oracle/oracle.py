"""CPU oracle for the grid-method step -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker / the timed CPU baseline.  The
product package (barnes-hut_amd/) never imports it.

Two restatements of python-bh's sequential step live here:

* ``libbh_oracle.so`` (bh_oracle.c): FP64 C restatement of every phase,
  with OpenMP over target particles for the forces.
* numpy restatements of the placement phase that replay the reference's own
  numpy calls (``place_numpy``) -- they pin the C placement bit for bit.

References (python-bh/...):
  bounding box / sizing   barneshut/kernels/helpers/grid.py:5-26,
                          barneshut/implementations/base.py:211-239
  cell ids + sort         barneshut/implementations/sequential.py:32-50
  COM                     barneshut/internals/cloud.py:9-17
  forces                  barneshut/kernels/grid_decomposition/sequential/evaluation.py:73-109,
                          barneshut/kernels/gravity/cpu_numba.py:30-92,
                          barneshut/kernels/grid_decomposition/gpu/grid.py:116-268
  direct sum              barneshut/implementations/base.py:16-32
  cloud-to-cloud gravity  barneshut/kernels/gravity/cpu_numba.py:8-92 (the gravity plugin)
  grid potential          (no reference counterpart: the energy diagnostic of long runs)
  tick                    barneshut/kernels/grid_decomposition/gpu/grid.py:159-174
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbh_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compile libbh_oracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.bho_grid_dim.restype = ctypes.c_int
        L.bho_grid_dim.argtypes = [ctypes.c_int64, ctypes.c_int64]
        L.bho_bbox.argtypes = [ctypes.c_int64, _dp, _dp, ctypes.c_int, _dp]
        L.bho_place.argtypes = [ctypes.c_int64, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, _i32p, _i32p, _i32p]
        L.bho_sort_perm.argtypes = [ctypes.c_int64, _i32p, ctypes.POINTER(_dp), _i64p]
        L.bho_counts.argtypes = [ctypes.c_int64, _i32p, ctypes.c_int64, _i64p, _i64p]
        L.bho_summarize.argtypes = [ctypes.c_int64, _i64p, _dp, _dp, _dp, _dp]
        L.bho_forces.argtypes = [ctypes.c_int, _i64p, _i32p, ctypes.c_int64, _dp, _dp, _dp, _dp,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 _i64p, ctypes.c_int64, ctypes.c_int, _dp, _dp]
        L.bho_direct.argtypes = [ctypes.c_int64, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, _i64p, ctypes.c_int64, ctypes.c_int, _dp, _dp]
        L.bho_tick.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double]
        L.bho_interactions.argtypes = [ctypes.c_int, _i64p, _i64p]
        L.bho_energy.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, _dp]
        L.bho_potential.argtypes = [ctypes.c_int, _i64p, _i32p, ctypes.c_int64, _dp, _dp, _dp,
                                    _dp, ctypes.c_double, _i64p, ctypes.c_int64, ctypes.c_int,
                                    _dp]
        L.bho_cloud_gravity.argtypes = [ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int64, _dp,
                                        ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i32(a):
    return a.ctypes.data_as(_i32p)


def _i64(a):
    return a.ctypes.data_as(_i64p)


def _f64c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# C restatement
# --------------------------------------------------------------------------
def grid_dim(n: int, ppb: int) -> int:
    """base.py:221-231 + helpers/grid.py:5-8."""
    return int(lib().bho_grid_dim(int(n), int(ppb)))


def bbox(px, py, G):
    """helpers/grid.py:10-26, base.py:237 -> (min_x, min_y, max_x, max_y, step)."""
    px, py = _f64c(px), _f64c(py)
    out = np.zeros(5)
    lib().bho_bbox(len(px), _d(px), _d(py), int(G), _d(out))
    return tuple(float(v) for v in out)


def place(px, py, min_x, min_y, step, G):
    """sequential.py:32-34 -> (gx, gy, key) int32 arrays, key = gx*G + gy."""
    px, py = _f64c(px), _f64c(py)
    n = len(px)
    gx = np.empty(n, np.int32)
    gy = np.empty(n, np.int32)
    key = np.empty(n, np.int32)
    lib().bho_place(n, _d(px), _d(py), float(min_x), float(min_y), float(step), int(G),
                    _i32(gx), _i32(gy), _i32(key))
    return gx, gy, key


def sort_perm(key, fields):
    """sequential.py:37 permutation; fields = (px, py, m, vx, vy, ax, ay, pid)."""
    key = np.ascontiguousarray(key, dtype=np.int32)
    fs = [_f64c(f) for f in fields]
    assert len(fs) == 8
    arr = (_dp * 8)(*[_d(f) for f in fs])
    perm = np.empty(len(key), np.int64)
    lib().bho_sort_perm(len(key), _i32(key), arr, _i64(perm))
    return perm


def _check_keys(key, ncells):
    """The C routines index per-cell arrays by key: refuse keys outside [0, ncells)."""
    if len(key) and (int(key.min()) < 0 or int(key.max()) >= ncells):
        raise ValueError(f"cell keys must lie in [0, {ncells})")


def _check_offsets(offsets, n):
    """Exclusive per-cell offsets: start at 0, non-decreasing, end at n."""
    if len(offsets) < 1 or offsets[0] != 0 or offsets[-1] != n or np.any(np.diff(offsets) < 0):
        raise ValueError(f"offsets must rise from 0 to {n}")


def counts_offsets(key_sorted, ncells):
    """Per-cell counts and exclusive offsets (ncells+1) in x-major key order."""
    key_sorted = np.ascontiguousarray(key_sorted, dtype=np.int32)
    _check_keys(key_sorted, ncells)
    c = np.empty(ncells, np.int64)
    o = np.empty(ncells + 1, np.int64)
    lib().bho_counts(len(key_sorted), _i32(key_sorted), int(ncells), _i64(c), _i64(o))
    return c, o


def summarize(offsets, x, y, m):
    """cloud.py:9-17 per cell -> com (ncells, 3) = (X, Y, M); empty -> 0."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    x, y, m = _f64c(x), _f64c(y), _f64c(m)
    _check_offsets(offsets, len(x))
    if not len(x) == len(y) == len(m):
        raise ValueError("x, y and m must have one entry per particle")
    ncells = len(offsets) - 1
    com = np.empty((ncells, 3))
    lib().bho_summarize(ncells, _i64(offsets), _d(x), _d(y), _d(m), _d(com))
    return com


def forces(G, offsets, key_sorted, x, y, m, com, Gc=1.0, eps2=0.0, r2_cut=0.0,
           targets=None, nthreads=0):
    """Grid-method accelerations (near 3x3 direct + far non-empty COMs)."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    key_sorted = np.ascontiguousarray(key_sorted, dtype=np.int32)
    x, y, m = _f64c(x), _f64c(y), _f64c(m)
    com = _f64c(com)
    n = len(x)
    G = int(G)
    if not len(y) == len(m) == len(key_sorted) == n or len(offsets) != G * G + 1 \
            or com.size != 3 * G * G:
        raise ValueError("forces: arrays must match n particles and G*G cells")
    _check_offsets(offsets, n)
    _check_keys(key_sorted, G * G)
    if targets is not None:
        targets = np.ascontiguousarray(targets, dtype=np.int64)
        if len(targets) and (int(targets.min()) < 0 or int(targets.max()) >= n):
            raise ValueError(f"targets must lie in [0, {n})")
        nt = len(targets)
        tp = _i64(targets)
    else:
        nt = n
        tp = None
    ax = np.empty(nt)
    ay = np.empty(nt)
    lib().bho_forces(int(G), _i64(offsets), _i32(key_sorted), n, _d(x), _d(y), _d(m), _d(com),
                     float(Gc), float(eps2), float(r2_cut), tp, nt, int(nthreads), _d(ax), _d(ay))
    return ax, ay


def direct(x, y, m, Gc=1.0, eps2=0.0, r2_cut=0.0, targets=None, nthreads=0):
    """base.py:16-32 O(S*N) direct sum (self skipped by index)."""
    x, y, m = _f64c(x), _f64c(y), _f64c(m)
    n = len(x)
    if targets is not None:
        targets = np.ascontiguousarray(targets, dtype=np.int64)
        nt = len(targets)
        tp = _i64(targets)
    else:
        nt = n
        tp = None
    ax = np.empty(nt)
    ay = np.empty(nt)
    lib().bho_direct(n, _d(x), _d(y), _d(m), float(Gc), float(eps2), float(r2_cut), tp, nt,
                     int(nthreads), _d(ax), _d(ay))
    return ax, ay


def energy(x, y, m, vx, vy, Gc=1.0, eps2=0.0, nthreads=0):
    """(U, K): U = Gc * sum_{i<j} m_i m_j log(r^2 + eps2) / 2, K = sum m v^2 / 2.
    No reference counterpart (python-bh reports no energy)."""
    x, y, m, vx, vy = (_f64c(a) for a in (x, y, m, vx, vy))
    out = np.zeros(2)
    lib().bho_energy(len(x), _d(x), _d(y), _d(m), _d(vx), _d(vy), float(Gc), float(eps2),
                     int(nthreads), _d(out))
    return float(out[0]), float(out[1])


def potential(G, offsets, key_sorted, x, y, m, com, eps2=0.0, targets=None, nthreads=0):
    """Grid-method potential per target: phi_i = sum over the 3x3 cells' other
    particles of m_j log(r^2 + eps2) plus sum over the other non-empty cells of
    M_k log(|X_k - x_i|^2 + eps2).  U_grid = Gc / 4 * sum_i m_i phi_i.  No
    reference counterpart (python-bh reports no energy)."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    key_sorted = np.ascontiguousarray(key_sorted, dtype=np.int32)
    x, y, m, com = _f64c(x), _f64c(y), _f64c(m), _f64c(com)
    n = len(x)
    G = int(G)
    if not len(y) == len(m) == len(key_sorted) == n or len(offsets) != G * G + 1 \
            or com.size != 3 * G * G:
        raise ValueError("potential: arrays must match n particles and G*G cells")
    _check_offsets(offsets, n)
    _check_keys(key_sorted, G * G)
    if targets is not None:
        targets = np.ascontiguousarray(targets, dtype=np.int64)
        if len(targets) and (int(targets.min()) < 0 or int(targets.max()) >= n):
            raise ValueError(f"targets must lie in [0, {n})")
        nt, tp = len(targets), _i64(targets)
    else:
        nt, tp = n, None
    phi = np.empty(nt)
    lib().bho_potential(G, _i64(offsets), _i32(key_sorted), n, _d(x), _d(y), _d(m), _d(com),
                        float(eps2), tp, nt, int(nthreads), _d(phi))
    return phi


def tick(x, y, vx, vy, ax, ay, dt):
    """gpu/grid.py:159-174, in place on float64 arrays."""
    for a in (x, y, vx, vy, ax, ay):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    lib().bho_tick(len(x), _d(x), _d(y), _d(vx), _d(vy), _d(ax), _d(ay), float(dt))


def cloud_gravity(A, B=None, Gc=1.0, eps2=0.0, r2_cut=1e-5, update_other=False):
    """cpu_numba.py:8-92: the fields cpu_numba_kernel adds to the clouds.

    A, B: (n, ld >= 3) float64 rows (px, py, mass first).  B is None: one
    cloud (self_self_numba).  Returns (accA, accB) as (n, 2) arrays; accB is
    None unless B is given and update_other."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    if A.ndim != 2 or A.shape[1] < 3:
        raise ValueError("rows need shape (n, >= 3)")
    accA = np.zeros((len(A), 2))
    same = B is None
    if same:
        B = A
    B = np.ascontiguousarray(B, dtype=np.float64)
    if B.ndim != 2 or B.shape[1] < 3:
        raise ValueError("rows need shape (n, >= 3)")
    upd = bool(update_other) and not same
    accB = np.zeros((len(B), 2)) if upd else None
    lib().bho_cloud_gravity(len(A), _d(A), A.shape[1], len(B), _d(B), B.shape[1], int(same),
                            float(Gc), float(eps2), float(r2_cut), int(upd), _d(accA),
                            _d(accB) if upd else _d(np.zeros(2)))
    return accA, accB


def interactions(G, counts):
    """Exact (near, far) interaction counts of one evaluation."""
    counts = np.ascontiguousarray(counts, dtype=np.int64)
    out = np.zeros(2, np.int64)
    lib().bho_interactions(int(G), _i64(counts), _i64(out))
    return int(out[0]), int(out[1])


# --------------------------------------------------------------------------
# Whole step
# --------------------------------------------------------------------------
def step(px, py, m, vx, vy, pid, G, Gc=1.0, dt=0.0, eps2=0.0, r2_cut=0.0,
         ax=None, ay=None, do_tick=True, nthreads=0):
    """One create_tree -> summarize -> evaluate -> timestep pass
    (base.py:94-135).  Returns a dict with the sorted state, placement
    outputs, COMs, pre-tick accelerations and (if do_tick) the post-tick
    state.  All arrays are in the new sorted order."""
    n = len(px)
    ax = np.zeros(n) if ax is None else _f64c(ax)
    ay = np.zeros(n) if ay is None else _f64c(ay)
    pidf = np.asarray(pid, dtype=np.float64)
    bb = bbox(px, py, G)
    gx, gy, key = place(px, py, bb[0], bb[1], bb[4], G)
    perm = sort_perm(key, (px, py, m, vx, vy, ax, ay, pidf))
    s = lambda a: np.ascontiguousarray(np.asarray(a)[perm])
    out = dict(bbox=bb, perm=perm, key=s(key), gx=s(gx), gy=s(gy),
               px=s(px), py=s(py), m=s(m), vx=s(vx), vy=s(vy),
               ax_in=s(ax), ay_in=s(ay), pid=s(pid))
    cnt, off = counts_offsets(out["key"], G * G)
    out["counts"], out["offsets"] = cnt, off
    com = summarize(off, out["px"], out["py"], out["m"])
    out["com"] = com
    fx, fy = forces(G, off, out["key"], out["px"], out["py"], out["m"], com, Gc, eps2, r2_cut,
                    nthreads=nthreads)
    # evaluate accumulates into the existing accelerations (gpu/grid.py:227-230)
    out["ax"] = out["ax_in"] + fx
    out["ay"] = out["ay_in"] + fy
    if do_tick:
        x2, y2 = out["px"].copy(), out["py"].copy()
        vx2, vy2 = out["vx"].copy(), out["vy"].copy()
        ax2, ay2 = out["ax"].copy(), out["ay"].copy()
        tick(x2, y2, vx2, vy2, ax2, ay2, dt)
        out.update(px_next=x2, py_next=y2, vx_next=vx2, vy_next=vy2)
    return out


# --------------------------------------------------------------------------
# numpy replay of the reference's own placement calls
# --------------------------------------------------------------------------
FIELDS = "f8,f8,f8,f8,f8,f8,f8,f8,f8,f8"   # internals/particle.py:14


def next_perfect_square_numpy(n):
    """helpers/grid.py:5-8 with numpy's sqrt/power/ceil."""
    if n % np.sqrt(n) != 0:
        return np.power(np.ceil(np.sqrt(n)), 2)
    return n


def grid_dim_numpy(n, ppb):
    """base.py:221-231."""
    if ppb == 1:
        nleaves = ppb
    else:
        nleaves = n / (0.7 * ppb)
    nleaves = next_perfect_square_numpy(nleaves)
    return int(nleaves ** 0.5)


def place_numpy(particles, G):
    """Replays helpers/grid.py:10-26, base.py:237 and sequential.py:32-37 on an
    N x 10 AoS array (columns px,py,mass,vx,vy,ax,ay,gx,gy,pid); returns the
    sorted copy and (min_x, min_y, step)."""
    a = np.array(particles, dtype=np.float64, copy=True)
    pts = a[:, 0:2]
    max_x, max_y = np.max(pts, axis=0)[:2]
    min_x, min_y = np.min(pts, axis=0)[:2]
    x_edge, y_edge = max_x - min_x, max_y - min_y
    if x_edge > y_edge:
        max_y += (x_edge - y_edge)
    elif y_edge > x_edge:
        max_x += (y_edge - x_edge)
    bottom_left = np.array((min_x, min_y))
    top_right = np.array((max_x, max_y))
    step = (top_right[0] - bottom_left[0]) / G
    a[:, 7:9] = a[:, 0:2]
    a[:, 7:9] = (a[:, 7:9] - bottom_left) / step
    a[:, 7:9] = np.clip(np.floor(a[:, 7:9]), 0, G - 1)
    a.view(FIELDS).sort(order=["f7", "f8"], axis=0, kind="stable")
    return a, (float(min_x), float(min_y), float(step))
