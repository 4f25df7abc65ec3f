"""ctypes binding of libbh_hip.so (the C ABI declared in include/bh_hip.h).

PyTorch is used only for device buffers and the current HIP stream.  There is
no CPU fallback: if the shared library is missing, or no GPU is present, the
calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                     "csrc")
LIB_PATH = os.path.join(_CSRC, "libbh_hip.so")

BH_OK, BH_EINVAL, BH_EHIP, BH_EDEVICE, BH_ESTATE = 0, 1, 2, 3, 4
(PREC_FP64, PREC_FP32_KAHAN, PREC_FP32_KAHAN_TERM, PREC_FP32_PLAIN, PREC_FP32_MFMA,
 PREC_FP64_IEEE) = range(6)
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32_KAHAN, "fp32_kahan": PREC_FP32_KAHAN,
              "fp32_kahan_term": PREC_FP32_KAHAN_TERM, "fp32_plain": PREC_FP32_PLAIN,
              "fp32_mfma": PREC_FP32_MFMA, "fp64_ieee": PREC_FP64_IEEE}
FP64_PRECISIONS = ("fp64", "fp64_ieee")
(EXPORT_COUNTS, EXPORT_OFFSETS, EXPORT_NEPREFIX, EXPORT_PERM, EXPORT_KEYS, EXPORT_COM,
 EXPORT_BBOX, EXPORT_ITEMS, EXPORT_COST, EXPORT_KEYS_IN) = range(10)
MLP_NPARAM = 1377
MLP_NFEAT = 8
ST_ERR, ST_MAXCOUNT, ST_NITEMS, ST_QUEUE, ST_NCELLS_NE = 0, 1, 2, 3, 4
ST_NRECOMP, ST_RCGROUPS = 6, 7

# every entry point declared in include/bh_hip.h
EXPORTED = (
    "bh_last_error", "bh_version", "bh_create", "bh_destroy", "bh_set_systems", "bh_set_tuning",
    "bh_get_tuning", "bh_set_cost_log", "bh_place_particles", "bh_calculate_box_cumm",
    "bh_sort_particles", "bh_summarize", "bh_build_items", "bh_evaluate_boxes",
    "bh_force_pass", "bh_tick_particles", "bh_step", "bh_step_graph", "bh_step_learned", "bh_read_stats", "bh_export",
    "bh_interactions", "bh_interactions_async", "bh_item_features", "bh_mlp_forward", "bh_mlp_backward", "bh_mlp_adam",
    "bh_order_by_cost", "bh_learned_order", "bh_query", "bh_energy", "bh_set_recompute",
    "bh_recompute", "bh_step_noalloc", "bh_energy_grid", "bh_force_pass_part",
    "bh_cloud_gravity",
)
# limits of the allocation-free step (bh_step_noalloc)
NOALLOC_MAX_N = 262144
NOALLOC_MAX_GRID = 256


class BhError(RuntimeError):
    """A libbh_hip call returned an error code."""

    def __init__(self, code, msg):
        super().__init__(f"libbh_hip error {code}: {msg}")
        self.code = code


class BhParticles(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("px", "py", "mass", "vx", "vy", "ax", "ay", "pid")]


class BhLaw(ctypes.Structure):
    _fields_ = [("G", ctypes.c_double), ("eps2", ctypes.c_double), ("r2_cut", ctypes.c_double)]


class BhCloud(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_void_p), ("n", ctypes.c_int64), ("ld", ctypes.c_int64)]


class BhTuning(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int32), ("yspan", ctypes.c_int32), ("waves_per_cu", ctypes.c_int32),
                ("tail_split", ctypes.c_int32), ("tail_items", ctypes.c_int32)]


_lib = None


def load(path: str = LIB_PATH):
    """Load libbh_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it first (python -c "
                           f"'import __graft_entry__ as g; g.build()' or make -C {_CSRC})")
    # torch ships its own libamdhip64.  Loaded first, it satisfies the
    # library's HIP dependency, so both share one runtime (and torch's device
    # pointers are valid in the library's calls).  Loaded after the library,
    # it would add a second runtime to the process, and the library's would no
    # longer see the device.
    _torch()
    L = ctypes.CDLL(path)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "bh_last_error": (ctypes.c_char_p, []),
        "bh_version": (ctypes.c_int, []),
        "bh_create": (ctypes.c_int, [ctypes.c_int, P(vp)]),
        "bh_destroy": (ctypes.c_int, [vp]),
        "bh_set_systems": (ctypes.c_int, [vp, i32, P(i64), P(i32)]),
        "bh_set_tuning": (ctypes.c_int, [vp, P(BhTuning)]),
        "bh_get_tuning": (ctypes.c_int, [vp, P(BhTuning)]),
        "bh_set_cost_log": (ctypes.c_int, [vp, i32]),
        "bh_place_particles": (ctypes.c_int, [vp, P(BhParticles), vp]),
        "bh_calculate_box_cumm": (ctypes.c_int, [vp, vp]),
        "bh_sort_particles": (ctypes.c_int, [vp, P(BhParticles), P(BhParticles), vp]),
        "bh_summarize": (ctypes.c_int, [vp, P(BhParticles), vp]),
        "bh_build_items": (ctypes.c_int, [vp, vp]),
        "bh_evaluate_boxes": (ctypes.c_int, [vp, P(BhParticles), P(BhLaw), i32, vp, vp]),
        "bh_force_pass": (ctypes.c_int, [vp, P(BhParticles), P(BhLaw), i32, vp, vp]),
        "bh_force_pass_part": (ctypes.c_int, [vp, P(BhParticles), P(BhLaw), i32, vp, i32, i32,
                                              vp]),
        "bh_tick_particles": (ctypes.c_int, [vp, P(BhParticles), dbl, vp]),
        "bh_step": (ctypes.c_int, [vp, P(BhParticles), P(BhParticles), P(BhLaw), i32, dbl, i32,
                                   vp]),
        "bh_step_graph": (ctypes.c_int, [vp, P(BhParticles), P(BhParticles), P(BhLaw), i32, dbl,
                                         i32, vp]),
        "bh_step_learned": (ctypes.c_int, [vp, P(BhParticles), P(BhParticles), P(BhLaw), i32,
                                           dbl, i32, vp, vp, i32, vp]),
        "bh_read_stats": (ctypes.c_int, [vp, P(i32), vp]),
        "bh_export": (ctypes.c_int, [vp, i32, vp, i64, vp]),
        "bh_interactions": (ctypes.c_int, [vp, P(i64), vp]),
        "bh_interactions_async": (ctypes.c_int, [vp, vp, vp]),
        "bh_item_features": (ctypes.c_int, [vp, vp, vp]),
        "bh_mlp_forward": (ctypes.c_int, [vp, vp, i64, vp, vp]),
        "bh_mlp_backward": (ctypes.c_int, [vp, vp, vp, i64, vp, vp, vp]),
        "bh_mlp_adam": (ctypes.c_int, [vp, vp, vp, vp, i32, ctypes.c_float, vp]),
        "bh_order_by_cost": (ctypes.c_int, [vp, vp, vp, vp]),
        "bh_learned_order": (ctypes.c_int, [vp, vp, vp, vp]),
        "bh_query": (ctypes.c_int, [vp, i32, P(i64)]),
        "bh_energy": (ctypes.c_int, [vp, P(BhParticles), dbl, dbl, P(dbl), vp]),
        "bh_energy_grid": (ctypes.c_int, [vp, P(BhParticles), dbl, dbl, P(dbl), vp]),
        "bh_set_recompute": (ctypes.c_int, [vp, dbl, i32]),
        "bh_recompute": (ctypes.c_int, [vp, P(BhParticles), P(BhLaw), vp]),
        "bh_step_noalloc": (ctypes.c_int, [P(BhParticles), P(BhParticles), i64, i32, P(BhLaw),
                                           i32, dbl, i32, P(BhTuning), dbl, vp, vp]),
        "bh_cloud_gravity": (ctypes.c_int, [P(BhCloud), P(BhCloud), i32, P(BhLaw), i32, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(code):
    if code != BH_OK:
        msg = _lib.bh_last_error().decode() if _lib is not None else "?"
        raise BhError(code, msg)


def _torch():
    import torch
    return torch


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def particles_struct(d, n=None):
    """BhParticles from a dict of device tensors px, py, mass, vx, vy, ax, ay,
    pid (each of n elements when n is given)."""
    for k in ("px", "py", "mass", "vx", "vy", "ax", "ay"):
        t = d[k]
        if t.dtype != _torch().float64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{k} must be a contiguous float64 device tensor")
        if n is not None and t.numel() != n:
            raise ValueError(f"{k} has {t.numel()} elements, the context holds {n} particles")
    t = d["pid"]
    if t.dtype != _torch().int32 or not t.is_cuda or not t.is_contiguous():
        raise ValueError("pid must be a contiguous int32 device tensor")
    if n is not None and t.numel() != n:
        raise ValueError(f"pid has {t.numel()} elements, the context holds {n} particles")
    return BhParticles(*[d[k].data_ptr() for k in ("px", "py", "mass", "vx", "vy", "ax", "ay",
                                                   "pid")])


def device_buffer(t, name, dtype, min_numel, exact=False, optional=False):
    """Checks a device tensor handed to the library as a raw pointer: dtype,
    contiguity, and at least (or exactly) min_numel elements."""
    if t is None:
        if optional:
            return
        raise ValueError(f"{name} is required")
    torch = _torch()
    want = getattr(torch, dtype) if isinstance(dtype, str) else dtype
    if t.dtype != want or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous {want} device tensor")
    n = t.numel()
    if (n != min_numel) if exact else (n < min_numel):
        raise ValueError(f"{name} has {n} elements, needs {'exactly' if exact else 'at least'} "
                         f"{min_numel}")


class Context:
    """Owns one bh_ctx (device scratch for a set of independent systems).
    Tensors passed as raw pointers are checked against the context's sizes
    (particles, work-item capacity) before any call."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = ctypes.c_void_p()
        check(self.L.bh_create(int(device), ctypes.byref(h)))
        self.h = h
        self.device = device
        self.n_sys = 0
        self.N = None
        self.max_items = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.bh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- setup
    def set_systems(self, n, grid_dim):
        n = np.ascontiguousarray(np.atleast_1d(n), dtype=np.int64)
        g = np.ascontiguousarray(np.atleast_1d(grid_dim), dtype=np.int32)
        if n.shape != g.shape:
            raise ValueError("n and grid_dim must have the same length")
        check(self.L.bh_set_systems(self.h, len(n), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    g.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        self.n_sys = len(n)
        self.n = n.copy()
        self.grid_dim = g.copy()
        self.N = int(n.sum())
        self.max_items = self.query(4)

    def _parts(self, d):
        return particles_struct(d, self.N)

    def _order(self, order, optional=True):
        device_buffer(order, "order", "int32", self.max_items, optional=optional)

    def _params(self, params):
        device_buffer(params, "params", "float32", MLP_NPARAM, exact=True)

    def set_tuning(self, p=2, yspan=512, waves_per_cu=32, tail_split=16, tail_items=1 << 20):
        t = BhTuning(int(p), int(yspan), int(waves_per_cu), int(tail_split), int(tail_items))
        check(self.L.bh_set_tuning(self.h, ctypes.byref(t)))

    def get_tuning(self):
        t = BhTuning()
        check(self.L.bh_get_tuning(self.h, ctypes.byref(t)))
        return dict(p=t.p, yspan=t.yspan, waves_per_cu=t.waves_per_cu, tail_split=t.tail_split,
                    tail_items=t.tail_items)

    def set_cost_log(self, on: bool):
        check(self.L.bh_set_cost_log(self.h, 1 if on else 0))

    def set_recompute(self, kappa=1.5, deferred=False):
        """FP64 recompute after FP32 passes of the particles whose rounding-
        error bound sqrt(sum |term|^2) exceeds kappa * |a| (kappa <= 0: off).
        deferred: force_pass stops after the FP32 kernels and the caller runs
        recompute() (the step calls still run it)."""
        check(self.L.bh_set_recompute(self.h, float(kappa), 1 if deferred else 0))

    def recompute(self, parts, G=1.0, eps2=0.0, r2_cut=0.0, stream=None):
        p = self._parts(parts)
        law = BhLaw(float(G), float(eps2), float(r2_cut))
        check(self.L.bh_recompute(self.h, ctypes.byref(p), ctypes.byref(law),
                                  stream_handle(stream)))

    def query(self, what: int) -> int:
        out = ctypes.c_int64()
        check(self.L.bh_query(self.h, int(what), ctypes.byref(out)))
        return int(out.value)

    # ---- phases
    def place(self, parts, stream=None):
        p = self._parts(parts)
        check(self.L.bh_place_particles(self.h, ctypes.byref(p), stream_handle(stream)))

    def cumm(self, stream=None):
        check(self.L.bh_calculate_box_cumm(self.h, stream_handle(stream)))

    def sort(self, src, dst, stream=None):
        a, b = self._parts(src), self._parts(dst)
        check(self.L.bh_sort_particles(self.h, ctypes.byref(a), ctypes.byref(b),
                                       stream_handle(stream)))

    def summarize(self, parts, stream=None):
        p = self._parts(parts)
        check(self.L.bh_summarize(self.h, ctypes.byref(p), stream_handle(stream)))

    def build_items(self, stream=None):
        check(self.L.bh_build_items(self.h, stream_handle(stream)))

    def evaluate(self, parts, G=1.0, eps2=0.0, r2_cut=0.0, precision="fp64", order=None,
                 stream=None):
        p = self._parts(parts)
        self._order(order)
        law = BhLaw(float(G), float(eps2), float(r2_cut))
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        check(self.L.bh_evaluate_boxes(self.h, ctypes.byref(p), ctypes.byref(law), prec,
                                       ptr(order), stream_handle(stream)))

    def force_pass(self, parts, G=1.0, eps2=0.0, r2_cut=0.0, precision="fp64", order=None,
                   stream=None, part=None, nparts=1):
        """The force kernels on the current items.  part / nparts: one of
        nparts launches covering contiguous ranges of the work queue (call
        part = 0 .. nparts-1 in order; the result equals one launch)."""
        p = self._parts(parts)
        self._order(order)
        law = BhLaw(float(G), float(eps2), float(r2_cut))
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        if part is None:
            check(self.L.bh_force_pass(self.h, ctypes.byref(p), ctypes.byref(law), prec,
                                       ptr(order), stream_handle(stream)))
        else:
            check(self.L.bh_force_pass_part(self.h, ctypes.byref(p), ctypes.byref(law), prec,
                                            ptr(order), int(part), int(nparts),
                                            stream_handle(stream)))

    def tick(self, parts, dt, stream=None):
        p = self._parts(parts)
        check(self.L.bh_tick_particles(self.h, ctypes.byref(p), float(dt), stream_handle(stream)))

    def step(self, a, b, G=1.0, eps2=0.0, r2_cut=0.0, precision="fp64", dt=0.0, do_tick=True,
             graph=False, stream=None):
        pa, pb = self._parts(a), self._parts(b)
        law = BhLaw(float(G), float(eps2), float(r2_cut))
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        fn = self.L.bh_step_graph if graph else self.L.bh_step
        check(fn(self.h, ctypes.byref(pa), ctypes.byref(pb), ctypes.byref(law), prec, float(dt),
                 1 if do_tick else 0, stream_handle(stream)))

    def step_learned(self, a, b, params, order, G=1.0, eps2=0.0, r2_cut=0.0, precision="fp32",
                     dt=0.0, do_tick=True, graph=False, stream=None):
        pa, pb = self._parts(a), self._parts(b)
        self._params(params)
        self._order(order, optional=False)
        law = BhLaw(float(G), float(eps2), float(r2_cut))
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        check(self.L.bh_step_learned(self.h, ctypes.byref(pa), ctypes.byref(pb), ctypes.byref(law),
                                     prec, float(dt), 1 if do_tick else 0, ptr(params), ptr(order),
                                     1 if graph else 0, stream_handle(stream)))

    # ---- state
    def stats(self, stream=None, raise_on_error=True):
        out = np.zeros(16, np.int32)
        code = self.L.bh_read_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                    stream_handle(stream))
        if code == BH_EDEVICE and not raise_on_error:
            return out
        check(code)
        return out

    def export(self, what, count, dtype, stream=None):
        torch = _torch()
        out = torch.empty(int(count), dtype=dtype, device=f"cuda:{self.device}")
        check(self.L.bh_export(self.h, int(what), ptr(out), int(count), stream_handle(stream)))
        return out

    def export_host(self, what, count, dtype):
        out = np.empty(int(count), dtype=dtype)
        check(self.L.bh_export(self.h, int(what), ctypes.c_void_p(out.ctypes.data), int(count),
                               stream_handle()))
        return out

    def interactions(self, stream=None):
        out = np.zeros(2 * self.n_sys + 2, np.int64)
        check(self.L.bh_interactions(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     stream_handle(stream)))
        return out

    def energy(self, parts, G=1.0, eps2=0.0, stream=None):
        """[(U_s, K_s)] per system: direct FP64 potential of the law and kinetic energy."""
        p = self._parts(parts)
        out = np.zeros(2 * self.n_sys, np.float64)
        check(self.L.bh_energy(self.h, ctypes.byref(p), float(G), float(eps2),
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               stream_handle(stream)))
        return out.reshape(-1, 2)

    def energy_grid(self, parts, G=1.0, eps2=0.0, stream=None):
        """[(U_s, K_s)] per system with the potential by the grid method (near
        cells direct, far cells from their COMs; needs the current placement
        and summary of parts)."""
        p = self._parts(parts)
        out = np.zeros(2 * self.n_sys, np.float64)
        check(self.L.bh_energy_grid(self.h, ctypes.byref(p), float(G), float(eps2),
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    stream_handle(stream)))
        return out.reshape(-1, 2)

    def interactions_async(self, dev_out, stream=None):
        """dev_out: device int64 tensor [2*S+2] (near counts include self pairs)."""
        device_buffer(dev_out, "dev_out", "int64", 2 * self.n_sys + 2)
        check(self.L.bh_interactions_async(self.h, ptr(dev_out), stream_handle(stream)))

    # ---- cost model
    def item_features(self, feat, stream=None):
        """feat: device float32 of at least 8 x the work-item capacity."""
        device_buffer(feat, "feat", "float32", MLP_NFEAT * self.max_items)
        check(self.L.bh_item_features(self.h, ptr(feat), stream_handle(stream)))

    def order_by_cost(self, cost, order, stream=None):
        device_buffer(cost, "cost", "float32", self.max_items)
        self._order(order, optional=False)
        check(self.L.bh_order_by_cost(self.h, ptr(cost), ptr(order), stream_handle(stream)))

    def learned_order(self, params, order, stream=None):
        self._params(params)
        self._order(order, optional=False)
        check(self.L.bh_learned_order(self.h, ptr(params), ptr(order), stream_handle(stream)))


def step_noalloc(a, b, grid_dim, G=1.0, eps2=0.0, r2_cut=0.0, precision="fp64", dt=0.0,
                 do_tick=True, tuning=None, kappa=1.5, stats=None, stream=None):
    """bh_step_noalloc: one full round of one system (a placed and sorted
    into b, then summarize, forces, tick on b) with no device allocation; all
    scratch comes from the library's static per-GPU array.  tuning: dict of
    bh_tuning fields (tail_split must be 1) or None; stats: device int32[16]
    or None."""
    n = a["px"].numel()
    pa, pb = particles_struct(a, n), particles_struct(b, n)
    device_buffer(stats, "stats", "int32", 16, optional=True)
    law = BhLaw(float(G), float(eps2), float(r2_cut))
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    tun = None
    if tuning is not None:
        tun = ctypes.byref(BhTuning(int(tuning["p"]), int(tuning["yspan"]),
                                    int(tuning["waves_per_cu"]), int(tuning["tail_split"]),
                                    int(tuning["tail_items"])))
    L = load()
    check(L.bh_step_noalloc(ctypes.byref(pa), ctypes.byref(pb), int(n), int(grid_dim),
                            ctypes.byref(law), prec, float(dt), 1 if do_tick else 0, tun,
                            float(kappa), ptr(stats), stream_handle(stream)))


def cloud_struct(rows, name):
    """BhCloud over a (n, ld >= 7) float64 device tensor of particle rows
    (columns px, py, mass, vx, vy, ax, ay, ...) with unit column stride."""
    torch = _torch()
    if not isinstance(rows, torch.Tensor) or rows.dtype != torch.float64 or not rows.is_cuda:
        raise ValueError(f"{name} must be a float64 device tensor")
    if rows.dim() != 2 or rows.shape[1] < 7:
        raise ValueError(f"{name} must have shape (n, >= 7), got {tuple(rows.shape)}")
    n = rows.shape[0]
    if n and (rows.stride(1) != 1 or (n > 1 and rows.stride(0) < rows.shape[1])):
        raise ValueError(f"{name} needs unit column stride and non-overlapping rows")
    ld = rows.stride(0) if n > 1 else rows.shape[1]
    return BhCloud(rows.data_ptr() if n else None, n, ld)


def cloud_gravity(self_rows, other_rows=None, G=1.0, eps2=0.0, r2_cut=1e-5,
                  update_other=False, stream=None):
    """bh_cloud_gravity on device row tensors, in place: adds to self_rows'
    ax, ay columns the attraction of other_rows (other_rows None: of the other
    rows of self_rows), and with update_other also to other_rows'.  Default law:
    the reference's CPU kernel (r^2 <= 1e-5 dropped, no softening)."""
    same = other_rows is None
    a = cloud_struct(self_rows, "self_rows")
    b = a if same else cloud_struct(other_rows, "other_rows")
    if not same and self_rows.device != other_rows.device:
        raise ValueError("self_rows and other_rows must be on one device")
    law = BhLaw(float(G), float(eps2), float(r2_cut))
    L = load()
    check(L.bh_cloud_gravity(ctypes.byref(a), None if same else ctypes.byref(b), int(same),
                             ctypes.byref(law), 1 if update_other else 0,
                             stream_handle(stream)))


def _mlp_rows(x, name="x"):
    if x.dim() != 2 or x.shape[1] != MLP_NFEAT:
        raise ValueError(f"{name} must be [n, {MLP_NFEAT}]")
    device_buffer(x, name, "float32", x.shape[0] * MLP_NFEAT, exact=True)
    return int(x.shape[0])


def mlp_forward(params, x, y, stream=None):
    device_buffer(params, "params", "float32", MLP_NPARAM, exact=True)
    n = _mlp_rows(x)
    device_buffer(y, "y", "float32", n)
    L = load()
    check(L.bh_mlp_forward(ptr(params), ptr(x), int(x.shape[0]), ptr(y), stream_handle(stream)))


def mlp_backward(params, x, target, grad, loss=None, stream=None):
    device_buffer(params, "params", "float32", MLP_NPARAM, exact=True)
    n = _mlp_rows(x)
    device_buffer(target, "target", "float32", n)
    device_buffer(grad, "grad", "float32", MLP_NPARAM, exact=True)
    device_buffer(loss, "loss", "float32", 1, optional=True)
    L = load()
    check(L.bh_mlp_backward(ptr(params), ptr(x), ptr(target), int(x.shape[0]), ptr(grad),
                            ptr(loss), stream_handle(stream)))


def mlp_adam(params, grad, m, v, t, lr, stream=None):
    for name, b in (("params", params), ("grad", grad), ("m", m), ("v", v)):
        device_buffer(b, name, "float32", MLP_NPARAM, exact=True)
    L = load()
    check(L.bh_mlp_adam(ptr(params), ptr(grad), ptr(m), ptr(v), int(t), float(lr),
                        stream_handle(stream)))
