"""The raw ctypes binding a maintainer would add on the reference side (the
stub shown for the C ABI), exercised as written: no helper module, plain
pointers and a stream handle.  One FP64 round through the per-kernel entry
points and one through the captured whole-step graph, against the oracle, and
an FP32 pass with the FP64 recompute run inside the pass or by the caller."""
import ctypes
import os

import numpy as np
import pytest

from conftest import PKG, gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

from oracle import oracle as orc  # noqa: E402
from barneshut import inputs  # noqa: E402

LIB = os.path.join(PKG, "csrc", "libbh_hip.so")


class Particles(ctypes.Structure):          # bh_particles
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("px", "py", "mass", "vx", "vy", "ax", "ay", "pid")]


class Law(ctypes.Structure):                # bh_law
    _fields_ = [("G", ctypes.c_double), ("eps2", ctypes.c_double), ("r2_cut", ctypes.c_double)]


def _buffers(torch, d, n, dev):
    t = {k: torch.as_tensor(np.asarray(d[k], np.float64)).to(dev) for k in
         ("px", "py", "mass", "vx", "vy")}
    t["ax"] = torch.zeros(n, dtype=torch.float64, device=dev)
    t["ay"] = torch.zeros(n, dtype=torch.float64, device=dev)
    t["pid"] = torch.as_tensor(np.asarray(d["pid"], np.int32)).to(dev)
    return t


def _struct(t):
    return Particles(*[t[k].data_ptr() for k in ("px", "py", "mass", "vx", "vy", "ax", "ay", "pid")])


def test_stub_binding_phases_and_graph():
    import torch
    lib = ctypes.CDLL(LIB)
    for name in ("bh_create", "bh_set_systems", "bh_place_particles", "bh_calculate_box_cumm",
                 "bh_sort_particles", "bh_summarize", "bh_evaluate_boxes", "bh_tick_particles",
                 "bh_step_graph", "bh_destroy"):
        getattr(lib, name).restype = ctypes.c_int
    lib.bh_tick_particles.argtypes = [ctypes.c_void_p, ctypes.POINTER(Particles), ctypes.c_double,
                                      ctypes.c_void_p]
    lib.bh_step_graph.argtypes = [ctypes.c_void_p, ctypes.POINTER(Particles),
                                  ctypes.POINTER(Particles), ctypes.POINTER(Law), ctypes.c_int32,
                                  ctypes.c_double, ctypes.c_int32, ctypes.c_void_p]
    N, G = 6000, 24
    d = inputs.gaussian_mixture(N, 5, (0.02, 0.1), seed=44)
    dev = torch.device("cuda:0")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    law = Law(1.0, 1e-6, 0.0)

    ctx = ctypes.c_void_p()
    assert lib.bh_create(0, ctypes.byref(ctx)) == 0
    n = (ctypes.c_int64 * 1)(N)
    g = (ctypes.c_int32 * 1)(G)
    assert lib.bh_set_systems(ctx, 1, n, g) == 0

    # per-kernel calls: a = current buffers, b = the other set
    ta, tb = _buffers(torch, d, N, dev), _buffers(torch, d, N, dev)
    a, b = _struct(ta), _struct(tb)
    assert lib.bh_place_particles(ctx, ctypes.byref(a), stream) == 0
    assert lib.bh_calculate_box_cumm(ctx, stream) == 0
    assert lib.bh_sort_particles(ctx, ctypes.byref(a), ctypes.byref(b), stream) == 0
    assert lib.bh_summarize(ctx, ctypes.byref(b), stream) == 0
    assert lib.bh_evaluate_boxes(ctx, ctypes.byref(b), ctypes.byref(law), 0, None, stream) == 0
    torch.cuda.synchronize()
    h = {k: tb[k].cpu().numpy() for k in tb}
    ref = orc.step(d["px"], d["py"], d["mass"], d["vx"], d["vy"], d["pid"], G, 1.0, 0.0, 1e-6,
                   0.0, do_tick=False)
    np.testing.assert_array_equal(h["pid"], ref["pid"])
    err = np.hypot(h["ax"] - ref["ax"], h["ay"] - ref["ay"]) / np.hypot(ref["ax"], ref["ay"])
    assert err.max() < 1e-9
    # FP32 pass with the FP64 recompute run by the caller (deferred) equals the
    # same pass with the recompute inside bh_force_pass
    lib.bh_set_recompute.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]
    lib.bh_set_recompute.restype = ctypes.c_int
    lib.bh_recompute.argtypes = [ctypes.c_void_p, ctypes.POINTER(Particles), ctypes.POINTER(Law),
                                 ctypes.c_void_p]
    lib.bh_recompute.restype = ctypes.c_int
    lib.bh_read_stats.restype = ctypes.c_int
    fp32 = []
    for deferred in (0, 1):
        tb["ax"].zero_()
        tb["ay"].zero_()
        assert lib.bh_set_recompute(ctx, ctypes.c_double(2.0), deferred) == 0
        assert lib.bh_force_pass(ctx, ctypes.byref(b), ctypes.byref(law), 1, None, stream) == 0
        if deferred:
            assert lib.bh_recompute(ctx, ctypes.byref(b), ctypes.byref(law), stream) == 0
        stats = (ctypes.c_int32 * 16)()
        assert lib.bh_read_stats(ctx, stats, stream) == 0
        fp32.append((tb["ax"].cpu().numpy(), tb["ay"].cpu().numpy(), stats[6]))
    np.testing.assert_array_equal(fp32[0][0], fp32[1][0])
    np.testing.assert_array_equal(fp32[0][1], fp32[1][1])
    assert fp32[0][2] == fp32[1][2]
    err = np.hypot(fp32[0][0] - ref["ax"], fp32[0][1] - ref["ay"]) / np.hypot(ref["ax"], ref["ay"])
    assert err.max() < 1e-5
    # without a prepared FP32 pass the recompute is refused (BH_ESTATE)
    assert lib.bh_set_recompute(ctx, ctypes.c_double(2.0), 1) == 0
    assert lib.bh_recompute(ctx, ctypes.byref(b), ctypes.byref(law), stream) == 4
    assert lib.bh_set_recompute(ctx, ctypes.c_double(2.0), 0) == 0

    # back to the FP64 accelerations for the tick below
    tb["ax"].zero_()
    tb["ay"].zero_()
    assert lib.bh_force_pass(ctx, ctypes.byref(b), ctypes.byref(law), 0, None, stream) == 0
    assert lib.bh_tick_particles(ctx, ctypes.byref(b), ctypes.c_double(1e-4), stream) == 0
    torch.cuda.synchronize()
    after_phases = {k: tb[k].cpu().numpy() for k in ("px", "py", "vx", "vy", "pid")}

    # the whole round from a captured graph, on fresh buffers: same result
    tc, td = _buffers(torch, d, N, dev), _buffers(torch, d, N, dev)
    c_, d_ = _struct(tc), _struct(td)
    assert lib.bh_step_graph(ctx, ctypes.byref(c_), ctypes.byref(d_), ctypes.byref(law), 0,
                             ctypes.c_double(1e-4), 1, stream) == 0
    torch.cuda.synchronize()
    for k in ("px", "py", "vx", "vy", "pid"):
        np.testing.assert_array_equal(td[k].cpu().numpy(), after_phases[k])
    assert lib.bh_destroy(ctx) == 0
