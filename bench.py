#!/usr/bin/env python3
"""Throughput of the grid-method step on one MI355X.

Workload (`value`): N = 4,194,304 particles in 2-D, a mixture of 64 isotropic
Gaussians (centres U[0,1)^2, sigma log-uniform [0.005, 0.05], clipped to the
unit box), masses U[1,2], zero initial velocity, a 1024 x 1024 grid; the reference's
law a_i = G sum m_j (x_j - x_i) / (r^2 + eps^2) with softening eps = 1e-5 (a
third of the mean particle spacing in the densest cluster), no cutoff,
G = 1e-6, dt = 1e-5.

The timed window is always steps 1..K of that initial distribution: the
warm-up steps run first, then the initial state is loaded again.  The run is
resolved: the softening caps any pair's acceleration at G m / (2 eps) = 0.1,
whose orbital time (~1e-2) is a thousand steps, and over 25 steps (2.5e-4
time units, < 1 % of the clusters' free-fall time of ~0.03) particles move
~1e-7 of the box, so every step places the same distribution and does the
same interactions (1.99e12 per step), whatever --steps and --warmup are.
Softening costs nothing: the kernels add eps^2 to r^2 either way.

A "step" is one full round: bounding box + placement + ordered permutation,
per-cell COM, forces, tick.  Paths on the same input:
  * plain FP64, static work-queue order -- the rebuilt reference path, each
    weight m / r^2 from the hardware reciprocal plus one third-order
    refinement (within about one ulp of the quotient);
  * the same with the IEEE division per weight (`fp64_ieee`, secondary);
  * FP32 cell-relative forces, Kahan-compensated tile sums, static order;
  * the same FP32 path with the learned work-queue order (the cost model in
    barnes-hut_amd/costmodels/clusters256.npz; --cost-model none disables).
The FP32 paths include the FP64 recompute of the particles whose FP32 terms
nearly cancel (kappa 1.5, --recompute-kappa 0 disables); it is timed apart from
the force kernel.  `value` is the FP32 + learned-order path's whole-job
throughput.

Secondary blocks (not in `value`): the cost-model workload (256 clusters,
sigma [0.001, 0.03], lognormal masses, 1024^2: FP64, FP32 static, FP32
learned), 512 independent systems stepped by one graph launch per step, the
reference's CPU-runnable case (16384 uniform particles, 32^2, eps = 1e-3: the
oracle's full step on the host cores and the GPU step), and the reference's
own logged configuration (N = 10M uniform in [0,499)^2, 120 x 120 grid, FP64,
G = 1, dt = 0.005) against its published 166.837 s/step, and the gravity
plugin's cloud-to-cloud kernel (one cloud of 131,072 particles; a 64-particle
leaf against 16,384 COMs, on device and on host rows).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "barnes-hut_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Mparticle-steps/s at N=4M 2D (FP32+learned sched vs plain); force rel err"
N_DEFAULT = 4194304
GRID_DEFAULT = 1024
N_CLUSTERS = 64
SIGMA = (0.005, 0.05)
SEED = 2024
GC, DT, EPS2, R2CUT = 1e-6, 1e-5, 1e-10, 0.0
# cost-model workload: 256 clusters, sigma log-uniform [0.001, 0.03], lognormal masses
C4_CLUSTERS, C4_SIGMA, C4_SEED = 256, (0.001, 0.03), 91
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
FLOP_PER_INTERACTION = 11
# HBM bytes per FP32 force-kernel launch on this workload: rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (KiB per dispatch), scaled by the readings of a
# 1 GiB device copy (FETCH x2.0, WRITE x1.0), measured on the first two
# steps of this workload with the default launch shape
# (profiles/force32_gauss64_n4M_g1024_hbm_traffic.md).
TRAFFIC = 5.09e9
# Issue-limited ceiling of the FP32 far-field term: 6 VALU instructions per
# interaction (v_rcp_f32 co-issued), no packed FP32, 256 CUs x 4 SIMDs x 16
# lanes x 2.4 GHz = 39.3e12 lane-instructions/s -> 6.55e12 interactions/s.
LANE_OPS = 256 * 4 * 16 * 2.4e9
ISSUE_CEILING_TFLOPS = LANE_OPS / 6 * FLOP_PER_INTERACTION / 1e12
# FP64 far-field term, two entries per reciprocal: 10.5 VALU instructions
# (2 differences, 2 fma for r^2, half of Q = q0 q1 and of its 3-instruction
# refinement, 2 products for the weight, 2 fma into the sums) + half a
# v_rcp_f64, which occupies 4 issue slots: 12.5 slots per term
# (profiles/force64_gauss64_n4M_g1024_valu_issue.md)
ISSUE_CEILING64_TFLOPS = LANE_OPS / 12.5 * FLOP_PER_INTERACTION / 1e12
# FP64 recompute threshold of the FP32 paths (--recompute-kappa)
RECOMPUTE_KAPPA = 1.5


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_path(eng, d, precision, steps, warmup, order_fn=None, kappa=RECOMPUTE_KAPPA):
    """Warm-up from the initial state `d`, then `d` again and `steps` timed
    full steps.  Returns (seconds per step, force-kernel ms per step,
    recompute ms per step, near and far interactions per step).  The FP32
    force pass runs deferred, so that HIP events split the force kernels (with
    the slice combine) from the FP64 recompute stage."""
    import torch
    ctx = eng.ctx
    fp32 = not precision.startswith("fp64")
    if fp32:
        ctx.set_recompute(kappa, deferred=True)

    n_tot = int(eng.n.sum())
    counts = torch.zeros((steps, 4), dtype=torch.int64, device=eng.dev)

    def one(events=None, slot=None):
        eng.create_tree()
        if slot is not None:
            ctx.interactions_async(counts[slot])
        eng.summarize()
        ctx.build_items()
        order = order_fn() if order_fn is not None else None
        if events is not None:
            events[0].record()
        ctx.force_pass(eng.parts, GC, EPS2, R2CUT, precision, order)
        if events is not None:
            events[1].record()
        if fp32 and kappa > 0:
            ctx.recompute(eng.parts, GC, EPS2, R2CUT)
        if events is not None:
            events[2].record()
        eng.timestep(DT)

    for _ in range(warmup):
        one()
    eng.check()
    eng.load(d)   # the timed window starts from the initial state
    torch.cuda.synchronize()
    evs = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(steps)]
    t0 = time.perf_counter()
    for s in range(steps):
        one(evs[s], s)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    eng.check()
    kms = np.array([a.elapsed_time(b) for a, b, _ in evs])
    rms = np.array([b.elapsed_time(c) for _, b, c in evs])
    c = counts.cpu().numpy()
    near = c[:, 2] - n_tot
    far = c[:, 3]
    return (t1 - t0) / steps, kms, rms, near, far


def measure(d, G, precision, tuning, steps, warmup, model=None, kappa=RECOMPUTE_KAPPA,
            autotune_first=False):
    """One path on workload `d` (G x G grid): see run_path."""
    import torch
    from barneshut.engine import Engine
    n = len(d["px"])
    eng = Engine([n], [G], tuning=tuning)
    eng.load(d)
    if autotune_first:
        from barneshut import autotune
        eng.create_tree()
        eng.summarize()
        tuning = autotune.tune(eng, (GC, EPS2, R2CUT), precision, log=log)["tuning"]
        eng.load(d)
    order_fn = None
    if model is not None:
        order_buf = torch.empty(eng.ctx.query(4), dtype=torch.int32, device=eng.dev)

        def order_fn():
            eng.ctx.learned_order(model.params, order_buf)
            return order_buf
    sec, kms, rms, near, far = run_path(eng, d, precision, steps, warmup, order_fn, kappa)
    st = eng.ctx.stats()
    n_items, n_rc = int(st[2]), int(st[6])
    eng.close()
    flops = FLOP_PER_INTERACTION * (near + far).astype(np.float64)
    return dict(seconds_per_step=sec, value=n / sec / 1e6, kms=kms,
                force_kernel_ms=float(kms.mean()), recompute_ms=float(rms.mean()),
                recomputed_last_step=n_rc, near=near, far=far,
                flop_per_launch=float(flops.mean()),
                achieved=float(flops.sum() / (kms.sum() * 1e-3) / 1e12), tuning=tuning,
                work_items=n_items, steps=steps)


def path_summary(r, peak):
    return {"value": r["value"], "ms_per_step": r["seconds_per_step"] * 1e3,
            "steps": r["steps"], "tuning": r["tuning"], "force_kernel_ms": r["force_kernel_ms"],
            "recompute_ms": r["recompute_ms"], "achieved_tflops": r["achieved"],
            "frac_of_peak": r["achieved"] / peak,
            "interactions_first_last": [float(r["near"][0] + r["far"][0]),
                                        float(r["near"][-1] + r["far"][-1])]}


def force_errors(d, G, runs, kappa=RECOMPUTE_KAPPA):
    """Accelerations of the initial state (the first timed step's): FP32
    precisions and the IEEE-division FP64 kernel against the default FP64
    kernel on the same placement, normwise, p99 and max per-particle relative
    error over all N particles.  The FP64 path itself is checked against the
    oracle by the tests, on these workloads at full size."""
    import torch
    from barneshut.engine import Engine
    eng = Engine([len(d["px"])], [G])
    eng.load(d)
    eng.create_tree()
    eng.summarize()
    eng.ctx.build_items()
    eng.ctx.force_pass(eng.parts, GC, EPS2, R2CUT, "fp64")
    a64 = (eng.parts["ax"].clone(), eng.parts["ay"].clone())
    out = {}
    for name, prec, kap in runs:
        eng.ctx.set_recompute(kap)
        eng.parts["ax"].zero_()
        eng.parts["ay"].zero_()
        eng.ctx.force_pass(eng.parts, GC, EPS2, R2CUT, prec)
        num = torch.hypot(eng.parts["ax"] - a64[0], eng.parts["ay"] - a64[1])
        den = torch.hypot(a64[0], a64[1])
        per = (num / den).cpu().numpy()
        out[name] = dict(normwise=float(torch.sqrt((num ** 2).sum() / (den ** 2).sum())),
                         p99=float(np.quantile(per, 0.99)), max=float(per.max()),
                         n_over_1e5=int((per > 1e-5).sum()))
        if kap > 0 and not prec.startswith("fp64"):
            out[name]["recomputed"] = int(eng.ctx.stats()[6])
    eng.check()
    eng.close()
    return out


C2_ERROR_RUNS = [("fp32", "fp32", RECOMPUTE_KAPPA), ("fp32_no_recompute", "fp32", 0.0),
                 ("fp32_kahan_term", "fp32_kahan_term", RECOMPUTE_KAPPA),
                 ("fp32_kahan_term_no_recompute", "fp32_kahan_term", 0.0),
                 ("fp32_plain_no_recompute", "fp32_plain", 0.0),
                 ("fp64_ieee", "fp64_ieee", 0.0)]


def cpu_baseline(d, G, budget_s=12.0):
    """The C restatement (oracle/) of python-bh's sequential step on this host:
    full placement, COM and tick over all N; forces on a random sample of
    target particles (OpenMP over targets), extrapolated to N.  The GPU
    placement of the same state is then checked against the timed oracle
    placement (bounding box, cell keys, counts, offsets, permutation)."""
    from barneshut.engine import Engine
    from oracle import oracle as orc
    cores = len(os.sched_getaffinity(0))
    n = len(d["px"])
    z = np.zeros(n)
    t0 = time.perf_counter()
    bb = orc.bbox(d["px"], d["py"], G)
    _, _, key = orc.place(d["px"], d["py"], bb[0], bb[1], bb[4], G)
    perm = orc.sort_perm(key, (d["px"], d["py"], d["mass"], d["vx"], d["vy"], z, z,
                               d["pid"].astype(np.float64)))
    ks = key[perm]
    x, y, m = d["px"][perm], d["py"][perm], d["mass"][perm]
    _, off = orc.counts_offsets(ks, G * G)
    t1 = time.perf_counter()
    com = orc.summarize(off, x, y, m)
    t2 = time.perf_counter()
    rng = np.random.default_rng(11)
    probe = np.sort(rng.choice(n, 256, replace=False))
    ta = time.perf_counter()
    orc.forces(G, off, ks, x, y, m, com, GC, EPS2, R2CUT, targets=probe, nthreads=cores)
    per_target = (time.perf_counter() - ta) / len(probe)
    ns = int(min(n, max(512, budget_s / max(per_target, 1e-9))))
    tg = np.sort(rng.choice(n, ns, replace=False))
    ta = time.perf_counter()
    orc.forces(G, off, ks, x, y, m, com, GC, EPS2, R2CUT, targets=tg, nthreads=cores)
    t_force = (time.perf_counter() - ta) * n / ns
    vx, vy = d["vx"][perm].copy(), d["vy"][perm].copy()
    ax, ay = np.zeros(n), np.zeros(n)
    xx, yy = x.copy(), y.copy()
    tb = time.perf_counter()
    orc.tick(xx, yy, vx, vy, ax, ay, DT)
    t_tick = time.perf_counter() - tb
    t_step = (t1 - t0) + (t2 - t1) + t_force + t_tick
    # after the timing: the oracle's placement as the checker
    cnt, _ = orc.counts_offsets(ks, G * G)
    eng = Engine([n], [G])
    eng.load(d)
    eng.create_tree()
    eng.check()
    gbb = eng.bbox()[0]
    check = {"n": n,
             "bbox": bool(gbb[0] == bb[0] and gbb[1] == bb[1] and gbb[2] == bb[2]
                          and gbb[4] == bb[4]),
             "keys": bool(np.array_equal(eng.keys_in(), key)),
             "counts": bool(np.array_equal(eng.counts(), cnt)),
             "offsets": bool(np.array_equal(eng.offsets(), off)),
             "permutation": bool(np.array_equal(eng.perm(), perm))}
    eng.close()
    return {"value": n / t_step / 1e6, "unit": "Mparticle-steps/s", "cores": cores,
            "kind": "port",
            "sample": (f"C restatement of the sequential step (oracle/bh_oracle.c) on the initial "
                       f"state (the first timed step's): placement ({t1 - t0:.2f}s, 1 thread), "
                       f"COM ({t2 - t1:.2f}s) and tick over all {n} particles; forces on {ns} "
                       f"random particles with {cores} OpenMP threads, x{n / ns:.1f} "
                       f"extrapolated ({t_force:.1f}s/step)"),
            "seconds_per_step": t_step, "placement_check": check}


def c1_line(steps):
    """The reference's CPU-runnable case: 16384 particles U[0,1)^2, masses
    U[1,2], zero velocity, G = 1, eps = 1e-3 (eps2 = 1e-6), 32 x 32 grid,
    dt = 1e-4, one step.  The oracle's full step (placement, COM, forces of
    every particle, tick) on the host cores, the GPU FP64 step, and the grid
    method's own error against the FP64 direct sum (both sides)."""
    from barneshut import inputs
    from barneshut.engine import Engine
    from oracle import oracle as orc
    n, G, gc, eps2, dt = 16384, 32, 1.0, 1e-6, 1e-4
    d = inputs.uniform(n, seed=1)
    cores = len(os.sched_getaffinity(0))
    t = []
    for _ in range(3):
        t0 = time.perf_counter()
        ref = orc.step(d["px"], d["py"], d["mass"], d["vx"], d["vy"], d["pid"], G, gc, dt, eps2,
                       0.0, nthreads=cores)
        t.append(time.perf_counter() - t0)
    t_cpu = min(t)
    t1 = time.perf_counter()
    ref1 = orc.step(d["px"], d["py"], d["mass"], d["vx"], d["vy"], d["pid"], G, gc, dt, eps2,
                    0.0, nthreads=1)
    t_cpu1 = time.perf_counter() - t1
    ex, ey = orc.direct(ref["px"], ref["py"], ref["m"], gc, eps2, 0.0, nthreads=cores)
    e_orc = float(np.sqrt(((ref["ax"] - ex) ** 2 + (ref["ay"] - ey) ** 2).sum()
                          / (ex ** 2 + ey ** 2).sum()))
    import torch
    eng = Engine([n], [G], tuning=dict(p=1, yspan=512, waves_per_cu=32, tail_split=1,
                                       tail_items=0))
    eng.load(d)
    eng.step(gc, eps2, 0.0, "fp64", dt, do_tick=False)
    eng.check()
    g = eng.fetch()
    e_gpu = float(np.sqrt(((g["ax"] - ex) ** 2 + (g["ay"] - ey) ** 2).sum()
                          / (ex ** 2 + ey ** 2).sum()))
    t_gpu = {}
    for graph in (False, True):   # kernel by kernel, and one graph replay per step
        eng.load(d)
        eng.step(gc, eps2, 0.0, "fp64", dt, graph=graph)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.step(gc, eps2, 0.0, "fp64", dt, graph=graph)
        torch.cuda.synchronize()
        t_gpu[graph] = (time.perf_counter() - t0) / steps
        eng.check()
    eng.close()
    return {"n": n, "grid": G, "G": gc, "eps2": eps2, "dt": dt,
            "cpu_full_step_s": t_cpu, "cpu_value": n / t_cpu / 1e6, "cpu_cores": cores,
            "cpu_full_step_1thread_s": t_cpu1, "cpu_kind": "port (oracle/bh_oracle.c)",
            "gpu_fp64_ms_per_step": t_gpu[False] * 1e3, "gpu_value": n / t_gpu[False] / 1e6,
            "gpu_fp64_graph_ms_per_step": t_gpu[True] * 1e3,
            "gpu_graph_value": n / t_gpu[True] / 1e6,
            "method_error_vs_direct_sum": {"oracle": e_orc, "gpu_fp64": e_gpu}}


# launch shapes the autotuner tries for the batched line (untimed, cached)
BATCHED_CANDIDATES = dict(p=(1, 2), yspan=(512,), waves_per_cu=(16, 32), tail_split=(1, 16))


def batched_line(steps, model_path, seed=0):
    """512 independent uniform systems (N log-uniform [1024, 262144], grid
    side ceil(sqrt(N/16))), FP32, one graph launch per step, launch shape
    picked by the autotuner on the loaded state (not timed)."""
    from barneshut import costmodel, inputs
    from barneshut.batched import BatchedSystems
    sizes, grids = inputs.batched_sizes(512, 1024, 262144, seed=seed)
    data = inputs.batched_uniform(sizes, seed=seed)
    out = {"systems": 512, "particles": int(sizes.sum()), "cells": int((grids.astype(np.int64) ** 2).sum())}
    law = (GC, 1e-12, 0.0)
    for name in ("static", "learned"):
        model = None
        if name == "learned":
            if not model_path or not os.path.exists(model_path):
                continue
            import torch
            model = costmodel.CostModel.load(model_path, device=torch.device("cuda"))
        b = BatchedSystems(sizes, grids, data, law, "fp32", DT, graph=True, model=model,
                           tuning="auto", autotune_candidates=BATCHED_CANDIDATES)
        out["tuning"] = b.tune_result["tuning"]
        sec = b.run(steps, warmup=1)
        out["recomputed_last_step"] = int(b.eng.ctx.stats()[6])
        inter = b.eng.interactions()[-1].astype(np.float64)    # placement of the last step
        out[name] = {"ms_per_step": sec * 1e3, "value": float(sizes.sum()) / sec / 1e6,
                     "interactions_last_step": float(inter.sum()),
                     "tflops_whole_step": FLOP_PER_INTERACTION * float(inter.sum()) / sec / 1e12}
        b.close()
    return out


# fixed launch shapes of the batched line's shape sweep (--batched-shapes)
BATCHED_SWEEP = (dict(p=1, yspan=512, waves_per_cu=32, tail_split=1, tail_items=0),
                 dict(p=2, yspan=512, waves_per_cu=32, tail_split=1, tail_items=0),
                 dict(p=2, yspan=512, waves_per_cu=32, tail_split=16, tail_items=1 << 20),
                 dict(p=1, yspan=512, waves_per_cu=16, tail_split=1, tail_items=0))


def batched_shapes(steps, seed=0):
    """The batched line's workload (static order, graph replay) under each
    fixed launch shape of BATCHED_SWEEP: ms per step, for comparing shapes
    (e.g. under rocprofv3 counters)."""
    from barneshut import inputs
    from barneshut.batched import BatchedSystems
    sizes, grids = inputs.batched_sizes(512, 1024, 262144, seed=seed)
    data = inputs.batched_uniform(sizes, seed=seed)
    out = []
    for tun in BATCHED_SWEEP:
        b = BatchedSystems(sizes, grids, data, (GC, 1e-12, 0.0), "fp32", DT, graph=True,
                           tuning=tun)
        sec = b.run(steps, warmup=1)
        b.close()
        log(f"batched {tun}: {sec * 1e3:.2f} ms/step")
        out.append({"tuning": tun, "ms_per_step": sec * 1e3})
    return out


def reference_config_line(steps, tuning64):
    """The reference's logged run (python-bh/data/frontera/nbody_out_singlegpu.dat):
    N = 10,000,000 uniform in [0,499)^2 (bin/gen_input.py), 120 x 120 grid,
    G = 1, dt = 0.005, FP64, one step per run; 166.837 s/step published."""
    from barneshut import inputs
    from barneshut.engine import Engine
    n, G = 10_000_000, 120
    d = inputs.reference_normal(n, seed=1)
    out = {"n": n, "grid": G}
    import torch
    for prec in ("fp64", "fp64_ieee"):
        e = Engine([n], [G], tuning=tuning64)
        e.load(d)
        e.step(1.0, 0.0, 0.0, prec, 0.005)
        e.check()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            e.step(1.0, 0.0, 0.0, prec, 0.005)
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / steps
        e.check()
        e.close()
        out[prec] = {"ms_per_step": sec * 1e3, "value": n / sec / 1e6}
    published = 166.83713   # s/step, nbody_out_singlegpu.dat:35
    sec = out["fp64"]["ms_per_step"] * 1e-3
    out.update({"precision": "fp64", "ms_per_step": sec * 1e3, "value": n / sec / 1e6,
                "published_s_per_step": published, "published_value": n / published / 1e6,
                "speedup_vs_published": published / sec,
                "published_hardware": "one GPU of the reference's Frontera runs (model not "
                                      "recorded)"})
    return out


def gravity_plugin_line():
    """The gravity plugin (bh_cloud_gravity, cpu_numba.py:8-92), FP64, the
    reference CPU law (r^2 <= 1e-5 dropped):
    * one cloud of 131,072 uniform particles (device rows): every ordered pair,
      1.718e10 terms per call, timed with HIP events on the launch stream;
    * a leaf of 64 particles against a COM pseudo-cloud of 16,384 rows -- the
      reference's per-leaf far-field call (seq/evaluation.py:95-98) -- on
      device rows, and on numpy rows (the host copies included).
    The CPU side is gravity_plugin_cpu (timed with the other baselines)."""
    import torch
    from barneshut.kernels import hiplib
    from barneshut.kernels.gravity import hip_cloud
    rng = np.random.default_rng(21)

    def rows(n, lo=0.0, hi=1.0):
        return _cloud_rows(rng, n, lo, hi)

    def timed(fn, calls):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(calls):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e-3 / calls

    n = 131072
    big = torch.from_numpy(rows(n)).to("cuda")
    t_big = timed(lambda: hiplib.cloud_gravity(big), 5)
    terms = float(n) * (n - 1)
    leaf, coms = rows(64, 0.0, 0.01), rows(16384)
    dl, dc = torch.from_numpy(leaf).to("cuda"), torch.from_numpy(coms).to("cuda")
    t_leaf = timed(lambda: hiplib.cloud_gravity(dl, dc), 200)
    t0 = time.perf_counter()
    for _ in range(50):
        hip_cloud.apply_force(leaf, coms, 1.0, False)
    t_host = (time.perf_counter() - t0) / 50
    ach = FLOP_PER_INTERACTION * terms / t_big / 1e12
    return {"kernel": "k_cloud (FP64, refined reciprocal)",
            "one_cloud": {"n": n, "terms": terms, "ms_per_call": t_big * 1e3,
                          "terms_per_s": terms / t_big, "achieved_tflops": ach,
                          "frac_of_fp64_peak": ach / FP64_PEAK_TFLOPS},
            "leaf_vs_coms": {"leaf": 64, "coms": 16384, "device_us_per_call": t_leaf * 1e6,
                             "host_rows_us_per_call": t_host * 1e6}}


def _cloud_rows(rng, n, lo=0.0, hi=1.0):
    r = np.zeros((n, 10))
    r[:, 0:2] = rng.uniform(lo, hi, (n, 2))
    r[:, 2] = rng.uniform(1.0, 2.0, n)
    return r


def gravity_plugin_cpu(n_gpu_cloud):
    """The oracle's restatement of cpu_numba's one-cloud loop (each pair once,
    both rows updated) on one host core, on a cloud of 8,192 uniform rows;
    extrapolated per pair to the GPU line's cloud."""
    from oracle import oracle as orc
    ns = 8192
    sample = _cloud_rows(np.random.default_rng(22), ns)
    t0 = time.perf_counter()
    orc.cloud_gravity(sample)
    t = time.perf_counter() - t0
    per_pair = t / (ns * (ns - 1) / 2.0)
    return {"kind": "port (oracle/bh_oracle.c, the loops of cpu_numba.py)", "cores": 1,
            "sample": f"one cloud of {ns} rows, {ns * (ns - 1) // 2} pairs, {t:.2f} s",
            "ns_per_pair": per_pair * 1e9,
            "one_cloud_s_extrapolated": per_pair * n_gpu_cloud * (n_gpu_cloud - 1) / 2.0}


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    ap.add_argument("--grid", type=int, default=GRID_DEFAULT)
    ap.add_argument("--seed", type=int, default=SEED)
    ap.add_argument("--p", type=int, default=2, help="FP32 path: particles per lane (1, 2, 4)")
    ap.add_argument("--yspan", type=int, default=512, help="max rows per work item")
    ap.add_argument("--waves-per-cu", type=int, default=32)
    ap.add_argument("--tail-split", type=int, default=16,
                    help="FP32 path: far-field slices per split item (1 = off)")
    ap.add_argument("--tail-items", type=int, default=1 << 20,
                    help="FP32 path: split the last this many queue items (>= items: all)")
    ap.add_argument("--p64", type=int, default=2, help="FP64 path: particles per lane")
    ap.add_argument("--yspan64", type=int, default=512, help="FP64 path: max rows per item")
    ap.add_argument("--waves64", type=int, default=32, help="FP64 path: waves per CU")
    ap.add_argument("--ieee-steps", type=int, default=3,
                    help="timed steps of the IEEE-division FP64 line (0 = skip)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick each path's launch shape with barneshut.autotune first")
    ap.add_argument("--cost-model", default=os.path.join(PKG, "costmodels", "clusters256.npz"),
                    help="trained cost-model file (.npz) or 'none'")
    ap.add_argument("--batched-model", default=os.path.join(PKG, "costmodels", "batched512.npz"))
    ap.add_argument("--c4-steps", type=int, default=5,
                    help="timed steps of each cost-model-workload path (0 = skip the block)")
    ap.add_argument("--batched-shapes", action="store_true",
                    help="only time the batched line under fixed launch shapes (one JSON line)")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-static", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-errors", action="store_true")
    ap.add_argument("--skip-extra", action="store_true",
                    help="skip the batched, reference-config and CPU-case blocks")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--recompute-kappa", type=float, default=RECOMPUTE_KAPPA,
                    help="FP32 paths: FP64 recompute of particles with error bound > kappa*|a| "
                         "(0 = off)")
    args = ap.parse_args()
    kappa = args.recompute_kappa

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              f"the multi-GPU slab decomposition is a design, not a runnable path", file=sys.stderr)
        sys.exit(2)

    import torch
    from barneshut import inputs

    if args.batched_shapes:
        print(json.dumps({"batched_shapes": batched_shapes(args.steps)}), flush=True)
        return

    if args.calibrate_copy_mib:
        src = torch.ones(args.calibrate_copy_mib * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy_mib} MiB read + written")
        del src, dst
    log(f"generating N={args.n} clusters={N_CLUSTERS} seed={args.seed}")
    d = inputs.gaussian_mixture(args.n, N_CLUSTERS, SIGMA, seed=args.seed)
    G = args.grid
    tuning = dict(p=args.p, yspan=args.yspan, waves_per_cu=args.waves_per_cu,
                  tail_split=args.tail_split, tail_items=args.tail_items)
    tuning64 = dict(p=args.p64, yspan=args.yspan64, waves_per_cu=args.waves64, tail_split=1,
                    tail_items=0)
    tuning64_ieee = dict(p=1, yspan=512, waves_per_cu=32, tail_split=1, tail_items=0)
    model = None
    if args.cost_model and args.cost_model != "none" and os.path.exists(args.cost_model):
        from barneshut import costmodel
        model = costmodel.CostModel.load(args.cost_model, device=torch.device("cuda"))

    def run(name, dd, precision, tun, steps, warmup, mdl=None):
        log(f"{name}: {precision}{' learned order' if mdl is not None else ''}")
        r = measure(dd, G, precision, tun, steps, warmup, mdl, kappa, args.autotune)
        log(f"{name}: {r['seconds_per_step'] * 1e3:.1f} ms/step, force kernel "
            f"{r['force_kernel_ms']:.1f} ms, {r['achieved']:.1f} TFLOP/s")
        return r

    res64 = None if args.skip_fp64 else run("c2 fp64", d, "fp64", tuning64, args.steps,
                                            args.warmup)
    res64i = None
    if not args.skip_fp64 and args.ieee_steps > 0:
        res64i = run("c2 fp64_ieee", d, "fp64_ieee", tuning64_ieee, args.ieee_steps, 1)
    res_static = None if args.skip_static else run("c2 fp32 static", d, "fp32", tuning,
                                                   args.steps, args.warmup)
    if model is not None:
        res = run("c2 fp32 learned", d, "fp32", tuning, args.steps, args.warmup, model)
        sched = "learned"
    else:
        res = res_static if res_static is not None else run("c2 fp32", d, "fp32", tuning,
                                                            args.steps, args.warmup)
        sched = "static"

    out = {
        "metric": METRIC,
        "value": res["value"],
        "unit": "Mparticle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["seconds_per_step"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "2-D grid-method step, N=4,194,304, 64 Gaussian clusters, "
                               "1024x1024 grid, timed steps 1..K of the initial state",
                   "n": args.n, "grid": G, "clusters": N_CLUSTERS, "sigma": list(SIGMA),
                   "mass": "U[1,2]", "G": GC, "dt": DT, "eps2": EPS2, "r2_cut": R2CUT,
                   "seed": args.seed, "schedule": sched, "cost_model":
                       os.path.relpath(args.cost_model, ROOT) if model is not None else None,
                   "tuning": res["tuning"], "work_items": res["work_items"],
                   "recompute_kappa": kappa},
        "interactions_per_step": {"near": float(res["near"].mean()),
                                  "far": float(res["far"].mean()),
                                  "first_step": float(res["near"][0] + res["far"][0]),
                                  "last_step": float(res["near"][-1] + res["far"][-1])},
        "recompute": {"kappa": kappa, "ms_per_step": res["recompute_ms"],
                      "share_of_step": res["recompute_ms"] / (res["seconds_per_step"] * 1e3),
                      "particles_last_step": res["recomputed_last_step"]},
        "roofline": {"bound": "valu", "achieved": res["achieved"], "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": res["achieved"] / FP32_PEAK_TFLOPS,
                     "traffic": TRAFFIC, "kernel": "k_force32<P,0,1>",
                     "force_kernel_ms": res["force_kernel_ms"],
                     "flop_per_launch": res["flop_per_launch"],
                     "issue_ceiling": ISSUE_CEILING_TFLOPS,
                     "frac_of_issue_ceiling": res["achieved"] / ISSUE_CEILING_TFLOPS},
    }
    if res_static is not None:
        out["fp32_static"] = path_summary(res_static, FP32_PEAK_TFLOPS)
        out["learned_vs_static"] = res_static["seconds_per_step"] / res["seconds_per_step"]
    if res64 is not None:
        out["fp64_plain"] = path_summary(res64, FP64_PEAK_TFLOPS)
        out["fp64_plain"]["kernel"] = "k_force64<P,0,1> (refined hardware reciprocal)"
        out["fp64_plain"]["issue_ceiling"] = ISSUE_CEILING64_TFLOPS
        out["fp64_plain"]["frac_of_issue_ceiling"] = res64["achieved"] / ISSUE_CEILING64_TFLOPS
        out["speedup_vs_fp64"] = res64["seconds_per_step"] / res["seconds_per_step"]
    if res64i is not None:
        out["fp64_ieee"] = path_summary(res64i, FP64_PEAK_TFLOPS)
        out["fp64_ieee"]["kernel"] = "k_force64<P,0,0> (IEEE division)"
        out["speedup_vs_fp64_ieee"] = res64i["seconds_per_step"] / res["seconds_per_step"]
    if not args.skip_errors:
        log("force errors")
        out["force_rel_err"] = force_errors(d, G, C2_ERROR_RUNS, kappa=kappa)
    if args.c4_steps > 0:
        log(f"cost-model workload: {C4_CLUSTERS} clusters, lognormal masses")
        d4 = inputs.gaussian_mixture(args.n, C4_CLUSTERS, C4_SIGMA, seed=C4_SEED,
                                     mass="lognormal")
        c4 = {"workload": f"N={args.n}, {C4_CLUSTERS} Gaussian clusters, sigma "
                          f"{list(C4_SIGMA)}, lognormal(0,0.5) masses, {G}x{G} grid",
              "seed": C4_SEED, "steps": args.c4_steps}
        r64 = run("c4 fp64", d4, "fp64", tuning64, args.c4_steps, 1)
        rst = run("c4 fp32 static", d4, "fp32", tuning, args.c4_steps, 1)
        c4["fp64"] = path_summary(r64, FP64_PEAK_TFLOPS)
        c4["fp32_static"] = path_summary(rst, FP32_PEAK_TFLOPS)
        c4["speedup_static_vs_fp64"] = r64["seconds_per_step"] / rst["seconds_per_step"]
        if model is not None:
            rle = run("c4 fp32 learned", d4, "fp32", tuning, args.c4_steps, 1, model)
            c4["fp32_learned"] = path_summary(rle, FP32_PEAK_TFLOPS)
            c4["speedup_learned_vs_fp64"] = r64["seconds_per_step"] / rle["seconds_per_step"]
            c4["learned_vs_static"] = rst["seconds_per_step"] / rle["seconds_per_step"]
        if not args.skip_errors:
            c4["force_rel_err"] = force_errors(d4, G, C2_ERROR_RUNS[:2], kappa=kappa)
        out["c4"] = c4
    if not args.skip_extra:
        log("batched systems")
        out["batched_512"] = batched_line(args.steps, args.batched_model)
        log("reference configuration")
        out["reference_config"] = reference_config_line(3, tuning64)
        log("reference CPU-runnable case")
        out["c1"] = c1_line(50)
        log("gravity plugin")
        out["gravity_plugin"] = gravity_plugin_line()
    if not args.skip_cpu:
        log("cpu baseline")
        if "gravity_plugin" in out:
            out["gravity_plugin"]["cpu"] = gravity_plugin_cpu(out["gravity_plugin"]["one_cloud"]["n"])
        cb = cpu_baseline(d, G)
        out["placement_bit_exact"] = cb.pop("placement_check")
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
