"""Autotuner for the force kernel's launch shape.

Candidates are (p, yspan, waves_per_cu, tail_split): particles per lane, the
largest row span of a work item, persistent waves per compute unit, and the
far-field slices each FP32 item is split into.  Each candidate
is timed with HIP events on the engine's current state (after create_tree and
summarize).  For FP32 paths a candidate is admitted only if its accelerations
stay within `max_rel_err` (normwise) of an FP64 pass on the same state: the
item span sets the FP32 origin distance, so it trades accuracy for occupancy.
Results are cached per (device, n, grid_dim, precision) in a JSON file.
"""
from __future__ import annotations

import itertools
import json
import os

DEFAULT_CANDIDATES = dict(p=(1, 2, 4), yspan=(128, 512), waves_per_cu=(16, 32),
                          tail_split=(1, 8, 16))
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "bh_mi355x_autotune.json")


def _key(eng, precision):
    """Device, systems, precision and the library build (a rebuilt kernel is
    tuned again)."""
    import torch
    from .kernels import hiplib
    name = torch.cuda.get_device_name(eng.dev)
    n = ",".join(str(int(v)) for v in eng.n[:8])
    g = ",".join(str(int(v)) for v in eng.grid_dim[:8])
    try:
        st = os.stat(hiplib.LIB_PATH)
        build = f"{st.st_size}:{int(st.st_mtime)}"
    except OSError:
        build = "?"
    return f"{name}|S={len(eng.n)}|n={n}|G={g}|{precision}|lib={build}"


def _time_force(eng, law, precision, order=None, reps=2):
    import torch
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        eng.parts["ax"].zero_()
        eng.parts["ay"].zero_()
        start.record()
        eng.ctx.force_pass(eng.parts, *law, precision, order)
        end.record()
        end.synchronize()
        best = min(best, start.elapsed_time(end))
    return best


def tune(eng, law=(1.0, 0.0, 0.0), precision="fp32", candidates=None, max_rel_err=1e-6,
         reps=2, cache=CACHE, use_cache=True, log=None):
    """Pick the fastest admissible tuning for `eng`'s current state and set it.
    law = (G, eps2, r2_cut).  Returns dict(tuning, ms, table)."""
    import torch
    key = _key(eng, precision)
    if use_cache and cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                hit = json.load(f).get(key)
            if hit:
                eng.ctx.set_tuning(**hit["tuning"])
                return hit
        except (OSError, ValueError):
            pass
    cand = dict(DEFAULT_CANDIDATES)
    if candidates:
        cand.update(candidates)
    saved = {k: eng.parts[k].clone() for k in ("ax", "ay")}
    ref = None
    if precision != "fp64":
        eng.ctx.build_items()
        eng.parts["ax"].zero_()
        eng.parts["ay"].zero_()
        eng.ctx.force_pass(eng.parts, *law, "fp64")
        ref = (eng.parts["ax"].clone(), eng.parts["ay"].clone())
        den = torch.sqrt((ref[0] ** 2 + ref[1] ** 2).sum())
    table = []
    best = None
    splits = cand["tail_split"] if precision != "fp64" else (1,)
    for p, ys, w, sp in itertools.product(cand["p"], cand["yspan"], cand["waves_per_cu"], splits):
        tun = dict(p=int(p), yspan=int(ys), waves_per_cu=int(w), tail_split=int(sp),
                   tail_items=(1 << 20) if sp > 1 else 0)
        eng.ctx.set_tuning(**tun)
        eng.ctx.build_items()
        ms = _time_force(eng, law, precision, reps=reps)
        err = 0.0
        if ref is not None:
            num = torch.sqrt(((eng.parts["ax"] - ref[0]) ** 2
                              + (eng.parts["ay"] - ref[1]) ** 2).sum())
            err = float(num / den)
        ok = err <= max_rel_err
        table.append(dict(tuning=tun, ms=ms, rel_err=err, admitted=ok))
        if log:
            log(f"tune {tun} {ms:.2f} ms err {err:.2e}{'' if ok else ' (rejected)'}")
        if ok and (best is None or ms < best["ms"]):
            best = dict(tuning=tun, ms=ms, rel_err=err)
    for k in ("ax", "ay"):
        eng.parts[k].copy_(saved[k])
    if best is None:
        raise RuntimeError("no admissible tuning (all exceeded max_rel_err)")
    eng.ctx.set_tuning(**best["tuning"])
    eng.ctx.build_items()
    best["table"] = table
    if cache:
        try:
            os.makedirs(os.path.dirname(cache), exist_ok=True)
            data = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    data = json.load(f)
            data[key] = {"tuning": best["tuning"], "ms": best["ms"], "rel_err": best["rel_err"]}
            with open(cache, "w") as f:
                json.dump(data, f, indent=1)
        except (OSError, ValueError):
            pass
    return best
