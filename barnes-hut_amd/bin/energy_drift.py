#!/usr/bin/env python3
"""Energy drift of a run of the large clustered workload (a mixture of
Gaussian clusters, sigma log-uniform [0.001, 0.03], plus a uniform
background; masses U[1,2]), measured with the grid-method energy
(bh_energy_grid: about one FP32 force pass per evaluation, so it scales to
the full 67,108,864-particle workload where the direct O(N^2) sum cannot).

    energy_drift.py --n 67108864 --grid 4096 --clusters 4096 --steps 20 \
        --energy-every 2 --budget-s 1000

Each step: placement, COM, FP32 cell-relative forces with the FP64 recompute
(learned queue order when a cost model is given), FP64 tick.  The energy is
evaluated before the first step and every --energy-every steps; with
--budget-s the run stops taking steps when the next step and a final energy
evaluation would not fit.  Prints a line per step and energy, then one JSON
line with the energy history and the relative drift (E - E0) / |E0|.
--direct also evaluates the direct FP64 energy (O(N^2): small N only).
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.realpath(__file__))))

import numpy as np  # noqa: E402

PKG = os.path.dirname(os.path.dirname(os.path.realpath(__file__)))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=1 << 18)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--clusters", type=int, default=16)
    ap.add_argument("--background", type=float, default=0.2)
    ap.add_argument("--seed", type=int, default=71)
    ap.add_argument("--G", type=float, default=1e-6)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--eps2", type=float, default=1e-10,
                    help="softening^2 (default: eps = 1e-5; clipping the mixture to the box "
                         "stacks particles on the corners, where the unsoftened law is 0/0)")
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--cost-model", default=os.path.join(PKG, "costmodels", "clusters256.npz"),
                    help="cost model (.npz) for the learned queue order, or 'none'")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--energy-every", type=int, default=1)
    ap.add_argument("--budget-s", type=float, default=0.0)
    ap.add_argument("--direct", action="store_true")
    ap.add_argument("--parts", type=int, default=4,
                    help="force pass in this many launches (progress between them)")
    a = ap.parse_args(argv)

    import torch
    from barneshut import inputs
    from barneshut.engine import Engine
    t_start = time.perf_counter()

    def log(msg):
        print(f"[drift {time.perf_counter() - t_start:7.1f}s] {msg}", flush=True)

    d = inputs.gaussian_mixture(a.n, a.clusters, (0.001, 0.03), seed=a.seed,
                                background=a.background)
    log(f"generated N={a.n} clusters={a.clusters} background={a.background} seed={a.seed}")
    eng = Engine([a.n], [a.grid])
    eng.load(d)
    del d
    order = model = None
    if a.cost_model != "none" and not a.precision.startswith("fp64"):
        from barneshut.costmodel import CostModel
        model = CostModel.load(a.cost_model, device=eng.dev)
        order = torch.empty(eng.ctx.query(4), dtype=torch.int32, device=eng.dev)

    def energy(step):
        t0 = time.perf_counter()
        eng.create_tree()
        eng.summarize()
        eng.check()
        log(f"energy at step {step}: placed")
        U, K = eng.ctx.energy_grid(eng.parts, a.G, a.eps2)[0]
        rec = {"step": step, "U": float(U), "K": float(K), "E": float(U + K)}
        if a.direct:
            Ud, Kd = eng.energy(a.G, a.eps2)[0]
            rec.update(U_direct=float(Ud), E_direct=float(Ud + Kd))
        dt_e = time.perf_counter() - t0
        log(f"energy at step {step}: U {U:.12e} K {K:.6e} E {U + K:.12e} ({dt_e:.1f}s)")
        return rec, dt_e

    hist = []
    rec, t_energy = energy(0)
    hist.append(rec)
    t_step = 0.0
    done = 0
    for s in range(1, a.steps + 1):
        if a.budget_s > 0:
            # this step (the first one estimated by an energy evaluation, which
            # costs about one force pass) and one energy after it must fit
            left = a.budget_s - (time.perf_counter() - t_start)
            need = 1.05 * ((t_step if t_step > 0 else t_energy) + t_energy)
            if left < need:
                log(f"budget: stopping after {done} steps ({left:.0f}s left)")
                break
        t0 = time.perf_counter()
        eng.create_tree()
        eng.summarize()
        eng.ctx.build_items()
        if model is not None:
            eng.ctx.learned_order(model.params, order)
        eng.check()
        log(f"step {s}: placed")
        for part in range(a.parts):   # launches over ranges of the work queue
            eng.ctx.force_pass(eng.parts, a.G, a.eps2, 0.0, a.precision, order, part=part,
                               nparts=a.parts)
            eng.check()
            log(f"step {s}: forces {part + 1}/{a.parts}")
        eng.timestep(a.dt)
        eng.check()
        t_step = time.perf_counter() - t0
        done = s
        log(f"step {s}/{a.steps} {t_step:.1f}s")
        if s % a.energy_every == 0:
            rec, t_energy = energy(s)
            hist.append(rec)
    if hist[-1]["step"] != done:
        rec, _ = energy(done)
        hist.append(rec)
    eng.close()
    E0 = hist[0]["E"]
    out = {"n": a.n, "grid": a.grid, "clusters": a.clusters, "background": a.background,
           "seed": a.seed, "G": a.G, "dt": a.dt, "eps2": a.eps2, "precision": a.precision,
           "learned_order": model is not None, "steps_done": done,
           "energy": hist, "drift": [(h["E"] - E0) / abs(E0) for h in hist],
           "seconds_per_step": t_step, "seconds_per_energy": t_energy}
    if a.direct:
        Ed0 = hist[0]["E_direct"]
        out["drift_direct"] = [(h["E_direct"] - Ed0) / abs(Ed0) for h in hist]
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
