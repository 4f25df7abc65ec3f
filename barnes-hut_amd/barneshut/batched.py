"""Many independent systems stepped together on one GPU.

The reference runs K independent BarnesHut objects in Python threads
(python-bh/bin/run_2d_concurrent.py:29-90).  Here all systems share one
context: their particles are concatenated, every kernel covers all systems
(cells of system s live at global indices [c_off_s, c_off_s + G_s^2)), the
force kernel drains one work queue that holds the items of every system, and
a whole step is replayed as one captured HIP graph.
"""
from __future__ import annotations

import time

import numpy as np

from .engine import Engine


# Many small systems: 64-particle items and no far-field slicing (most
# systems' COM lists are shorter than one slice would be worth).
BATCHED_TUNING = dict(p=1, yspan=512, waves_per_cu=32, tail_split=1, tail_items=0)


class BatchedSystems:

    def __init__(self, sizes, grid_dims, data, law=(1.0, 0.0, 0.0), precision="fp32", dt=0.0,
                 tuning=None, device=0, model=None, graph=True, autotune_candidates=None,
                 recompute_kappa=1.5):
        """tuning: None (BATCHED_TUNING), a dict, or "auto": the autotuner picks
        the launch shape on the loaded state (autotune_candidates narrows it).
        recompute_kappa: FP32 precisions re-evaluate in FP64 the particles whose
        rounding-error bound exceeds kappa * |a| (0 turns it off)."""
        self.sizes = np.asarray(sizes, np.int64)
        self.grid_dims = np.asarray(grid_dims, np.int32)
        auto = isinstance(tuning, str)
        if auto and tuning != "auto":
            raise ValueError(f"tuning must be None, a dict or 'auto', not {tuning!r}")
        self.eng = Engine(self.sizes, self.grid_dims, device=device,
                          tuning=BATCHED_TUNING if (tuning is None or auto) else tuning)
        self.eng.ctx.set_recompute(recompute_kappa)
        self.eng.load(data)
        self.tune_result = None
        if auto:
            from . import autotune
            self.eng.create_tree()
            self.eng.summarize()
            self.tune_result = autotune.tune(self.eng, law, precision,
                                             candidates=autotune_candidates)
            self.eng.load(data)
        self.law = law
        self.precision = precision
        self.dt = dt
        self.graph = graph
        self.model = model
        self.order = None
        if model is not None:
            import torch
            self.order = torch.empty(self.eng.ctx.query(4), dtype=torch.int32, device=self.eng.dev)

    def step(self):
        e = self.eng
        G, eps2, r2 = self.law
        if self.model is None:
            e.ctx.step(e.parts, e.other, G, eps2, r2, self.precision, self.dt, True, self.graph)
        else:
            e.ctx.step_learned(e.parts, e.other, self.model.params, self.order, G, eps2, r2,
                               self.precision, self.dt, True, self.graph)
        e.cur = 1 - e.cur

    def run(self, steps, warmup=1):
        """Returns seconds per step (synchronised wall clock)."""
        import torch
        for _ in range(warmup):
            self.step()
        self.eng.check()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        self.eng.check()
        return dt

    def system_slices(self):
        off = np.concatenate([[0], np.cumsum(self.sizes)])
        return [slice(int(off[i]), int(off[i + 1])) for i in range(len(self.sizes))]

    def close(self):
        self.eng.close()
