"""The MI355X implementation behind ``implementation = singlegpu | hip``.

Mirrors SingleGPUBarnesHut (python-bh/barneshut/implementations/singlegpu.py:
16-112): device arrays are created on the first create_tree
(singlegpu.py:28-39), every phase synchronises before returning so the Timer
phases measure device time (singlegpu.py:64,77,92,99), and
ensure_particles_id_ordered / get_particles return rows in pid order
(singlegpu.py:101-112).

Differences from the reference path, each deliberate:
  * the bounding box comes from the current device positions every step
    (the reference reads the stale host copy, base.py:215);
  * cell order is x-major and the order inside a cell is numpy's
    structured-sort order of the sequential path (sequential.py:37), so the
    permutation is deterministic (the reference's atomic scatter is not);
  * empty cells are skipped in the far field (seq/evaluation.py:90), which
    also avoids the reference kernel's 0/0 for a particle at (0, 0).

Extra keys, section [hip]:
  precision   fp64 (default) | fp64_ieee | fp32 | fp32_kahan_term | fp32_plain
  eps2        softening added to r^2 (default 0 = reference GPU law)
  r2_cut      drop pairs with r^2 <= r2_cut (reference CPU law: 1e-5)
  grid_dim    explicit grid side (0: from max_particles_per_box)
  schedule    static (default) | learned
  cost_model  .npz file of a trained cost model (schedule = learned)
  p, yspan, waves_per_cu, tail_split, tail_items   force-kernel launch shape of
              the FP32 precisions (autotune = True to tune)
  p64, yspan64, waves_per_cu64   launch shape of the FP64 precisions (default
              2, 512, 32: the fastest on the 4M-particle workloads; FP64 items
              are never split, so no slice planes are allocated)
  recompute_kappa  FP32 precisions: FP64 recompute of the particles whose
              rounding-error bound exceeds kappa * |a| (default 1.5; 0 = off)
"""
import numpy as np

import barneshut.internals.particle as p
from barneshut.internals.config import Config

from .base import BaseBarnesHut


class HipBarnesHut(BaseBarnesHut):

    def __init__(self):
        super().__init__()
        self.G, self.eps2, self.r2_cut = self.law()
        self.tick_seconds = float(Config.get("bh", "tick_seconds"))
        self.precision = Config.get("hip", "precision", fallback="fp64")
        self.schedule = Config.get("hip", "schedule", fallback="static")
        self.cost_model_path = Config.get("hip", "cost_model", fallback="")
        self.autotune = Config.get("hip", "autotune", fallback="False") == "True"
        self.tuning = dict(p=int(Config.get("hip", "p", fallback="2")),
                           yspan=int(Config.get("hip", "yspan", fallback="512")),
                           waves_per_cu=int(Config.get("hip", "waves_per_cu", fallback="32")),
                           tail_split=int(Config.get("hip", "tail_split", fallback="16")),
                           tail_items=int(Config.get("hip", "tail_items", fallback="1048576")))
        if self.precision in ("fp64", "fp64_ieee"):
            self.tuning = dict(p=int(Config.get("hip", "p64", fallback="2")),
                               yspan=int(Config.get("hip", "yspan64", fallback="512")),
                               waves_per_cu=int(Config.get("hip", "waves_per_cu64",
                                                           fallback="32")),
                               tail_split=1, tail_items=0)
        self.device = int(Config.get("hip", "device", fallback="0"))
        self.recompute_kappa = float(Config.get("hip", "recompute_kappa", fallback="1.5"))
        self.engine = None
        self.model = None
        self.order = None
        self.host_dirty = False

    def _init_device_arrays(self):
        """Allocates the device state and copies the particles once."""
        if self.engine is not None:
            return
        from barneshut.engine import Engine
        self.grid_dim_ = self.grid_dim()
        self.engine = Engine([self.n_particles], [self.grid_dim_], device=self.device,
                             tuning=self.tuning)
        self.engine.ctx.set_recompute(self.recompute_kappa)
        a = self.particles
        self.engine.load(dict(px=a[:, p.px], py=a[:, p.py], mass=a[:, p.mass], vx=a[:, p.vx],
                              vy=a[:, p.vy], ax=a[:, p.ax], ay=a[:, p.ay],
                              pid=a[:, p.pid].astype(np.int32)))
        if self.schedule == "learned":
            import torch
            from barneshut.costmodel import CostModel
            self.model = CostModel.load(self.cost_model_path, device=self.engine.dev)
            self.order = torch.empty(self.engine.ctx.query(4), dtype=torch.int32,
                                     device=self.engine.dev)

    def create_tree(self):
        self._init_device_arrays()
        self.engine.create_tree()
        self.engine.check()
        self.host_dirty = True

    def summarize(self):
        self.engine.summarize()
        self.engine.check()

    def evaluate(self):
        eng = self.engine
        if self.autotune:
            from barneshut import autotune
            autotune.tune(eng, (self.G, self.eps2, self.r2_cut), self.precision)
            self.autotune = False
        eng.ctx.build_items()
        order = None
        if self.model is not None:
            eng.ctx.learned_order(self.model.params, self.order)
            order = self.order
        eng.ctx.force_pass(eng.parts, self.G, self.eps2, self.r2_cut, self.precision, order)
        eng.check()
        self.host_dirty = True

    def timestep(self):
        self.engine.timestep(self.tick_seconds)
        self.engine.check()
        self.host_dirty = True

    def ensure_particles_id_ordered(self):
        """Device state -> host N x 10 rows sorted by pid (singlegpu.py:101-103)."""
        if self.engine is None or not self.host_dirty:
            return
        h = self.engine.fetch()
        keys = self.engine.keys().astype(np.int64)
        G = self.grid_dim_
        n = len(h["px"])
        a = np.empty((n, p.nfields), dtype=p.ftype)
        a[:, p.px] = h["px"]
        a[:, p.py] = h["py"]
        a[:, p.mass] = h["mass"]
        a[:, p.vx] = h["vx"]
        a[:, p.vy] = h["vy"]
        a[:, p.ax] = h["ax"]
        a[:, p.ay] = h["ay"]
        a[:, p.gx] = keys // G
        a[:, p.gy] = keys % G
        a[:, p.pid] = h["pid"]
        order = np.argsort(h["pid"], kind="stable")
        self.particles = a[order]
        self.host_dirty = False

    def get_particles(self, sample_indices=None):
        self.ensure_particles_id_ordered()
        if sample_indices is None:
            return self.particles
        return {i: self.particles[i].copy() for i in sample_indices}

    def cleanup(self):
        pass

    def close(self):
        if self.engine is not None:
            self.engine.close()
            self.engine = None
