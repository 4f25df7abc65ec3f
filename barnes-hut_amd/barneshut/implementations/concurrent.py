"""K copies of one input stepped together: the MI355X form of the reference's
concurrent runner (python-bh/bin/run_2d_concurrent.py:53-87), which builds K
BarnesHut objects on the same input file and INI and runs them in K Python
threads.

Here the K copies are K independent systems of one context: their particles
are concatenated, every phase kernel covers all of them at once, and the
force kernel drains one work queue holding the items of every system.  Each
phase is still timed under the reference's Timer names (a phase now covers K
systems), and --check compares system 0 with the direct sum exactly as the
single-system implementation does.  All systems start from the same rows and
run the same kernels on the same cell structure, so they stay identical bit
for bit; `systems_identical()` checks that.
"""
import numpy as np

import barneshut.internals.particle as p

from .hip import HipBarnesHut


class ConcurrentHipBarnesHut(HipBarnesHut):

    def __init__(self, concurrent):
        super().__init__()
        if concurrent < 1:
            raise ValueError("concurrent must be >= 1")
        self.k = int(concurrent)

    def _init_device_arrays(self):
        if self.engine is not None:
            return
        from barneshut.engine import Engine
        self.grid_dim_ = self.grid_dim()
        n, k = self.n_particles, self.k
        self.engine = Engine([n] * k, [self.grid_dim_] * k, device=self.device,
                             tuning=self.tuning)
        self.engine.ctx.set_recompute(self.recompute_kappa)
        a = self.particles
        rep = lambda col: np.tile(a[:, col], k)   # noqa: E731
        self.engine.load(dict(px=rep(p.px), py=rep(p.py), mass=rep(p.mass), vx=rep(p.vx),
                              vy=rep(p.vy), ax=rep(p.ax), ay=rep(p.ay),
                              pid=rep(p.pid).astype(np.int32)))
        if self.schedule == "learned":
            import torch
            from barneshut.costmodel import CostModel
            self.model = CostModel.load(self.cost_model_path, device=self.engine.dev)
            self.order = torch.empty(self.engine.ctx.query(4), dtype=torch.int32,
                                     device=self.engine.dev)

    def _rows(self, h, keys, s):
        """N x 10 rows of system s in pid order (gx, gy from its cell keys)."""
        n, G = self.n_particles, self.grid_dim_
        sl = slice(s * n, (s + 1) * n)
        c_off = s * G * G
        a = np.empty((n, p.nfields), dtype=p.ftype)
        for col, name in ((p.px, "px"), (p.py, "py"), (p.mass, "mass"), (p.vx, "vx"),
                          (p.vy, "vy"), (p.ax, "ax"), (p.ay, "ay")):
            a[:, col] = h[name][sl]
        kk = keys[sl].astype(np.int64) - c_off
        a[:, p.gx] = kk // G
        a[:, p.gy] = kk % G
        a[:, p.pid] = h["pid"][sl]
        return a[np.argsort(h["pid"][sl], kind="stable")]

    def ensure_particles_id_ordered(self):
        """System 0's rows, in pid order, are the implementation's particles."""
        if self.engine is None or not self.host_dirty:
            return
        h = self.engine.fetch()
        self.particles = self._rows(h, self.engine.keys(), 0)
        self.host_dirty = False

    def systems_identical(self):
        """True when every system's rows equal system 0's bit for bit."""
        h = self.engine.fetch()
        keys = self.engine.keys()
        ref = self._rows(h, keys, 0)
        return all(np.array_equal(self._rows(h, keys, s), ref) for s in range(1, self.k))
