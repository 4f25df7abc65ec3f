"""Device-resident state of one or more independent systems and the per-step
phase calls into libbh_hip.so.

Particles live in two SoA buffer sets (A, B) of torch device tensors; the sort
phase permutes the current set into the other one and the two swap, like the
reference's d_particles / d_particles_sort pair
(python-bh/barneshut/implementations/singlegpu.py:48,62-67).
"""
from __future__ import annotations

import numpy as np

from .kernels import hiplib

FIELDS_F64 = ("px", "py", "mass", "vx", "vy", "ax", "ay")
FIELDS = FIELDS_F64 + ("pid",)


def _torch():
    import torch
    return torch


def alloc_particles(n, device):
    torch = _torch()
    d = {k: torch.zeros(n, dtype=torch.float64, device=device) for k in FIELDS_F64}
    d["pid"] = torch.zeros(n, dtype=torch.int32, device=device)
    return d


class Engine:
    """Grid-method step on the GPU for `n[s]` particles with `grid_dim[s]`^2
    cells per system s (a single system is the common case)."""

    def __init__(self, n, grid_dim, device=0, tuning=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("the HIP grid-method engine needs a GPU (torch.cuda is unavailable)")
        self.device = device
        self.dev = torch.device(f"cuda:{device}")
        self.ctx = hiplib.Context(device)
        self.n = np.ascontiguousarray(np.atleast_1d(n), dtype=np.int64)
        self.grid_dim = np.ascontiguousarray(np.atleast_1d(grid_dim), dtype=np.int32)
        self.ctx.set_systems(self.n, self.grid_dim)
        if tuning:
            self.ctx.set_tuning(**tuning)
        self.N = int(self.n.sum())
        self.C = int((self.grid_dim.astype(np.int64) ** 2).sum())
        self.bufs = [alloc_particles(self.N, self.dev), alloc_particles(self.N, self.dev)]
        self.cur = 0

    # ---- data movement
    @property
    def parts(self):
        return self.bufs[self.cur]

    @property
    def other(self):
        return self.bufs[1 - self.cur]

    def load(self, d):
        """Host dict (px, py, mass, vx, vy[, ax, ay], pid) -> current buffers."""
        torch = _torch()
        p = self.parts
        for k in FIELDS_F64:
            if k in d:
                p[k].copy_(torch.as_tensor(np.ascontiguousarray(d[k], dtype=np.float64)))
            else:
                p[k].zero_()
        p["pid"].copy_(torch.as_tensor(np.ascontiguousarray(d["pid"], dtype=np.int32)))
        torch.cuda.synchronize(self.dev)

    def fetch(self, which=None):
        """Current (or given) buffers -> host dict of numpy arrays."""
        p = self.parts if which is None else which
        return {k: p[k].cpu().numpy() for k in FIELDS}

    # ---- phases
    def create_tree(self, stream=None):
        """bbox + keys + histogram, offsets, ordered permutation into the other
        buffer set; then swap."""
        self.ctx.place(self.parts, stream)
        self.ctx.cumm(stream)
        self.ctx.sort(self.parts, self.other, stream)
        self.cur = 1 - self.cur

    def summarize(self, stream=None):
        self.ctx.summarize(self.parts, stream)

    def evaluate(self, G=1.0, eps2=0.0, r2_cut=0.0, precision="fp64", order=None, stream=None):
        self.ctx.evaluate(self.parts, G, eps2, r2_cut, precision, order, stream)

    def timestep(self, dt, stream=None):
        self.ctx.tick(self.parts, dt, stream)

    def step(self, G=1.0, eps2=0.0, r2_cut=0.0, precision="fp64", dt=0.0, do_tick=True,
             graph=False, stream=None):
        """One full round through bh_step / bh_step_graph."""
        self.ctx.step(self.parts, self.other, G, eps2, r2_cut, precision, dt, do_tick, graph,
                      stream)
        self.cur = 1 - self.cur

    def check(self, stream=None):
        """Synchronise and raise on a device-side error flag."""
        return self.ctx.stats(stream)

    # ---- exported state (copies)
    def counts(self):
        return self.ctx.export_host(hiplib.EXPORT_COUNTS, self.C, np.int32)

    def offsets(self):
        return self.ctx.export_host(hiplib.EXPORT_OFFSETS, self.C + 1, np.int32)

    def perm(self):
        return self.ctx.export_host(hiplib.EXPORT_PERM, self.N, np.int32)

    def keys(self):
        return self.ctx.export_host(hiplib.EXPORT_KEYS, self.N, np.int32)

    def keys_in(self):
        return self.ctx.export_host(hiplib.EXPORT_KEYS_IN, self.N, np.int32)

    def com(self):
        return self.ctx.export_host(hiplib.EXPORT_COM, 3 * self.C, np.float64).reshape(-1, 3)

    def bbox(self):
        return self.ctx.export_host(hiplib.EXPORT_BBOX, 5 * len(self.n), np.float64).reshape(-1, 5)

    def items(self):
        st = self.ctx.stats()
        n = int(st[hiplib.ST_NITEMS])
        return self.ctx.export_host(hiplib.EXPORT_ITEMS, 8 * n, np.int32).reshape(-1, 8)

    def energy(self, G=1.0, eps2=0.0):
        """[(U_s, K_s)] per system for the current state: potential of the law
        (direct FP64 sum, O(n^2): a diagnostic for small and medium systems)
        and kinetic energy."""
        return self.ctx.energy(self.parts, G, eps2)

    def energy_grid(self, G=1.0, eps2=0.0):
        """[(U_s, K_s)] per system with the potential evaluated by the grid
        method (about one FP32 force pass of work): placement and summary of
        the current state first, then the potential kernel."""
        self.create_tree()
        self.summarize()
        return self.ctx.energy_grid(self.parts, G, eps2)

    def interactions(self):
        """(near, far) per system and total, exact, from the current counts."""
        out = self.ctx.interactions()
        return out.reshape(-1, 2)

    def close(self):
        self.ctx.close()
