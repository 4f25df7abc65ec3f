"""Implementation contract and run loop, python-bh/barneshut/implementations/
base.py:34-258 (BaseBarnesHut): input reader, the create_tree -> summarize ->
evaluate -> timestep round with the same Timer phase names, and the sampled
O(S*N) accuracy check behind ``--check``.

One deliberate difference: the reference compares the implementation's
accelerations after the tick has zeroed them (base.py:124-133 with
gpu/grid.py:170-171), which only works with skip_timestep = True.  Here the
accelerations are snapshotted after evaluate and before the tick.
"""
import logging
import random

import numpy as np

import barneshut.internals.particle as p
from barneshut.helpers import grid_dim_for
from barneshut.inputs import read_particles_file
from barneshut.internals.config import Config
from timer import Timer


def nsquared_acc_check(particles, samples, G, eps2=0.0, r2_cut=0.0):
    """Direct sum for the sampled rows (base.py:16-32), vectorised over the
    other particles; self is skipped by pid as in the reference.  Same law
    switches as the implementation (softening eps2, cutoff r2_cut)."""
    pos = particles[:, p.px:p.py + 1]
    m = particles[:, p.mass]
    pids = particles[:, p.pid]
    for i in range(samples.shape[0]):
        d = pos - samples[i, p.px:p.py + 1]
        r2 = (d * d).sum(axis=1)
        with np.errstate(divide="ignore", invalid="ignore"):
            w = m / (r2 + eps2)
        if r2_cut > 0:
            w = np.where(r2 > r2_cut, w, 0.0)
        w = np.where(pids == samples[i, p.pid], 0.0, w)
        samples[i, p.ax] += G * (d[:, 0] * w).sum()
        samples[i, p.ay] += G * (d[:, 1] * w).sum()


class BaseBarnesHut:

    def __init__(self):
        self.particles = None
        self.n_particles = None
        self.checking_accuracy = False
        self.sample_size = int(Config.get("general", "sample_check_size", fallback="10"))
        self.particles_per_leaf = int(Config.get("grid", "max_particles_per_box"))
        self.skip_timestep = Config.get("general", "skip_timestep", fallback="False") == "True"
        self.evaluation_rounds = int(Config.get("general", "evaluation_rounds", fallback="1"))
        self.last_check = None
        print(f"Skiping timestep? {self.skip_timestep}")

    def read_particles_from_file(self, filename):
        """N x 10 FP64 rows px, py, mass, vx, vy, 0, 0, 0, 0, pid
        (base.py:47-72; the format of bin/gen_input.py)."""
        d = read_particles_file(filename)
        self.set_particles(d)

    def set_particles(self, d):
        """Host dict (px, py, mass, vx, vy[, pid]) -> the N x 10 row array."""
        n = len(d["px"])
        a = np.zeros((n, p.nfields), dtype=p.ftype)
        a[:, p.px] = d["px"]
        a[:, p.py] = d["py"]
        a[:, p.mass] = d["mass"]
        a[:, p.vx] = d.get("vx", 0.0)
        a[:, p.vy] = d.get("vy", 0.0)
        a[:, p.pid] = np.arange(n, dtype=np.float64) if "pid" not in d else d["pid"]
        self.particles = a
        self.n_particles = n

    def grid_dim(self):
        """[hip] grid_dim overrides max_particles_per_box sizing."""
        g = int(Config.get("hip", "grid_dim", fallback="0"))
        return g if g > 0 else grid_dim_for(self.n_particles, self.particles_per_leaf)

    def create_tree(self):
        raise NotImplementedError()

    def summarize(self):
        raise NotImplementedError()

    def evaluate(self):
        raise NotImplementedError()

    def timestep(self):
        raise NotImplementedError()

    def cleanup(self):
        pass

    def run(self, check_accuracy=False):
        """base.py:94-135: `rounds` iterations of the four phases, each timed."""
        n_iterations = int(Config.get("general", "rounds", fallback="1"))
        self.checking_accuracy = check_accuracy
        if self.checking_accuracy:
            sample_indices = self.generate_sample_indices()
        with Timer.get_handle("end-to-end"):
            for _ in range(n_iterations):
                if self.checking_accuracy:
                    nsquared_sample = self.preround_accuracy_check(sample_indices)
                with Timer.get_handle("grid_creation"):
                    self.create_tree()
                with Timer.get_handle("summarization"):
                    self.summarize()
                for _ in range(self.evaluation_rounds):
                    with Timer.get_handle("evaluation"):
                        self.evaluate()
                if self.checking_accuracy:
                    self.ensure_particles_id_ordered()
                    impl_sample = self.get_particles(sample_indices)
                if not self.skip_timestep:
                    with Timer.get_handle("timestep"):
                        self.timestep()
                if self.checking_accuracy:
                    self.last_check = self.check_accuracy(sample_indices, nsquared_sample,
                                                          impl_sample)
        self.cleanup()

    def get_particles(self, sample_indices=None):
        raise NotImplementedError()

    def generate_sample_indices(self):
        """base.py:149-154: fixed seed 0, random.sample of sample_check_size."""
        random.seed(0)
        samples = random.sample(range(self.n_particles), self.sample_size)
        logging.debug(f"acc_check: sample is {samples}")
        return samples

    def law(self):
        return (float(Config.get("bh", "grav_constant")),
                float(Config.get("hip", "eps2", fallback="0")),
                float(Config.get("hip", "r2_cut", fallback="0")))

    def preround_accuracy_check(self, sample_indices):
        """base.py:156-179: direct sum for the sampled particles before the round."""
        particles = self.get_particles()
        samples_nd = np.zeros((len(sample_indices), p.nfields))
        for i, si in enumerate(sample_indices):
            samples_nd[i] = particles[si].copy()
            samples_nd[i, p.ax] = 0
            samples_nd[i, p.ay] = 0
        G, eps2, r2_cut = self.law()
        nsquared_acc_check(particles, samples_nd, G, eps2, r2_cut)
        return {samples_nd[i, p.pid]: samples_nd[i].copy() for i in range(samples_nd.shape[0])}

    def check_accuracy(self, sample_indices, nsquared_sample, impl_sample=None):
        """base.py:181-205: mean over samples of |a_nsq - a_impl| / |a_nsq| per
        axis (printed), plus the per-sample vector relative error (returned)."""
        if impl_sample is None:
            impl_sample = self.get_particles(sample_indices)
        cum_err = np.zeros(2)
        vec = []
        for i in sample_indices:
            nsq = nsquared_sample[i][p.ax:p.ay + 1]
            aprox = impl_sample[i][p.ax:p.ay + 1]
            assert nsquared_sample[i][p.pid] == impl_sample[i][p.pid]
            with np.errstate(divide="ignore", invalid="ignore"):
                cum_err += np.fabs((nsq - aprox) / nsq)
            vec.append(float(np.hypot(*(nsq - aprox)) / np.hypot(*nsq)))
        cum_err /= float(self.sample_size)
        err = ", ".join([str(x) for x in cum_err])
        print(f"avg relative {self.sample_size} points on each axis:\n\t{err}")
        return {"per_axis_mean": cum_err.tolist(), "vector_rel": vec}
