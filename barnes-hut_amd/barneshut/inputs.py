"""Synthetic particle sets and the reference's binary input format.

File format (python-bh/bin/gen_input.py:12-13,27-54; reader
python-bh/barneshut/implementations/base.py:47-72): a native int32 N followed
by N records of five native float64 values (x, y, mass, vx, vy).

Generators return a dict of float64 numpy arrays px, py, mass, vx, vy and an
int32 pid (0..N-1); every generator is seeded.
"""
from __future__ import annotations

import math
import struct

import numpy as np

_UPPER = float(np.nextafter(1.0, 0.0))   # largest double below 1.0


def _pack(px, py, mass, vx=None, vy=None):
    n = len(px)
    z = np.zeros(n)
    return dict(px=np.ascontiguousarray(px, dtype=np.float64),
                py=np.ascontiguousarray(py, dtype=np.float64),
                mass=np.ascontiguousarray(mass, dtype=np.float64),
                vx=np.ascontiguousarray(z if vx is None else vx, dtype=np.float64),
                vy=np.ascontiguousarray(z if vy is None else vy, dtype=np.float64),
                pid=np.arange(n, dtype=np.int32))


def uniform(n, seed=0, mass=(1.0, 2.0)):
    """Positions U[0,1)^2, masses U[mass], zero velocity."""
    rng = np.random.default_rng(seed)
    p = rng.random((n, 2))
    m = rng.uniform(mass[0], mass[1], n)
    return _pack(p[:, 0], p[:, 1], m)


def reference_normal(n, seed=0):
    """The reference generator's "normal" distribution (gen_input.py:40-51):
    x, y U[0, 499), mass U[0.5, 5], velocity U[-1, 1]."""
    rng = np.random.default_rng(seed)
    x = rng.random(n) * 499.0
    y = rng.random(n) * 499.0
    vx = (rng.random(n) - 0.5) * 2
    vy = (rng.random(n) - 0.5) * 2
    m = 0.5 + rng.random(n) * 4.5
    return _pack(x, y, m, vx, vy)


def gaussian_mixture(n, n_clusters, sigma=(0.005, 0.05), seed=0, mass="uniform",
                     background=0.0):
    """Mixture of isotropic Gaussians: centres U[0,1)^2, sigma log-uniform in
    `sigma`, points clipped to [0, 1); a `background` fraction is uniform.
    mass: "uniform" -> U[1,2]; "lognormal" -> lognormal(0, 0.5)."""
    rng = np.random.default_rng(seed)
    nb = int(round(n * background))
    nc = n - nb
    c = rng.random((n_clusters, 2))
    s = np.exp(rng.uniform(math.log(sigma[0]), math.log(sigma[1]), n_clusters))
    lab = rng.integers(0, n_clusters, nc)
    p = c[lab] + rng.standard_normal((nc, 2)) * s[lab][:, None]
    if nb:
        p = np.concatenate([p, rng.random((nb, 2))])
        p = p[rng.permutation(n)]
    p = np.clip(p, 0.0, _UPPER)
    if mass == "uniform":
        m = rng.uniform(1.0, 2.0, n)
    elif mass == "lognormal":
        m = rng.lognormal(0.0, 0.5, n)
    else:
        raise ValueError(mass)
    return _pack(p[:, 0], p[:, 1], m)


def batched_sizes(n_sys=512, lo=1024, hi=262144, seed=0):
    """System sizes log-uniform in [lo, hi] and grid side ceil(sqrt(N/16))."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(math.log(lo), math.log(hi), n_sys)).astype(np.int64)
    n = np.clip(n, lo, hi)
    g = np.ceil(np.sqrt(n / 16.0)).astype(np.int32)
    return n, g


def batched_uniform(sizes, seed=0):
    """Concatenated independent uniform systems (positions U[0,1)^2, m U[1,2]);
    pid restarts at 0 in every system."""
    parts = [uniform(int(n), seed=seed * 100003 + i) for i, n in enumerate(sizes)]
    out = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    return out


def write_particles_file(fname, d):
    """Writes the gen_input.py binary format."""
    n = len(d["px"])
    rec = np.empty((n, 5), dtype=np.float64)
    rec[:, 0] = d["px"]
    rec[:, 1] = d["py"]
    rec[:, 2] = d["mass"]
    rec[:, 3] = d["vx"]
    rec[:, 4] = d["vy"]
    with open(fname, "wb") as fp:
        fp.write(struct.pack("i", n))
        fp.write(rec.tobytes())


def read_particles_file(fname):
    """Reads the gen_input.py binary format (base.py:47-72): pid = row index."""
    with open(fname, "rb") as fp:
        (n,) = struct.unpack("i", fp.read(struct.calcsize("i")))
        rec = np.frombuffer(fp.read(n * 40), dtype=np.float64)
    if rec.size != 5 * n:
        raise ValueError(f"{fname}: expected {n} records, file holds {rec.size // 5}")
    rec = rec.reshape(n, 5)
    return _pack(rec[:, 0], rec[:, 1], rec[:, 2], rec[:, 3], rec[:, 4])
