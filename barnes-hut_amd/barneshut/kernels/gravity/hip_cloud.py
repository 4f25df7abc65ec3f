"""The HIP cloud-to-cloud kernel behind the gravity plugin interface
``fn(self_cloud, other_cloud, G, update_other)``
(python-bh/barneshut/kernels/gravity/cpu_numba.py:8-28).

A cloud is anything with a ``particles`` attribute -- the reference's Cloud
(internals/cloud.py:39-41), rows in the N x 10 layout of internals/particle.py
-- or a bare array of such rows.  Rows held in a float64 device tensor are
updated in place with no copy.  Rows held on the host (numpy) go through
pinned staging memory kept per device: their positions and masses are packed
into it (accelerations zeroed), copied to the device in one transfer,
evaluated there, and the sums come back in one transfer and are added to the
cloud's ax, ay -- one add per cloud in turn, so clouds that are overlapping
views of one array both receive theirs, as in cpu_numba_kernel.  As in
cpu_numba_kernel the accelerations are added to, never reset, and one cloud
passed twice means its pairs among themselves.
"""
from __future__ import annotations

import os
import threading

import numpy as np

try:
    from .. import hiplib
except ImportError:
    # loaded by file path from another project's registry (whose own package
    # may also be called barneshut): take hiplib from this file's tree
    import importlib.util as _ilu
    _spec = _ilu.spec_from_file_location(
        "bh_hip_hiplib", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                      "hiplib.py"))
    hiplib = _ilu.module_from_spec(_spec)
    _spec.loader.exec_module(hiplib)

R2_CUT = 1e-5   # cpu_numba.py:36,44,59,67,81,89: pairs with r^2 <= 1e-5 are dropped
NCOL = 7        # px, py, mass, vx, vy, ax, ay: what the kernel reads or writes
AX = 5

_stage = {}                 # device index -> (pinned host rows, device rows)
_stage_lock = threading.Lock()


def get_kernel_function(r2_cut=R2_CUT, eps2=0.0):
    """fn(self_cloud, other_cloud, G, update_other=False) with the given law."""
    def hip_kernel(self_cloud, other_cloud, G, update_other=False):
        apply_force(self_cloud, other_cloud, G, update_other, r2_cut=r2_cut, eps2=eps2)
    return hip_kernel


def _rows(cloud):
    return cloud.particles if hasattr(cloud, "particles") else cloud


def _check_host(rows, name):
    if not isinstance(rows, np.ndarray) or rows.dtype != np.float64 or rows.ndim != 2 \
            or rows.shape[1] < NCOL:
        raise ValueError(f"{name}: rows must be a float64 array of shape (n, >= {NCOL})")
    if not rows.flags.writeable:
        raise ValueError(f"{name}: rows must be writeable (accelerations are added in place)")


def _staging(dev, nrows):
    """Pinned host and device buffers of at least nrows x NCOL, grown by doubling."""
    torch = hiplib._torch()
    h, d = _stage.get(dev.index, (None, None))
    if h is None or h.shape[0] < nrows:
        cap = max(nrows, 2 * (h.shape[0] if h is not None else 0), 4096)
        h = torch.empty((cap, NCOL), dtype=torch.float64, pin_memory=True)
        d = torch.empty((cap, NCOL), dtype=torch.float64, device=dev)
        _stage[dev.index] = (h, d)
    return h, d


def apply_force(self_cloud, other_cloud, G, update_other=False, r2_cut=R2_CUT, eps2=0.0,
                stream=None):
    """Adds the attraction of other_cloud's rows to self_cloud's accelerations
    (and, with update_other, the reverse), cpu_numba.py:8-28."""
    torch = hiplib._torch()
    same = self_cloud is other_cloud   # cpu_numba.py:9 compares the clouds themselves
    A = _rows(self_cloud)
    B = A if same else _rows(other_cloud)
    tA, tB = isinstance(A, torch.Tensor), isinstance(B, torch.Tensor)
    for t, name in ((A, "self_cloud"), (B, "other_cloud")):
        if isinstance(t, torch.Tensor):
            if not t.is_cuda:
                raise ValueError(f"{name}: a tensor of rows must be on the GPU")
        else:
            _check_host(t, name)
    if tA and tB:
        hiplib.cloud_gravity(A, None if same else B, G=G, eps2=eps2, r2_cut=r2_cut,
                             update_other=update_other, stream=stream)
        return
    dev = A.device if tA else (B.device if tB else
                               torch.device("cuda", torch.cuda.current_device()))
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    host = [] if tA else [A]
    if not same and not tB:
        host.append(B)
    with _stage_lock:
        nrows = sum(len(r) for r in host)
        h, d = _staging(dev, nrows)
        hv = h.numpy()
        views, o = [], 0
        for r in host:
            hv[o:o + len(r), :AX] = r[:, :AX]
            hv[o:o + len(r), AX:AX + 2] = 0.0   # the device forms the sums alone
            views.append(d[o:o + len(r)])
            o += len(r)
        with torch.cuda.stream(s):
            d[:nrows].copy_(h[:nrows], non_blocking=True)
        dA = A if tA else views[0]
        dB = dA if same else (B if tB else views[-1])
        hiplib.cloud_gravity(dA, None if same else dB, G=G, eps2=eps2, r2_cut=r2_cut,
                             update_other=update_other, stream=s)
        # staged rows that receive sums: all of them with update_other, else only
        # self's (staged first) if self is on the host
        back = nrows if update_other else (0 if tA else len(A))
        if back:
            with torch.cuda.stream(s):
                h[:back].copy_(d[:back], non_blocking=True)
        s.synchronize()
        o = 0
        for r in host:   # `accelerations += acc` per cloud, as cpu_numba.py:13,24-25,28
            if r is A or update_other:
                r[:, AX:AX + 2] += hv[o:o + len(r), AX:AX + 2]
            o += len(r)
