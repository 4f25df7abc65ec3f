"""The HIP cloud-to-cloud kernel behind the gravity plugin interface
``fn(self_cloud, other_cloud, G, update_other)``
(python-bh/barneshut/kernels/gravity/cpu_numba.py:8-28).

A cloud is anything with a ``particles`` attribute -- the reference's Cloud
(internals/cloud.py:39-41), rows in the N x 10 layout of internals/particle.py
-- or a bare array of such rows.  Rows held on the host (numpy) are copied to
the device (their first 7 columns), evaluated there, and only their ax, ay
columns are written back; rows held in a float64 device tensor are updated in
place with no copy.  As in cpu_numba_kernel, the accelerations are added to,
never reset, and one cloud passed twice means its pairs among themselves.
"""
from __future__ import annotations

import numpy as np

from barneshut.kernels import hiplib

R2_CUT = 1e-5   # cpu_numba.py:36,44,59,67,81,89: pairs with r^2 <= 1e-5 are dropped
NCOL = 7        # px, py, mass, vx, vy, ax, ay: what the kernel reads or writes
AX = 5


def get_kernel_function(r2_cut=R2_CUT, eps2=0.0):
    """fn(self_cloud, other_cloud, G, update_other=False) with the given law."""
    def hip_kernel(self_cloud, other_cloud, G, update_other=False):
        apply_force(self_cloud, other_cloud, G, update_other, r2_cut=r2_cut, eps2=eps2)
    return hip_kernel


def _rows(cloud):
    return cloud.particles if hasattr(cloud, "particles") else cloud


def _to_device(rows, name, dev):
    """(device tensor, host array to write back or None)."""
    torch = hiplib._torch()
    if isinstance(rows, torch.Tensor):
        if not rows.is_cuda:
            raise ValueError(f"{name}: a tensor of rows must be on the GPU")
        return rows, None
    if not isinstance(rows, np.ndarray) or rows.dtype != np.float64 or rows.ndim != 2 \
            or rows.shape[1] < NCOL:
        raise ValueError(f"{name}: rows must be a float64 array of shape (n, >= {NCOL})")
    if not rows.flags.writeable:
        raise ValueError(f"{name}: rows must be writeable (accelerations are added in place)")
    d = torch.from_numpy(np.ascontiguousarray(rows[:, :NCOL])).to(dev)
    return d, rows


def apply_force(self_cloud, other_cloud, G, update_other=False, r2_cut=R2_CUT, eps2=0.0,
                stream=None):
    """Adds the attraction of other_cloud's rows to self_cloud's accelerations
    (and, with update_other, the reverse), cpu_numba.py:8-28."""
    torch = hiplib._torch()
    same = self_cloud is other_cloud   # cpu_numba.py:9 compares the clouds themselves
    A = _rows(self_cloud)
    B = A if same else _rows(other_cloud)
    dev = A.device if isinstance(A, torch.Tensor) else (
        B.device if isinstance(B, torch.Tensor) else torch.device("cuda",
                                                                  torch.cuda.current_device()))
    dA, hA = _to_device(A, "self_cloud", dev)
    dB, hB = (dA, hA) if same else _to_device(B, "other_cloud", dev)
    hiplib.cloud_gravity(dA, None if same else dB, G=G, eps2=eps2, r2_cut=r2_cut,
                         update_other=update_other, stream=stream)
    if stream is not None and (hA is not None or hB is not None):
        stream.synchronize()   # the copies below run on the current stream
    if hA is not None:
        hA[:, AX:AX + 2] = dA[:, AX:AX + 2].cpu().numpy()
    if hB is not None and not same and update_other:
        hB[:, AX:AX + 2] = dB[:, AX:AX + 2].cpu().numpy()
