"""Gravity kernel registry, python-bh/barneshut/kernels/gravity/__init__.py:6-41.

``[sequential] force_calculation`` selects the cloud-to-cloud kernel
``fn(self_cloud, other_cloud, G, update_other)`` that ``Cloud.apply_force``
calls (python-bh/barneshut/internals/cloud.py:113-114); the choice is made
once and cached on the function, as the reference does.

* ``hip`` and ``numba``: the HIP kernel (``bh_cloud_gravity``) with the law of
  the reference's default ``cpu_numba`` kernel: a = G m (x_j - x_i) / r^2,
  pairs with r^2 <= 1e-5 dropped (cpu_numba.py:36,44,59,67,81,89);
* ``nop``: does nothing (gravity/__init__.py:3-4,33-34).

The reference's other formulations (vect, blas, guvectorize-*, pykokkos) are
alternative CPU/GPU codes of the same law that its default.ini marks as
baselines or broken; they are not provided and raise NotImplementedError.
"""
from barneshut.internals import config as _config
from barneshut.internals.config import Config

PROVIDED = ("hip", "numba", "nop")


def nop(*args):
    pass


def get_gravity_kernel():
    if "kernel_fn" not in dir(get_gravity_kernel):
        # the reference's default.ini value when no configuration was read
        fc = ("numba" if _config.config_instance is None
              else Config.get("sequential", "force_calculation", fallback="numba"))
        if fc in ("hip", "numba"):
            from . import hip_cloud
            get_gravity_kernel.kernel_fn = hip_cloud.get_kernel_function()
        elif fc == "nop":
            get_gravity_kernel.kernel_fn = nop
        else:
            raise NotImplementedError(
                f"[sequential] force_calculation = {fc} is not provided by this package; "
                f"use one of {', '.join(PROVIDED)}")
    return get_gravity_kernel.kernel_fn


def reset():
    """Forget the cached choice (the next get_gravity_kernel() reads Config again)."""
    if "kernel_fn" in dir(get_gravity_kernel):
        del get_gravity_kernel.kernel_fn
