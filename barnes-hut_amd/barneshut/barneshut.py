"""Implementation factory, python-bh/barneshut/barneshut.py:4-32.

``[general] implementation`` selects the class; "singlegpu" (the reference's
GPU path) and "hip" both map to the HIP implementation here, so an existing
singlegpu.ini works unchanged.  The reference's CPU and multi-GPU
implementations are not part of this package.
"""
from barneshut.internals.config import Config

_GPU_NAMES = ("singlegpu", "hip", "mi355x")


class BarnesHut:

    def __init__(self, ini_file):
        Config.read_file(ini_file)
        impl = Config.get("general", "implementation", fallback="sequential")
        print(f"Using {impl} implementation")
        if impl in _GPU_NAMES:
            from .implementations.hip import HipBarnesHut
            self.bh = HipBarnesHut()
        else:
            raise NotImplementedError(
                f"implementation '{impl}' is not provided by this package; "
                f"use one of {', '.join(_GPU_NAMES)}")

    def read_input_file(self, file):
        self.bh.read_particles_from_file(file)

    def run(self, *args, **kwargs):
        self.bh.run(*args, **kwargs)
