"""Implementation factory, python-bh/barneshut/barneshut.py:4-32.

``[general] implementation`` selects the class; "singlegpu" (the reference's
GPU path) and "hip" both map to the HIP implementation here, so an existing
singlegpu.ini works unchanged.  The reference's CPU and multi-GPU
implementations are not part of this package.
"""
from barneshut.internals.config import Config

_GPU_NAMES = ("singlegpu", "hip", "mi355x")


class BarnesHut:

    def __init__(self, ini_file, concurrent=1):
        """concurrent > 1: that many copies of the input, stepped together as
        independent systems of one context (the form bin/run_2d_concurrent.py
        uses for the reference's K concurrent BarnesHut objects)."""
        Config.read_file(ini_file)
        impl = Config.get("general", "implementation", fallback="sequential")
        print(f"Using {impl} implementation")
        if impl in _GPU_NAMES and concurrent > 1:
            from .implementations.concurrent import ConcurrentHipBarnesHut
            self.bh = ConcurrentHipBarnesHut(concurrent)
        elif impl in _GPU_NAMES:
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
