"""MI355X-native rebuild of python-bh's grid-method step.

``from barneshut import BarnesHut`` mirrors python-bh/barneshut/__init__.py:1.
"""
from .barneshut import BarnesHut  # noqa: F401
