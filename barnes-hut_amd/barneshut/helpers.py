"""Grid sizing helpers, python-bh/barneshut/kernels/helpers/grid.py:5-48 and
python-bh/barneshut/implementations/base.py:13,221-231."""
import numpy as np

LEAF_OCCUPANCY = 0.7


def next_perfect_square(n):
    """grid.py:5-8: if n % sqrt(n) != 0 -> ceil(sqrt(n))**2 (numpy float ops)."""
    if n % np.sqrt(n) != 0:
        return np.power(np.ceil(np.sqrt(n)), 2)
    return n


def grid_dim_for(n_particles, particles_per_leaf):
    """base.py:221-231; particles_per_leaf == 1 gives one leaf (G = 1), as in
    the reference."""
    if particles_per_leaf == 1:
        nleaves = particles_per_leaf
    else:
        nleaves = n_particles / (LEAF_OCCUPANCY * particles_per_leaf)
    nleaves = next_perfect_square(nleaves)
    return int(nleaves ** 0.5)


def get_neighbor_cells(point, grid_dim):
    """grid.py:28-37: the (up to 8) cells around `point`, clipped to the grid."""
    x, y = point
    xs = range(max(x - 1, 0), min(x + 1, grid_dim - 1) + 1)
    ys = range(max(y - 1, 0), min(y + 1, grid_dim - 1) + 1)
    return [(a, b) for a in xs for b in ys if (a, b) != (x, y)]
