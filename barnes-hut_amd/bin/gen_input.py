#!/usr/bin/env python3
"""Input generator with the reference's interface and file format
(python-bh/bin/gen_input.py:1-57):

    gen_input.py {normal|uniform|gaussians} num_particles fout [--seed S]

normal    -- the reference's distribution: x, y U[0,499), mass U[0.5,5], v U[-1,1]
uniform   -- x, y U[0,1), mass U[1,2], v = 0
gaussians -- 64 clipped Gaussians in the unit box, mass U[1,2], v = 0
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.realpath(__file__))))

from barneshut import inputs  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("distribution")
    ap.add_argument("num_particles", type=int)
    ap.add_argument("fout")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    print(f"Using distribution {a.distribution}")
    if a.distribution == "normal":
        d = inputs.reference_normal(a.num_particles, a.seed)
    elif a.distribution == "uniform":
        d = inputs.uniform(a.num_particles, a.seed)
    elif a.distribution == "gaussians":
        d = inputs.gaussian_mixture(a.num_particles, 64, seed=a.seed)
    else:
        raise SystemExit(f"unknown distribution {a.distribution}")
    inputs.write_particles_file(a.fout, d)


if __name__ == "__main__":
    main()
