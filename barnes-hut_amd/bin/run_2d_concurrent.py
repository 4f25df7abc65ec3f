#!/usr/bin/env python3
"""K independent systems stepped together (the reference's
python-bh/bin/run_2d_concurrent.py:29-90 runs K BarnesHut objects in Python
threads; here they share one context, one work queue and one graph per step).

    run_2d_concurrent.py K nruns [--n-min 1024 --n-max 262144 --seed 0
                                  --precision fp32 --steps 10]

System sizes are log-uniform in [n-min, n-max], grid side ceil(sqrt(N/16)),
positions U[0,1)^2, masses U[1,2].
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.realpath(__file__))))

from barneshut import inputs  # noqa: E402
from barneshut.batched import BatchedSystems  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("k", type=int)
    ap.add_argument("nruns", type=int)
    ap.add_argument("--n-min", type=int, default=1024)
    ap.add_argument("--n-max", type=int, default=262144)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--G", type=float, default=1e-6)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--no-graph", action="store_true")
    a = ap.parse_args(argv)
    sizes, grids = inputs.batched_sizes(a.k, a.n_min, a.n_max, a.seed)
    data = inputs.batched_uniform(sizes, a.seed)
    b = BatchedSystems(sizes, grids, data, (a.G, 0.0, 0.0), a.precision, a.dt,
                       graph=not a.no_graph)
    sec = b.run(a.nruns)
    print(f"{a.k} systems, {int(sizes.sum())} particles: {sec * 1e3:.2f} ms/step, "
          f"{sizes.sum() / sec / 1e6:.2f} Mparticle-steps/s")
    b.close()


if __name__ == "__main__":
    main()
