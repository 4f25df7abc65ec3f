#!/usr/bin/env python3
"""Concurrent runner with the reference's interface
(python-bh/bin/run_2d_concurrent.py:29-90):

    run_2d_concurrent.py input_file nwarmups nruns concurrent configfile [--debug] [--check]

The reference builds `concurrent` BarnesHut objects on the same input file and
INI and runs each round of them in Python threads.  Here the `concurrent`
copies are independent systems of one context (one work queue for all of
them): nwarmups runs, "warmup done", Timer reset, nruns runs, then the median
report.  A run is the INI's `rounds` steps of all copies; each Timer phase
covers every copy.  --check compares system 0 with the direct sum (the
single-system report) and checks that every copy equals system 0 bit for bit.

    run_2d_concurrent.py K nruns --synthetic [--n-min ...]

keeps the older synthetic form: K uniform systems of log-uniform sizes,
stepped by one graph launch per step.
"""
import argparse
import logging
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.realpath(__file__))))

from barneshut import BarnesHut  # noqa: E402
from timer import Timer  # noqa: E402


def main(argv=None):
    argv = sys.argv[1:] if argv is None else list(argv)
    if "--synthetic" in argv:
        return synthetic(argv)
    parser = argparse.ArgumentParser()
    parser.add_argument('input_file', help="Path to input file")
    parser.add_argument('nwarmups', help="How many runs to execute as warmup", type=int)
    parser.add_argument('nruns', help="Number of different runs to execute", type=int)
    parser.add_argument('concurrent', help="Number of different concurrent apps to run", type=int)
    parser.add_argument('configfile', help="Path to config file")
    parser.add_argument('--debug', help="Turn debug on", action="store_true")
    parser.add_argument('--check', help="Check accuracy after every round", action="store_true")
    args = parser.parse_args(argv)
    if args.debug:
        logging.basicConfig(level=logging.DEBUG)
    else:
        logging.disable(logging.CRITICAL)
    if args.concurrent < 1:
        parser.error("concurrent must be >= 1")
    bh = BarnesHut(args.configfile, concurrent=args.concurrent)
    bh.read_input_file(args.input_file)

    def one_run():
        bh.run(check_accuracy=args.check)
        if args.check and args.concurrent > 1:
            same = bh.bh.systems_identical()
            print(f"{args.concurrent} concurrent systems identical: {same}")
            if not same:
                raise RuntimeError("concurrent copies of one input diverged")

    for _ in range(args.nwarmups):
        one_run()
    print("warmup done")
    Timer.reset()
    for _ in range(args.nruns):
        one_run()
    Timer.print()
    return bh


def synthetic(argv):
    from barneshut import inputs
    from barneshut.batched import BatchedSystems
    ap = argparse.ArgumentParser()
    ap.add_argument("k", type=int)
    ap.add_argument("nruns", type=int)
    ap.add_argument("--synthetic", action="store_true")
    ap.add_argument("--n-min", type=int, default=1024)
    ap.add_argument("--n-max", type=int, default=262144)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--G", type=float, default=1e-6)
    ap.add_argument("--dt", type=float, default=1e-5)
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
    return b


if __name__ == "__main__":
    main()
