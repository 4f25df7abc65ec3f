#!/usr/bin/env python3
"""Command-line runner with the reference's interface
(python-bh/bin/run_2d.py:28-61):

    run_2d.py input_file nwarmups nruns configfile [--debug] [--check]

nwarmups runs, Timer reset, nruns runs, then the median report.
"""
import argparse
import logging
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.realpath(__file__))))

from barneshut import BarnesHut  # noqa: E402
from timer import Timer  # noqa: E402


def main(argv=None):
    parser = argparse.ArgumentParser()
    parser.add_argument('input_file', help="Path to input file")
    parser.add_argument('nwarmups', help="How many runs to execute as warmup", type=int)
    parser.add_argument('nruns', help="Number of different runs to execute", type=int)
    parser.add_argument('configfile', help="Path to config file")
    parser.add_argument('--debug', help="Turn debug on", action="store_true")
    parser.add_argument('--check', help="Check accuracy after every round", action="store_true")
    args = parser.parse_args(argv)
    if args.debug:
        logging.basicConfig(level=logging.DEBUG)
    else:
        logging.disable(logging.CRITICAL)
    bh = BarnesHut(args.configfile)
    bh.read_input_file(args.input_file)
    for _ in range(args.nwarmups):
        bh.run(check_accuracy=args.check)
    print("warmup done")
    Timer.reset()
    for _ in range(args.nruns):
        bh.run(check_accuracy=args.check)
    Timer.print()
    return bh


if __name__ == "__main__":
    main()
