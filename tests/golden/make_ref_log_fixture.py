"""Extracts the placement rows the reference printed in its run logs
(python-bh/data/frontera/nbody_out_parla{1,2,4}_eager.dat, "sample parts"
blocks: the first 5 rows of each placement slice, all 10 columns
px, py, mass, vx, vy, ax, ay, gx, gy, pid) into ref_log_placement.json.

These runs used N = 10,000,000 particles uniform in [0, 499)^2
(python-bh/bin/gen_input.py:9,40-42) on a 120 x 120 grid
(configs/parla*.ini max_particles_per_box = 1000).  The logs do not record
the bounding box, so the fixture is a weak known answer for the cell
formula.  Only the first round of each log is used (later rounds ran after
ticks, on a grown box).  Run from the repository root with the reference at
/root/reference; the output is committed.
"""
import json
import os
import re

REF = "/root/reference/python-bh/data/frontera"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_log_placement.json")


def blocks(text):
    for m in re.finditer(r"sample parts\d+ \[\[(.*?)\]\]", text, re.S):
        nums = [float(v) for v in re.findall(r"[-+]?\d+\.\d+e[-+]\d+", m.group(1))]
        for i in range(0, len(nums), 10):
            yield nums[i:i + 10]


def main():
    rows = []
    for name in ("nbody_out_parla1_eager.dat", "nbody_out_parla2_eager.dat",
                 "nbody_out_parla4_eager.dat"):
        with open(os.path.join(REF, name)) as f:
            text = f.read()
        # first round only: later rounds ran after ticks, on a grown box
        text = text.split("stopping end-to-end", 1)[0]
        for r in blocks(text):
            if True:
                if len(r) == 10 and r not in rows:
                    rows.append(r)
    with open(OUT, "w") as f:
        json.dump({"source": "python-bh/data/frontera/nbody_out_parla{1,2,4}_eager.dat",
                   "columns": ["px", "py", "mass", "vx", "vy", "ax", "ay", "gx", "gy", "pid"],
                   "grid_dim": 120, "domain": [0.0, 499.0], "rows": rows}, f, indent=1)
    print(f"{len(rows)} rows -> {OUT}")


if __name__ == "__main__":
    main()
