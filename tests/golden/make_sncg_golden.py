"""Writes the SNCG fixtures under tests/golden/ with the reference's own
exporter (src/bench/snc_export.cpp), built unchanged by oracle/Makefile into
oracle/_ref/snc_export.

Run in a container that has the reference tree (make -C oracle ref, then
python tests/golden/make_sncg_golden.py). Each fixture is the exporter's output
file for one command line; sncg_cases.json lists the command lines.
"""
from __future__ import annotations

import json
import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXPORTER = os.path.join(ROOT, "oracle", "_ref", "snc_export")

# (fixture name, exporter arguments): every structure, both delay modes.
CASES = [
    ("sncg_static_snc", "--structure static-snc --layers 60,40,8 --synapse-budget 500 --seed 3"),
    ("sncg_random_sparse_random_delays",
     "--structure random-sparse --layers 50,30,6 --synapse-budget 400 --seed 5 --delay-max 7"),
    ("sncg_dense_delay2", "--structure dense --layers 12,9,4 --delay 2"),
    ("sncg_static_snc_distance_delays",
     "--structure static-snc --layers 64,48 --synapse-budget 900 --seed 2 --delay-max 9 "
     "--delay-mode distance"),
]


def main():
    if not os.path.exists(EXPORTER):
        sys.exit(f"{EXPORTER} not built: run make -C oracle ref in a tree with the reference")
    for name, args in CASES:
        out = os.path.join(HERE, name + ".bin")
        subprocess.run([EXPORTER, *shlex.split(args), "--out", out], check=True,
                       stdout=subprocess.DEVNULL)
        print(f"{name}: {os.path.getsize(out)} bytes")
    with open(os.path.join(HERE, "sncg_cases.json"), "w") as f:
        json.dump([{"name": n, "args": a} for n, a in CASES], f, indent=1)


if __name__ == "__main__":
    main()
