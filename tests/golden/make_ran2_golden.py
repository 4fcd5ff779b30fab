"""Writes tests/golden/ran2_stream.npz from the reference's own RNG.

Runs oracle/_ref/ref_rng, which oracle/Makefile.ref compiles from the
reference sources src/mod_nrtype.f90, src/mod_nrutil.f90 and
src/mod_ranstate.f90 plus our driver oracle/ref_rng_driver.f90.  For each
seed it records, after init_rng_by_seed(seed), n pairs (rands(), ran_sysm())
as REAL(SP) bit patterns.
"""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
BIN = os.path.join(ROOT, "oracle", "_ref", "ref_rng")
SEEDS = (1, 7, 11, 42, 122, 1000, 1234, 2147483647)
N = 4096


def main():
    if not os.path.exists(BIN):
        sys.exit("build oracle/_ref/ref_rng first: make -f oracle/Makefile.ref")
    out = {}
    for s in SEEDS:
        r = subprocess.run([BIN, str(s), str(N)], capture_output=True, text=True, check=True)
        a = np.array([[int(v) for v in line.split()] for line in r.stdout.strip().splitlines()], np.int64)
        out[f"seed_{s}"] = a.astype(np.uint32).view(np.int32)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ran2_stream.npz"), **out)


if __name__ == "__main__":
    main()
