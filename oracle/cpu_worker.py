"""One process of the CPU baseline's ideal y split (bench.py's cpu_baseline
leg only).  TEST INFRASTRUCTURE: it times the oracle, never the product.

TiltRec-mpi splits the volume rows over MPI ranks (main-mpi.cpp:710-740);
at pitch 0 each row is independent, so N processes on disjoint rows are the
split without its margin bug.  Usage:
    python3 -m oracle.cpu_worker INPUT.npz VARIANT
INPUT holds coef [P][8], stack rows [P][rows][X], T and iterations; the
process runs SIRT on its rows with liborc_VARIANT.so and prints the
seconds the computation took as JSON.
"""
import json
import sys
import time

import numpy as np

from oracle import orc


def main():
    d = np.load(sys.argv[1], allow_pickle=False)
    L = orc.variant(sys.argv[2])
    T, iters = int(d["T"]), int(d["iters"])
    X = d["rows"].shape[2]
    t = time.perf_counter()
    orc.sirt_f32(d["coef"], d["rows"], T, orc.origin(X, T, T / 2), iters, 1.0, L=L)
    print(json.dumps({"seconds": time.perf_counter() - t}), flush=True)


if __name__ == "__main__":
    main()
