"""Writes tests/golden/shard_vectors.json: for the n=65536, d=16 graph of the
benchmark (seed 1), each rank's share (excess over the n(n-1) base, diameter)
for 2, 3 and 8 ranks under MATRIX and MATRIX_SAVING, as the reference's
MPI+CUDA flavour divides the work (aspl_mpi_cuda.cu:44-123).  The reference's
MPI flavours cannot be built here (no MPI, no CUDA), so the values come from
the oracle's restatement of that loop (oracle/odp_oracle.c oracle_eval_shard),
which tests/test_oracle.py pins against breadth-first searches and the whole
evaluation.  Run:  python tests/golden/make_shard_vectors.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import odp_amd as odp  # noqa: E402
import oracle  # noqa: E402


def main():
    n, d, seed = 65536, 16, 1
    adj = odp.generate_regular(n, d, seed)
    out = []
    for kind, name in ((oracle.MATRIX, "MATRIX"), (oracle.MATRIX_SAVING, "MATRIX_SAVING")):
        for procs in (2, 3, 8):
            parts = [list(oracle.evaluate_shard(adj, r, procs, nodes=n, kind=kind)) for r in range(procs)]
            out.append(dict(graph="regular", nodes=n, degree=d, seed=seed, kind=name, procs=procs, ranks=parts))
            print(name, procs, parts, flush=True)
    json.dump({"_comment": "one rank's (excess, diameter) per rank, from the oracle's restatement of "
                           "aspl_mpi_cuda.cu:44-123", "cases": out},
              open(os.path.join(HERE, "shard_vectors.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
