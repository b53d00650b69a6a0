"""Writes tests/golden/vectors.json: for graphs regenerated from
(generator, parameters, seed) by libodp_hip's seeded generators, the diameter,
sum and ASPL that the REFERENCE library (oracle/_ref, built from the reference
sources by oracle/Makefile) returns.  Run in the container that holds the
reference:  python tests/golden/make_vectors.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from goldens import C4_SAMPLE, C5_SAMPLE, build_case  # noqa: E402

CASES = [
    # single general graphs; the first three are the n=1024 d=6 workload
    dict(graph="regular", nodes=1024, degree=6, seed=1),
    dict(graph="regular", nodes=1024, degree=6, seed=2),
    dict(graph="regular", nodes=1024, degree=6, seed=3),
    dict(graph="regular", nodes=256, degree=3, seed=4),
    dict(graph="regular", nodes=1000, degree=5, seed=5),
    dict(graph="regular", nodes=2048, degree=8, seed=6),
    dict(graph="regular", nodes=4096, degree=16, seed=7),
    dict(graph="regular", nodes=8192, degree=12, seed=8),
    dict(graph="regular", nodes=16384, degree=16, seed=9),
    dict(graph="regular", nodes=16384, degree=24, seed=10),
    dict(graph="regular", nodes=20000, degree=4, seed=11),
    dict(graph="regular", nodes=30000, degree=10, seed=12),
    dict(graph="regular", nodes=4096, degree=3, seed=13),
    # 2-opt candidates of the batched workload
    dict(graph="regular", nodes=3000, degree=8, seed=14, two_opt=1),
    dict(graph="regular", nodes=12000, degree=24, seed=15, two_opt=1),
    dict(graph="regular", nodes=1500, degree=4, seed=16, two_opt=3),
    # circulant-symmetric general graphs
    dict(graph="general_s", nodes=512, degree=4, symmetries=2, seed=1),
    dict(graph="general_s", nodes=1024, degree=6, symmetries=4, seed=2),
    dict(graph="general_s", nodes=3000, degree=4, symmetries=3, seed=3),
    dict(graph="general_s", nodes=8192, degree=8, symmetries=8, seed=4),
    # rotationally symmetric grids
    dict(graph="grid", width=16, height=16, degree=4, length=3, symmetries=4, seed=1),
    dict(graph="grid", width=32, height=32, degree=6, length=4, symmetries=2, seed=2),
    dict(graph="grid", width=24, height=24, degree=8, length=4, symmetries=1, seed=3),
    dict(graph="grid", width=64, height=64, degree=8, length=4, symmetries=4, seed=4),
    dict(graph="grid", width=100, height=100, degree=6, length=3, symmetries=4, seed=5),
    # full-size workloads: n=65536 d=16 MATRIX, 256x256 grid s=4
    dict(graph="regular", nodes=65536, degree=16, seed=1),
    dict(graph="grid", width=256, height=256, degree=8, length=4, symmetries=4, seed=1),
]
# graphs of the batched workload (1024 candidates) and served chains after three
# requests of one swap each (tests/goldens.py builds them from these parameters)
CASES += [dict(graph="batch", index=g, batch_seed=2024) for g in C4_SAMPLE]
CASES += [dict(graph="chain", chain=i, swaps=3, rng=9000 + i) for i in C5_SAMPLE]


def _key(c):
    return json.dumps({k: v for k, v in c.items() if k not in ("diameter", "sum", "aspl")}, sort_keys=True)


def main():
    """Computes the cases not yet in vectors.json (all with --all)."""
    path = os.path.join(HERE, "vectors.json")
    have = {}
    if os.path.exists(path) and "--all" not in sys.argv:
        have = {_key(c): c for c in json.load(open(path))["cases"]}
    out = []
    for c in CASES:
        if _key(c) in have:
            out.append(have[_key(c)])
            continue
        kw, adj = build_case(c)
        t = time.time()
        (d, s, a), _ = oracle.run_reference(adj, threads=True, timeout=3600, **kw)
        rec = dict(c, diameter=d, sum=s, aspl=a.hex())
        print(json.dumps(rec), f"({time.time() - t:.1f}s)", flush=True)
        out.append(rec)
    json.dump({"_comment": "expected values from the reference library (oracle/_ref, OpenMP build, default kind) "
                           "on graphs regenerated from the listed generator parameters",
               "cases": out}, open(path, "w"), indent=1)


if __name__ == "__main__":
    main()
