"""Writes tests/golden/mutate_vectors.json: per case of tests/test_mutate.py, the
SHA-256 of the text that the reference library's generation and 2-opt moves
print through oracle/_ref/odp_ref_gen (built by `make -C oracle ref` from the
reference sources).  Run from the repository root: python tests/golden/make_mutate_vectors.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import test_mutate as T  # noqa: E402

out = {}
for case in T.CASES:
    out[T.case_id(case)] = hashlib.sha256(T.ref_text(*case).encode()).hexdigest()
with open(os.path.join(HERE, "mutate_vectors.json"), "w") as f:
    json.dump(out, f, indent=1, sort_keys=True)
print(f"{len(out)} cases written")
