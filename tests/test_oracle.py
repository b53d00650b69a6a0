"""Pins the oracle (oracle/odp_oracle.c) before anything is checked against it.

Each function of the restatement is checked against something it does not
contain: the reference's own golden outputs, the values the reference library
printed for its symmetric fixtures, the reference library itself on generated
graphs (committed vectors, plus live runs where oracle/_ref is built), closed
forms, networkx, and invariants (MATRIX == MATRIX_SAVING == BFS for every
chunk; symmetric table == its s=1 expansion).  test_planted_faults_are_caught
plants known mistakes in the oracle and shows these checks reject each one.
"""
import os

import numpy as np
import pytest

import odp_amd as odp
import oracle
from goldens import (TEST_GENERAL, TEST_GRID, GOLDEN, build_case, complete_graph, cycle_graph, cycle_sum,
                     general_fixture, grid_fixture, hypercube, parse_aspl_txt, path_graph, petersen,
                     symmetric_fixtures, two_triangles, vectors, case_nodes)

INT_MAX, LONG_MAX = 2**31 - 1, 2**63 - 1
KINDS = [(oracle.MATRIX, 1), (oracle.MATRIX_SAVING, 1), (oracle.MATRIX_SAVING, 3),
         (oracle.MATRIX_SAVING, 24), (oracle.BFS, 1)]


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()


def golden_general_ok():
    exp = parse_aspl_txt(os.path.join(GOLDEN, "general", "ASPL.txt"))
    for name, e in zip(TEST_GENERAL, exp):
        n, d, adj = general_fixture(name)
        for kind, chunk in KINDS:
            r = oracle.evaluate(adj, nodes=n, kind=kind, chunk=chunk)
            if (r[0], r[1], f"{r[2]:.10f}") != (e["diameter"], e["sum"], e["aspl10"]):
                return False
    return True


def golden_grid_ok():
    exp = parse_aspl_txt(os.path.join(GOLDEN, "grid", "ASPL.txt"))
    for name, e in zip(TEST_GRID, exp):
        w, h, d, _, adj = grid_fixture(name)
        for kind, chunk in KINDS:
            r = oracle.evaluate(adj, width=w, height=h, kind=kind, chunk=chunk)
            if (r[0], r[1], f"{r[2]:.10f}") != (e["diameter"], e["sum"], e["aspl10"]):
                return False
    return True


def symmetric_ok():
    fx = symmetric_fixtures()
    for f in fx["general"]:
        for s in f["symmetries"]:
            n, d, adj = general_fixture(f["file"], s)
            for kind, chunk in KINDS:
                r = oracle.evaluate(adj, nodes=n, symmetries=s, kind=kind, chunk=chunk)
                if (r[0], r[1], repr(r[2])) != (f["diameter"], f["sum"], f["aspl"]):
                    return False
    for f in fx["grid"]:
        for s in f["symmetries"]:
            w, h, d, _, adj = grid_fixture(f["file"], s)
            for kind, chunk in KINDS:
                r = oracle.evaluate(adj, width=w, height=h, symmetries=s, kind=kind, chunk=chunk)
                if (r[0], r[1], repr(r[2])) != (f["diameter"], f["sum"], f["aspl"]):
                    return False
    return True


def closed_forms_ok():
    for kind, chunk in KINDS:
        for n in (2, 5, 64, 65):
            if oracle.evaluate(complete_graph(n), nodes=n, kind=kind, chunk=chunk)[:2] != (1, n * (n - 1) // 2):
                return False
        for n in (3, 10, 64, 129):
            if oracle.evaluate(cycle_graph(n), nodes=n, kind=kind, chunk=chunk)[:2] != (n // 2, cycle_sum(n)):
                return False
        if oracle.evaluate(petersen(), nodes=10, kind=kind, chunk=chunk)[:2] != (2, 75):
            return False
        for k in (3, 6, 9):
            if oracle.evaluate(hypercube(k), nodes=1 << k, kind=kind, chunk=chunk)[:2] != (k, k * 4 ** k // 4):
                return False
        for n in (2, 7, 100):
            adj, nd = path_graph(n)
            r = oracle.evaluate(adj, nodes=n, num_degrees=nd, kind=kind, chunk=chunk)
            if r[:2] != (n - 1, (n - 1) * n * (n + 1) // 6):
                return False
    return True


def vectors_ok(max_nodes=20000):
    for c in vectors():
        if case_nodes(c) > max_nodes:
            continue
        kw, adj = build_case(c)
        r = oracle.evaluate(adj, kind=oracle.MATRIX_SAVING, chunk=8, **kw)
        if not same(r, (c["diameter"], c["sum"], float.fromhex(c["aspl"]))):
            return False
    return True


def test_golden_general():
    assert golden_general_ok()


def test_golden_grid():
    assert golden_grid_ok()


def test_symmetric_fixtures():
    assert symmetric_ok()


def test_closed_forms():
    assert closed_forms_ok()


def test_aspl_double_closed_form():
    # ASPL = sum_ordered / ((n-1) n) in double, sum returned halved (aspl.c:73-74)
    n, d, adj = general_fixture("n256d10")
    dia, s, a = oracle.evaluate(adj, nodes=n)
    assert a == (2 * s) / ((n - 1.0) * n)


def test_disconnected_sentinels():
    for kind, chunk in KINDS:
        assert oracle.evaluate(two_triangles(), nodes=6, kind=kind, chunk=chunk) == (INT_MAX, LONG_MAX,
                                                                                    np.finfo(np.float64).max)


def test_networkx_cross_check():
    nx = pytest.importorskip("networkx")
    for seed in range(4):
        n, d = 60 + 40 * seed, 3 + seed
        if (n * d) % 2:
            n += 1
        adj = odp.generate_regular(n, d, seed)
        g = nx.Graph()
        for i in range(n):
            for j in adj[i]:
                g.add_edge(i, int(j))
        lengths = dict(nx.all_pairs_shortest_path_length(g))
        total = sum(sum(v.values()) for v in lengths.values()) // 2
        diam = max(max(v.values()) for v in lengths.values())
        r = oracle.evaluate(adj, nodes=n)
        assert r[:2] == (diam, total)
        assert r[2] == nx.average_shortest_path_length(g) or abs(r[2] - nx.average_shortest_path_length(g)) < 1e-12


def test_kinds_and_chunks_agree():
    for seed in range(3):
        adj = odp.generate_regular(700, 4 + seed, seed)
        ref = oracle.evaluate(adj, nodes=700, kind=oracle.MATRIX)
        for kind, chunk in KINDS[1:] + [(oracle.MATRIX_SAVING, 11)]:
            assert same(oracle.evaluate(adj, nodes=700, kind=kind, chunk=chunk), ref)


def test_symmetric_equals_expanded():
    for (n, d, s, seed) in ((600, 4, 3, 1), (1024, 6, 4, 2), (960, 5, 6, 3)):
        base = odp.generate_general_s(n, d, s, seed)
        full = oracle.expand(base, nodes=n, symmetries=s)
        want = oracle.evaluate(full, nodes=n)
        for kind, chunk in KINDS:
            assert same(oracle.evaluate(base, nodes=n, symmetries=s, kind=kind, chunk=chunk), want)
    for (w, dg, L, s, seed) in ((20, 4, 3, 4, 1), (24, 6, 4, 2, 2), (16, 3, 2, 2, 3)):
        full = odp.generate_grid(w, w, dg, L, s, seed)
        base = odp.grid_base_table(w, w, dg, s, full)
        assert (oracle.expand(base, width=w, height=w, symmetries=s) == full).all()
        want = oracle.evaluate(full, width=w, height=w)
        for kind, chunk in KINDS:
            assert same(oracle.evaluate(base, width=w, height=w, symmetries=s, kind=kind, chunk=chunk), want)


def test_reference_vectors():
    assert vectors_ok()


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_live_reference():
    cases = [(odp.generate_regular(900, 5, 1), dict(nodes=900)),
             (odp.generate_general_s(800, 4, 4, 2), dict(nodes=800, symmetries=4))]
    full = odp.generate_grid(20, 20, 4, 3, 4, 3)
    cases.append((odp.grid_base_table(20, 20, 4, 4, full), dict(width=20, height=20, symmetries=4)))
    adj, nd = path_graph(50)
    cases.append((adj, dict(nodes=50, num_degrees=nd)))
    for adj, kw in cases:
        for kind in (oracle.MATRIX, oracle.MATRIX_SAVING, oracle.BFS):
            ref, _ = oracle.run_reference(adj, kind=kind, **kw)
            assert same(oracle.evaluate(adj, kind=kind, **kw), ref), (kw, kind)


FAULTS = {1: "drop the symmetry factor in MATRIX_SAVING", 2: "diameter off by one",
          3: "swap 90/270 rotations", 4: "wrong orbit shift", 5: "drop the last neighbour",
          6: "ASPL denominator n*n"}


@pytest.mark.parametrize("code", sorted(FAULTS))
def test_planted_faults_are_caught(code):
    oracle.set_fault(code)
    try:
        caught = not (golden_general_ok() and golden_grid_ok() and symmetric_ok() and closed_forms_ok()
                      and vectors_ok(max_nodes=5000))
    finally:
        oracle.set_fault(0)
    assert caught, FAULTS[code]
