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


# ---- one rank's share of the MPI+CUDA flavour --------------------------------

def _bfs_dists(full, sources):
    """Hop distances from each source (rows) to every vertex: breadth-first
    search, scipy's compiled unweighted shortest paths."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import shortest_path
    n = full.shape[0]
    rows = np.repeat(np.arange(n), full.shape[1])
    cols = full.ravel()
    keep = cols >= 0
    g = csr_matrix((np.ones(keep.sum()), (rows[keep], cols[keep])), shape=(n, n))
    return shortest_path(g, method="D", unweighted=True, indices=sources).astype(np.int64)


def _ref_source_vertices(nodes=None, width=None, height=None, symmetries=1):
    """Vertex of each source column (row of the bit matrix below n/s): the row
    itself, or for symmetric grids the base vertex the reference's itable puts
    there (utils.c:2103-2124)."""
    if width is None or symmetries == 1:
        return list(range((nodes or width * height) // symmetries))
    based = width * height // symmetries
    if symmetries == 2:
        return list(range(based))
    bh = height // 2
    return [(i // bh) * height + (i % bh) for i in range(based)]


def _rank_sources(based, kind, rank, procs):
    W = (based + 63) // 64
    if kind == oracle.MATRIX:
        chunk = (W + procs - 1) // procs
        blocks = [rank] if rank < (W + chunk - 1) // chunk else []
    else:
        chunk = 64
        blocks = list(range(rank, (W + chunk - 1) // chunk, procs))
    return [c for t in blocks for c in range(64 * chunk * t, min(based, 64 * chunk * (t + 1)))]


def shard_brute_ok(adj, procs_list=(1, 2, 3, 5), **kw):
    """Each rank's partial equals breadth-first searches from its own sources,
    and the partials combine to the whole evaluation."""
    s = kw.get("symmetries", 1)
    n = kw.get("nodes") or kw["width"] * kw["height"]
    full = oracle.expand(adj, **kw) if s > 1 else adj
    verts = _ref_source_vertices(**kw)
    dist = _bfs_dists(full, verts)
    whole = oracle.evaluate(adj, **kw)
    for kind in (oracle.MATRIX, oracle.MATRIX_SAVING):
        for procs in procs_list:
            tot, dmax = 0, 1
            for rank in range(procs):
                ex, dia = oracle.evaluate_shard(adj, rank, procs, kind=kind, **kw)
                src = _rank_sources(len(verts), kind, rank, procs)
                want_ex = s * sum(int(dist[c].sum()) - (n - 1) for c in src)
                want_d = max([1] + [int(dist[c].max()) for c in src])
                if (ex, dia) != (want_ex, want_d):
                    return False
                tot += ex
                dmax = max(dmax, dia)
            ordered = tot + n * (n - 1)
            if (dmax, ordered // 2, ordered / ((n - 1.0) * n)) != (whole[0], whole[1], whole[2]):
                return False
    return True


def test_shard_partials_bfs_and_combine():
    """oracle.evaluate_shard against breadth-first searches from each rank's
    sources, for both kinds and 1, 2, 3, 5 ranks, on a general graph of two
    4096-source blocks, a circulant-symmetric graph and 2- and 4-fold
    symmetric grids; the partials also combine to the whole evaluation."""
    adj = odp.generate_regular(4200, 3, 4)
    assert shard_brute_ok(adj, nodes=4200)
    base = odp.generate_general_s(1200, 4, 4, 6)
    assert shard_brute_ok(base, nodes=1200, symmetries=4)
    for s in (2, 4):
        full = odp.generate_grid(24, 24, 6, 3, s, 7)
        gb = odp.grid_base_table(24, 24, 6, s, full)
        assert shard_brute_ok(gb, width=24, height=24, symmetries=s)


def test_shard_planted_fault_is_caught():
    oracle.set_fault(7)           # MATRIX shards floor(W / procs) words wide
    try:
        assert not shard_brute_ok(odp.generate_regular(4200, 3, 4), procs_list=(3, 5), nodes=4200)
    finally:
        oracle.set_fault(0)
