"""The BFS kind (ODP_ASPL=BFS): one breadth-first search per source on the GPU
(k_bfs), the reference's CPU-only aspl_bfs (aspl.c:166-221).  Its results must
equal the reference's golden outputs, the oracle's BFS restatement and the
reference library's values, bit for bit, on every graph family."""
import os
import sys

import numpy as np
import pytest

import odp_amd as odp
import oracle
from goldens import (TEST_GENERAL, TEST_GRID, GOLDEN, build_case, case_nodes, complete_graph, cycle_graph,
                     cycle_sum, general_fixture, grid_fixture, hypercube, parse_aspl_txt, path_graph, petersen,
                     symmetric_fixtures, two_triangles, vectors)

pytestmark = pytest.mark.gpu

INT_MAX, LONG_MAX, DBL_MAX = 2**31 - 1, 2**63 - 1, sys.float_info.max


@pytest.fixture(autouse=True)
def bfs_kind(monkeypatch):
    monkeypatch.setenv("ODP_ASPL", "BFS")


def init(adj, nodes=None, width=None, height=None, symmetries=1, num_degrees=None):
    if width is None:
        if symmetries == 1:
            odp.init_aspl_general(nodes, adj.shape[1], num_degrees)
        else:
            odp.init_aspl_general_s(nodes, adj.shape[1], num_degrees, symmetries)
    elif symmetries == 1:
        odp.init_aspl_grid(width, height, adj.shape[1], num_degrees)
    else:
        odp.init_aspl_grid_s(width, height, adj.shape[1], num_degrees, symmetries)


def bfs_eval(adj, **kw):
    init(adj, **kw)
    try:
        return odp.set_aspl(adj)
    finally:
        odp.finalize_aspl()


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()


@pytest.mark.parametrize("idx", range(len(TEST_GENERAL)))
def test_golden_general(idx):
    exp = parse_aspl_txt(os.path.join(GOLDEN, "general", "ASPL.txt"))[idx]
    n, d, adj = general_fixture(TEST_GENERAL[idx])
    got = bfs_eval(adj, nodes=n)
    assert got[0] == exp["diameter"] and got[1] == exp["sum"] and f"{got[2]:.10f}" == exp["aspl10"]
    assert same(got, oracle.evaluate(adj, nodes=n, kind=oracle.BFS))


@pytest.mark.parametrize("idx", range(len(TEST_GRID)))
def test_golden_grid(idx):
    exp = parse_aspl_txt(os.path.join(GOLDEN, "grid", "ASPL.txt"))[idx]
    w, h, d, _, adj = grid_fixture(TEST_GRID[idx])
    got = bfs_eval(adj, width=w, height=h)
    assert got[0] == exp["diameter"] and got[1] == exp["sum"] and f"{got[2]:.10f}" == exp["aspl10"]
    assert same(got, oracle.evaluate(adj, width=w, height=h, kind=oracle.BFS))


def test_symmetric_fixtures():
    """The values the reference printed for its symmetric fixtures (general s
    and 2-/4-fold grids): the sources are one row per orbit."""
    fx = symmetric_fixtures()
    for f in fx["general"]:
        for s in f["symmetries"]:
            n, d, adj = general_fixture(f["file"], s)
            got = bfs_eval(adj, nodes=n, symmetries=s)
            assert (got[0], got[1], repr(got[2])) == (f["diameter"], f["sum"], f["aspl"]), (f["file"], s)
    for f in fx["grid"]:
        for s in f["symmetries"]:
            w, h, d, _, adj = grid_fixture(f["file"], s)
            got = bfs_eval(adj, width=w, height=h, symmetries=s)
            assert (got[0], got[1], repr(got[2])) == (f["diameter"], f["sum"], f["aspl"]), (f["file"], s)


def test_closed_forms_and_sentinels():
    for n in (2, 5, 33, 130):
        assert bfs_eval(complete_graph(n), nodes=n)[:2] == (1, n * (n - 1) // 2)
    for n in (3, 10, 129, 300):
        assert bfs_eval(cycle_graph(n), nodes=n)[:2] == (n // 2, cycle_sum(n))
    assert bfs_eval(petersen(), nodes=10)[:2] == (2, 75)
    for k in (3, 7, 10):
        assert bfs_eval(hypercube(k), nodes=1 << k)[:2] == (k, k * 4 ** k // 4)
    for n in (2, 7, 200):
        adj, nd = path_graph(n)
        assert bfs_eval(adj, nodes=n, num_degrees=nd)[:2] == (n - 1, (n - 1) * n * (n + 1) // 6)
    assert bfs_eval(two_triangles(), nodes=6) == (INT_MAX, LONG_MAX, DBL_MAX)
    adj = cycle_graph(40)                       # an isolated vertex through num_degrees
    nd = np.full(40, 2, np.int32)
    adj[3] = [3, 3]
    nd[3] = 0
    adj[2, 1] = 1
    adj[4, 0] = 5
    assert bfs_eval(adj, nodes=40, num_degrees=nd) == (INT_MAX, LONG_MAX, DBL_MAX)


@pytest.mark.parametrize("seed", range(3))
def test_random_vs_oracle(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 1500)) * 2
    d = int(rng.integers(3, 12))
    adj = odp.generate_regular(n, d, 300 + seed)
    assert same(bfs_eval(adj, nodes=n), oracle.evaluate(adj, nodes=n, kind=oracle.BFS))
    base = odp.generate_general_s(1024, 6, 4, 7 + seed)
    assert same(bfs_eval(base, nodes=1024, symmetries=4),
                oracle.evaluate(base, nodes=1024, symmetries=4, kind=oracle.BFS))


@pytest.mark.parametrize("case", vectors(), ids=lambda c: "-".join(str(v) for k, v in c.items()
                                                                 if k not in ("diameter", "sum", "aspl")))
def test_reference_vectors(case):
    """Every reference vector, the n=65536 d=16 graph and the 256x256 4-fold
    grid (diameter 128) included."""
    kw, adj = build_case(case)
    print(f"bfs vector n={case_nodes(case)}", flush=True)
    want = (case["diameter"], case["sum"], float.fromhex(case["aspl"]))
    assert same(bfs_eval(adj, **kw), want)


def test_device_swaps_and_profile(capfd, monkeypatch):
    """Device 2-opt swaps between evaluations under the BFS kind; the profile
    names the kind."""
    monkeypatch.setenv("ODP_PROFILE", "1")
    n, d = 2000, 6
    adj = odp.generate_regular(n, d, 5)
    odp.init_aspl_general(n, d)
    try:
        first = odp.set_aspl(adj)
        assert same(first, oracle.evaluate(adj, nodes=n, kind=oracle.BFS))
        rng = odp.Rng(11)
        cur = adj.copy()
        for _ in range(4):
            a, b, c, e = odp.two_opt(cur, rng)          # applied to the host copy
            assert odp.swap_device(a, b, c, e) == 0
            assert same(odp.set_aspl_resident(), oracle.evaluate(cur, nodes=n, kind=oracle.BFS))
        with pytest.raises(ValueError):
            odp.set_aspl_shard(0, 2)                # the MPI flavours divide bit-matrix columns only
        assert odp.autotune(cur) == -1              # the kind runs no plan
    finally:
        odp.finalize_aspl()
    out = capfd.readouterr().out
    assert "Algorithm       = BFS (HIP)" in out
