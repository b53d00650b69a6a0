"""HIP path vs the oracle and the reference's golden outputs, through the C ABI.

Every test compares integers exactly and the ASPL double bit-for-bit.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import odp_amd as odp
import oracle
from goldens import (TEST_GENERAL, TEST_GRID, GOLDEN, build_case, complete_graph, cycle_graph, cycle_sum,
                     general_fixture, grid_fixture, hypercube, parse_aspl_txt, path_graph, petersen,
                     symmetric_fixtures, two_triangles, vectors, case_nodes)

pytestmark = pytest.mark.gpu

INT_MAX, LONG_MAX, DBL_MAX = 2**31 - 1, 2**63 - 1, sys.float_info.max

RES = odp.PLAN_RESIDENT
LEV = odp.PLAN_LEVEL
# a spread of plans; each must return identical results
PLANS = [
    None,
    dict(plan=RES, wb=1, threads=256),
    dict(plan=RES, wb=2, threads=1024),
    dict(plan=RES, wb=8, threads=512),
    dict(plan=LEV, kw=0, upw=4, seg_major=1, xcd=0),
    dict(plan=LEV, kw=1, upw=1, seg_major=0, xcd=1),
    dict(plan=LEV, kw=3, upw=2, seg_major=1, xcd=1),
    dict(plan=LEV, kw=64, upw=8, seg_major=0, xcd=0),
    dict(plan=LEV, kw=2, upw=2, seg_major=0, xcd=1, group=2),
    dict(plan=LEV, kw=1, upw=16, seg_major=1, xcd=0, group=1),
    dict(plan=RES, wb=4, threads=1024),
    dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=-1),
    dict(plan=LEV, kw=2, upw=2, seg_major=1, xcd=1, direct=2),
    dict(plan=LEV, kw=4, upw=4, seg_major=0, xcd=1, group=2, direct=3),
    dict(plan=LEV, kw=128, upw=8, seg_major=0, xcd=1, direct=6),
    dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=-1, parts=4),
    dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=0, parts=2, group=3),
]


LDS_DYN_MAX = 163840 - 512          # LDS one resident workgroup may hold for its block


def plan_fits(plan, n):
    """Whether the library must accept `plan` for a graph of n vertices: a
    resident block of 32*wb sources needs n*wb*4 bytes of LDS, 16-bit ids
    (n <= 65535) and at most res_rpt(wb) rows per thread; level plans always fit."""
    if plan is None or plan["plan"] != RES:
        return True
    wb, th = plan["wb"], plan["threads"]
    rpt = {1: 40, 2: 20, 4: 10, 8: 2}[wb]
    return n <= 65535 and n * wb * 4 <= LDS_DYN_MAX and n <= rpt * th


def hip_eval(adj, nodes=None, width=None, height=None, symmetries=1, num_degrees=None, plan=None):
    """One Init / Set_aspl / Finalize cycle; None when `plan` does not fit the
    graph, after checking that the library refuses exactly those plans."""
    if width is None:
        if symmetries == 1:
            odp.init_aspl_general(nodes, adj.shape[1], num_degrees)
        else:
            odp.init_aspl_general_s(nodes, adj.shape[1], num_degrees, symmetries)
    else:
        if symmetries == 1:
            odp.init_aspl_grid(width, height, adj.shape[1], num_degrees)
        else:
            odp.init_aspl_grid_s(width, height, adj.shape[1], num_degrees, symmetries)
    try:
        if plan is not None:
            n = nodes if width is None else width * height
            if not plan_fits(plan, n):
                with pytest.raises(ValueError):
                    odp.set_plan(**plan)
                return None
            odp.set_plan(**plan)
        return odp.set_aspl(adj)
    finally:
        odp.finalize_aspl()


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()


@pytest.mark.parametrize("idx", range(len(TEST_GENERAL)))
@pytest.mark.parametrize("plan", [None, PLANS[4], PLANS[1]])
def test_golden_general(idx, plan):
    exp = parse_aspl_txt(os.path.join(GOLDEN, "general", "ASPL.txt"))[idx]
    n, d, adj = general_fixture(TEST_GENERAL[idx])
    got = hip_eval(adj, nodes=n, plan=plan)
    assert got[0] == exp["diameter"] and got[1] == exp["sum"] and f"{got[2]:.10f}" == exp["aspl10"]
    assert same(got, oracle.evaluate(adj, nodes=n))


@pytest.mark.parametrize("idx", range(len(TEST_GRID)))
@pytest.mark.parametrize("plan", [None, PLANS[4]])
def test_golden_grid(idx, plan):
    exp = parse_aspl_txt(os.path.join(GOLDEN, "grid", "ASPL.txt"))[idx]
    w, h, d, _, adj = grid_fixture(TEST_GRID[idx])
    got = hip_eval(adj, width=w, height=h, plan=plan)
    assert got[0] == exp["diameter"] and got[1] == exp["sum"] and f"{got[2]:.10f}" == exp["aspl10"]


@pytest.mark.parametrize("plan", [None, PLANS[4], PLANS[5]])
def test_symmetric_fixtures(plan):
    fx = symmetric_fixtures()
    for f in fx["general"]:
        for s in f["symmetries"]:
            n, d, adj = general_fixture(f["file"], s)
            got = hip_eval(adj, nodes=n, symmetries=s, plan=plan)
            assert (got[0], got[1], repr(got[2])) == (f["diameter"], f["sum"], f["aspl"]), (f["file"], s)
    for f in fx["grid"]:
        for s in f["symmetries"]:
            w, h, d, _, adj = grid_fixture(f["file"], s)
            got = hip_eval(adj, width=w, height=h, symmetries=s, plan=plan)
            assert (got[0], got[1], repr(got[2])) == (f["diameter"], f["sum"], f["aspl"]), (f["file"], s)


@pytest.mark.parametrize("plan", PLANS)
def test_closed_forms(plan):
    for n in (2, 5, 33, 64, 65, 130):
        got = hip_eval(complete_graph(n), nodes=n, plan=plan)
        if got is not None:
            assert got[:2] == (1, n * (n - 1) // 2)
    for n in (3, 10, 64, 129, 300):
        got = hip_eval(cycle_graph(n), nodes=n, plan=plan)
        if got is not None:
            assert got[:2] == (n // 2, cycle_sum(n))
    got = hip_eval(petersen(), nodes=10, plan=plan)
    assert got is None or got[:2] == (2, 75)
    for k in (3, 7, 10):
        got = hip_eval(hypercube(k), nodes=1 << k, plan=plan)
        if got is not None:
            assert got[:2] == (k, k * 4 ** k // 4)
    for n in (2, 7, 100, 200):
        adj, nd = path_graph(n)
        got = hip_eval(adj, nodes=n, num_degrees=nd, plan=plan)
        if got is not None:
            assert got[:2] == (n - 1, (n - 1) * n * (n + 1) // 6)


@pytest.mark.parametrize("plan", [None, PLANS[4], PLANS[2]])
def test_disconnected_sentinels(plan):
    got = hip_eval(two_triangles(), nodes=6, plan=plan)
    assert got == (INT_MAX, LONG_MAX, DBL_MAX)
    # isolated vertex via num_degrees
    adj = cycle_graph(40)
    nd = np.full(40, 2, np.int32)
    adj[3] = [3, 3]
    nd[3] = 0
    adj[2, 1] = 1
    adj[4, 0] = 5
    got = hip_eval(adj, nodes=40, num_degrees=nd, plan=plan)
    assert got == (INT_MAX, LONG_MAX, DBL_MAX)


SMALL_VECTORS = [c for c in vectors() if case_nodes(c) <= 30000]


@pytest.mark.parametrize("case", SMALL_VECTORS, ids=lambda c: "-".join(str(v) for k, v in c.items()
                                                                      if k not in ("diameter", "sum", "aspl")))
def test_reference_vectors(case):
    kw, adj = build_case(case)
    want = (case["diameter"], case["sum"], float.fromhex(case["aspl"]))
    for plan in (None, PLANS[1], PLANS[4], PLANS[6], PLANS[8]):
        got = hip_eval(adj, plan=plan, **kw)
        if got is not None:
            assert same(got, want), (plan, got, want)


@pytest.mark.parametrize("seed", range(6))
def test_random_vs_oracle(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 1500)) * 2
    d = int(rng.integers(3, 12))
    adj = odp.generate_regular(n, d, 100 + seed)
    want = oracle.evaluate(adj, nodes=n)
    for plan in PLANS:
        got = hip_eval(adj, nodes=n, plan=plan)
        if got is not None:
            assert same(got, want), (plan, got, want)


def test_symmetric_generated_vs_expanded():
    for (n, d, s, seed) in ((600, 4, 3, 1), (1024, 6, 4, 2), (2048, 5, 8, 3)):
        base = odp.generate_general_s(n, d, s, seed)
        full = oracle.expand(base, nodes=n, symmetries=s)
        want = oracle.evaluate(full, nodes=n)
        for plan in (None, PLANS[4], PLANS[5]):
            got = hip_eval(base, nodes=n, symmetries=s, plan=plan)
            assert same(got, want), (n, s, plan)
    for (w, dg, L, s, seed) in ((20, 4, 3, 4, 1), (24, 6, 4, 2, 2), (40, 8, 4, 4, 3)):
        full = odp.generate_grid(w, w, dg, L, s, seed)
        base = odp.grid_base_table(w, w, dg, s, full)
        want = oracle.evaluate(full, width=w, height=w)
        for plan in (None, PLANS[4], PLANS[5], PLANS[3]):
            got = hip_eval(base, width=w, height=w, symmetries=s, plan=plan)
            if got is not None:
                assert same(got, want), (w, s, plan)


def test_symmetric_resident_past_16384_rows():
    """Symmetric graphs past 16384 rows under the default plan, now resident in
    the wide kernel instances: a 4-fold 144x144 grid (20736 cells) and a 4-fold
    general graph of 36000 vertices, against the oracle on the expanded graphs."""
    w = 144
    full = odp.generate_grid(w, w, 8, 4, 4, 11)
    base = odp.grid_base_table(w, w, 8, 4, full)
    got = hip_eval(base, width=w, height=w, symmetries=4)
    assert same(got, oracle.evaluate(full, width=w, height=w))
    n = 36000
    base = odp.generate_general_s(n, 8, 4, 12)
    full = oracle.expand(base, nodes=n, symmetries=4)
    got = hip_eval(base, nodes=n, symmetries=4)
    assert same(got, oracle.evaluate(full, nodes=n))


def test_nonregular_symmetric_num_degrees():
    # sample/graph/grid/w10h10d4r4g4 with a vertex-orbit trimmed through num_degrees
    w, h, d, _, base = grid_fixture("w10h10d4r4g4", 4)
    nd = np.full(base.shape[0], d, np.int32)
    full = oracle.expand(base, width=w, height=h, symmetries=4, num_degrees=nd)
    want = oracle.evaluate(base, width=w, height=h, symmetries=4, num_degrees=nd)
    assert same(want, oracle.evaluate(full, width=w, height=h))
    got = hip_eval(base, width=w, height=h, symmetries=4, num_degrees=nd)
    assert same(got, want)


def test_repeated_lifecycle():
    adj = odp.generate_regular(500, 5, 3)
    want = oracle.evaluate(adj, nodes=500)
    for _ in range(5):
        assert same(hip_eval(adj, nodes=500), want)
        w, h, d, _, g = grid_fixture("w6h6d3r2g4", 4)
        assert hip_eval(g, width=w, height=h, symmetries=4)[:2] == (5, 1978)


def test_repeat_calls_same_init():
    n, d = 3000, 7
    adj = odp.generate_regular(n, d, 9) if (n * d) % 2 == 0 else None
    n, d = 3000, 8
    adj = odp.generate_regular(n, d, 9)
    rng = odp.Rng(5)
    for plan in (None, PLANS[4]):
        odp.init_aspl_general(n, d)
        if plan:
            odp.set_plan(**plan)
        a = adj.copy()
        for _ in range(4):
            odp.two_opt(a, rng)
            assert same(odp.set_aspl(a), oracle.evaluate(a, nodes=n))
        odp.finalize_aspl()


def test_device_swap_matches_host():
    n, d = 2000, 6
    adj = odp.generate_regular(n, d, 21)
    rng = odp.Rng(77)
    for plan in (None, PLANS[4]):
        odp.init_aspl_general(n, d)
        if plan:
            odp.set_plan(**plan)
        host = adj.copy()
        odp.upload_adjacency(host)
        for _ in range(5):
            a, b, c, dd = odp.two_opt(host, rng)
            assert odp.swap_device(a, b, c, dd) == 0
            assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
        assert odp.swap_device(0, 0, 0, 0) == -1      # missing edge is rejected
        # half valid: a-b is an edge, c-d is not -- nothing may change
        a, b = 5, int(host[5][0])
        c = next(v for v in range(n) if v not in (a, b) and v not in host[a] and v not in host[b])
        dd = next(v for v in range(n) if v not in (a, b, c) and v not in host[c])
        assert odp.swap_device(a, b, c, dd) == -1
        # ids that are not distinct (b == c, a == d) are not a 2-opt move
        c2 = int(host[b][0]) if int(host[b][0]) != a else int(host[b][1])
        assert odp.swap_device(a, b, b, c2) == -1
        assert odp.swap_device(a, b, int(host[c][0]), a) == -1
        assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
        # and the graph still takes valid moves afterwards
        a, b, c, dd = odp.two_opt(host, rng)
        assert odp.swap_device(a, b, c, dd) == 0
        assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
        odp.finalize_aspl()


def test_resident_table_follows_swaps_and_uploads():
    """The resident plan rebuilds its 16-bit neighbour table only when the
    adjacency changed by other means than device swaps (a swap patches the table
    as well): evaluations after swaps, after an upload of another graph, after
    ODP_Set_aspl and after a plan change must all equal the oracle."""
    n, d = 1500, 6
    g1 = odp.generate_regular(n, d, 41)
    g2 = odp.generate_regular(n, d, 42)
    rng = odp.Rng(5)
    odp.init_aspl_general(n, d)
    try:
        for plan in (PLANS[1], PLANS[10]):
            odp.set_plan(**plan)
            host = g1.copy()
            odp.upload_adjacency(host)
            assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
            for _ in range(4):
                a, b, c, dd = odp.two_opt(host, rng)
                assert odp.swap_device(a, b, c, dd) == 0
                assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
            odp.upload_adjacency(g2)
            assert same(odp.set_aspl_resident(), oracle.evaluate(g2, nodes=n))
            assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))
            assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
            a, b, c, dd = odp.two_opt(host, rng)
            assert odp.swap_device(a, b, c, dd) == 0
            assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
    finally:
        odp.finalize_aspl()


@pytest.mark.parametrize("plan", [
    dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=-1, parts=4),
    dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=0, parts=2, group=2),
    dict(plan=LEV, kw=4, upw=4, seg_major=0, xcd=1, direct=3),
])
def test_level_table_follows_swaps(plan):
    """The level plan keeps its row-neighbour tables (32-bit, and the 16-bit
    copy k_direct reads) and the row-part pair lists between evaluations: a
    device swap patches the tables and rebuilds the pair lists of the units
    (1024 rows) its four rows lie in; an upload, ODP_Set_aspl or a plan change
    makes the next evaluation rebuild them.  Every evaluation must equal the
    oracle, rejected moves included (they change nothing)."""
    n, d = 3000, 12                                   # 3 units; 47 words: 2 full 16-word blocks + 15
    g1 = odp.generate_regular(n, d, 91)
    g2 = odp.generate_regular(n, d, 92)
    rng = odp.Rng(17)
    odp.init_aspl_general(n, d)
    try:
        odp.set_plan(**plan)
        host = g1.copy()
        odp.upload_adjacency(host)
        for _ in range(3):                            # frontier first, then the plain levels
            assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
        for i in range(8):
            a, b, c, dd = odp.two_opt(host, rng)
            assert odp.swap_device(a, b, c, dd) == 0
            if i % 3 == 2:                            # two moves between evaluations
                a, b, c, dd = odp.two_opt(host, rng)
                assert odp.swap_device(a, b, c, dd) == 0
            assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n)), i
        assert odp.swap_device(0, 0, 0, 0) == -1      # not a move: nothing changes
        a, b = 5, int(host[5][0])                     # a-b an edge, c-d not: nothing changes
        c = next(v for v in range(n) if v not in (a, b) and v not in host[a] and v not in host[b])
        dd = next(v for v in range(n) if v not in (a, b, c) and v not in host[c])
        assert odp.swap_device(a, b, c, dd) == -1
        assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
        odp.upload_adjacency(g2)
        assert same(odp.set_aspl_resident(), oracle.evaluate(g2, nodes=n))
        assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))
        for _ in range(3):
            a, b, c, dd = odp.two_opt(host, rng)
            assert odp.swap_device(a, b, c, dd) == 0
            assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
        odp.set_plan(**plan)                          # a new plan rebuilds everything
        assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
        a, b, c, dd = odp.two_opt(host, rng)
        assert odp.swap_device(a, b, c, dd) == 0
        assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
    finally:
        odp.finalize_aspl()


def test_long_level_sequences_end_their_batches():
    """A batch of levels ends with the counters handed to host memory by a
    kernel that then writes a sequence number (small batches), or with copies
    and a stream synchronisation (batches of more than 4096 counters, and
    timed passes).  A 20000-cycle (diameter 10000) under a 1-word plan (313
    blocks, batches growing to 256 levels) takes both; a path and an
    unchanged-graph repeat check the sequence numbers stay in step."""
    n = 20000
    adj = cycle_graph(n)
    want = (n // 2, cycle_sum(n))
    for plan in (dict(plan=LEV, kw=1, upw=16, seg_major=0, xcd=1), dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1)):
        odp.init_aspl_general(n, 2)
        try:
            odp.set_plan(**plan)
            for _ in range(3):
                got = odp.set_aspl(adj)
                assert got[:2] == want, (plan, got)
            odp.stats_enable(True)                    # timed: the copies and a synchronisation
            assert odp.set_aspl(adj)[:2] == want
            odp.stats_enable(False)
            assert odp.set_aspl(adj)[:2] == want
        finally:
            odp.finalize_aspl()
    path, nd = path_graph(3000)
    odp.init_aspl_general(3000, path.shape[1], nd)
    try:
        for _ in range(2):
            assert odp.set_aspl(path)[:2] == (2999, 2999 * 3000 * 3001 // 6)
    finally:
        odp.finalize_aspl()


def test_drop_in_small_changes_as_patches():
    """ODP_Set_aspl on an array that differs from the last upload in a few
    entries (the annealing pattern: a move, then Set_aspl on the same array)
    queues those entries for the next launch instead of rebuilding the table.
    Moves, restores, an unchanged array, two moves at once, more changes than
    the launch takes, a device swap in between (the staging copy no longer
    mirrors the device) and a plan change with entries pending must all give
    the oracle's result."""
    n, d = 600, 4
    rng = odp.Rng(4)
    host = odp.generate_regular(n, d, 61)
    odp.init_aspl_general(n, d)
    try:
        odp.set_plan(**PLANS[1])
        assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))
        for step in range(12):
            saved = host.copy()
            odp.two_opt(host, rng)                       # four entries change
            assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n)), step
            if step % 3 == 2:
                host = saved                             # and back
                assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n)), step
        assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))      # unchanged
        odp.two_opt(host, rng)
        odp.two_opt(host, rng)
        assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))      # up to eight entries
        for _ in range(3):
            odp.two_opt(host, rng)
        assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))      # more: full upload
        dev = host.copy()
        a, b, c, dd = odp.two_opt(dev, rng)
        assert odp.swap_device(a, b, c, dd) == 0
        assert same(odp.set_aspl_resident(), oracle.evaluate(dev, nodes=n))
        assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))      # device had moved on
        h2 = host.copy()
        odp.two_opt(h2, rng)
        odp.upload_adjacency(h2)                         # queued as entries
        odp.set_plan(**PLANS[4])                         # stores them before the level plan reads
        assert same(odp.set_aspl_resident(), oracle.evaluate(h2, nodes=n))
    finally:
        odp.finalize_aspl()


@pytest.mark.parametrize("plan", [
    dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=-1, parts=4),
    dict(plan=LEV, kw=4, upw=4, seg_major=0, xcd=1, direct=3),
])
def test_level_plan_small_changes_as_patches(plan):
    """The level plan takes ODP_Set_aspl on an array that differs from the last
    upload in at most 8 entries as a patch: the entries, the row-neighbour
    tables and the row-part lists of their units are updated in one launch, no
    table is sent or rebuilt.  Moves, restores, an unchanged array, two moves
    at once, three (a full upload), a device swap in between (the staging copy
    no longer mirrors the device) and moves after it must all give the
    oracle's result."""
    n, d = 3000, 6
    rng = odp.Rng(8)
    host = odp.generate_regular(n, d, 63)
    odp.init_aspl_general(n, d)
    try:
        odp.set_plan(**plan)
        for _ in range(2):                               # frontier first, then the plain levels
            assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))
        for step in range(10):
            saved = host.copy()
            odp.two_opt(host, rng)                       # four entries change
            assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n)), step
            if step % 3 == 2:
                host = saved                             # and back
                assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n)), step
        assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))      # unchanged
        odp.two_opt(host, rng)
        odp.two_opt(host, rng)
        assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))      # up to eight entries
        for _ in range(3):
            odp.two_opt(host, rng)
        assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))      # more: full upload
        dev = host.copy()
        a, b, c, dd = odp.two_opt(dev, rng)
        assert odp.swap_device(a, b, c, dd) == 0
        assert same(odp.set_aspl_resident(), oracle.evaluate(dev, nodes=n))
        assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n))      # device had moved on
        for step in range(3):
            odp.two_opt(host, rng)
            assert same(odp.set_aspl(host), oracle.evaluate(host, nodes=n)), step
    finally:
        odp.finalize_aspl()


def test_small_upload_read_by_table_build():
    """Under the resident plan a small table reaches the device through the
    table build reading pinned host memory (no DMA); two uploads in a row, a
    plan change, a device swap and the drop-in path must all see the last
    upload."""
    n, d = 800, 4
    g1 = odp.generate_regular(n, d, 51)
    g2 = odp.generate_regular(n, d, 52)
    rng = odp.Rng(8)
    odp.init_aspl_general(n, d)
    try:
        odp.set_plan(**PLANS[1])
        odp.upload_adjacency(g1)
        odp.upload_adjacency(g2)
        assert same(odp.set_aspl_resident(), oracle.evaluate(g2, nodes=n))
        odp.upload_adjacency(g1)
        odp.set_plan(**PLANS[4])                      # level plan: the upload is copied first
        assert same(odp.set_aspl_resident(), oracle.evaluate(g1, nodes=n))
        odp.set_plan(**PLANS[1])
        host = g2.copy()
        odp.upload_adjacency(host)
        a, b, c, dd = odp.two_opt(host, rng)
        assert odp.swap_device(a, b, c, dd) == 0      # swaps the uploaded table, not a stale one
        assert same(odp.set_aspl_resident(), oracle.evaluate(host, nodes=n))
        assert same(odp.set_aspl(g1), oracle.evaluate(g1, nodes=n))
        assert same(odp.set_aspl_resident(), oracle.evaluate(g1, nodes=n))
        assert same(odp.set_aspl(g2), oracle.evaluate(g2, nodes=n))
    finally:
        odp.finalize_aspl()


def test_autotune_consistent():
    n, d = 4096, 8
    adj = odp.generate_regular(n, d, 5)
    want = oracle.evaluate(adj, nodes=n)
    odp.init_aspl_general(n, d)
    tried = odp.autotune(adj, repeats=1)      # exits the process if two plans disagree
    assert tried > 4
    assert same(odp.set_aspl(adj), want)
    odp.finalize_aspl()


def _batch_graphs(seed, count, nmax=3000):
    rng = np.random.default_rng(seed)
    graphs = []
    for g in range(count):
        n = int(rng.integers(8, nmax // 2)) * 2
        d = int(rng.choice([3, 4, 8, 12, 16, 24]))
        d = min(d, n - 1)
        if (n * d) % 2:
            d -= 1
        adj = odp.generate_regular(n, d, 1000 * seed + g)
        odp.two_opt(adj, odp.Rng(g))
        graphs.append(adj)
    return graphs


def test_batch_vs_oracle():
    graphs = _batch_graphs(1, 48)
    graphs.append(two_triangles())                       # a disconnected one in the mix
    graphs.append(complete_graph(40))
    packed = odp.pack(graphs)
    d, s, a = odp.batch_eval(packed)
    for g, adj in enumerate(graphs):
        want = oracle.evaluate(adj, nodes=adj.shape[0])
        assert same((int(d[g]), int(s[g]), float(a[g])), want), g
    d2, s2, a2 = odp.loop_eval(packed, odp.PLAN_LEVEL)
    assert (d2 == d).all() and (s2 == s).all() and (a2.view(np.int64) == a.view(np.int64)).all()
    d3, s3, a3 = odp.loop_eval(packed, odp.PLAN_AUTO)
    assert (d3 == d).all() and (s3 == s).all()
    ng = odp.batch_load(packed)
    d4, s4, a4 = odp.batch_run(ng)
    assert (d4 == d).all() and (s4 == s).all() and (a4.view(np.int64) == a.view(np.int64)).all()
    # alone vs in the mix
    for g in (0, 17, 48):
        one = odp.batch_eval(odp.pack([graphs[g]]))
        assert one[0][0] == d[g] and one[1][0] == s[g] and one[2][0] == a[g]


@pytest.mark.parametrize("kw", [1, 4, 16])
def test_frontier_and_dense_levels_agree(kw):
    """The first evaluation after Init runs the frontier-mode level kernel; later
    ones run it only for graphs that needed more than 16 levels, the plain one
    otherwise.  Every evaluation equals the oracle: small and large diameters,
    a non-regular table, a directed table, a disconnected graph."""
    cases = [(odp.generate_regular(3000, 5, 21), None), (cycle_graph(600), None)]
    pa, nd = path_graph(257)
    cases.append((pa, nd))
    # a directed table (steps +1 and +2 around a cycle): not symmetric, so the
    # frontier levels test neighbours' change flags instead of marks
    cases.append((np.array([[(i + 1) % 300, (i + 2) % 300] for i in range(300)], np.int32), None))
    cases.append((np.concatenate([cycle_graph(300), cycle_graph(200) + 300]).astype(np.int32), None))
    for adj, nd in cases:
        n = adj.shape[0]
        want = oracle.evaluate(adj, nodes=n, num_degrees=nd) if nd is not None else oracle.evaluate(adj, nodes=n)
        odp.init_aspl_general(n, adj.shape[1], nd)
        try:
            odp.set_plan(plan=LEV, kw=kw, upw=4, seg_major=0, xcd=1, group=0, direct=0)
            for _ in range(3):
                assert same(odp.set_aspl(adj), want), (n, kw)
        finally:
            odp.finalize_aspl()


def prism_graph(m):
    """Two m-cycles joined by rungs: 3-regular, diameter m // 2 + 1."""
    adj = np.zeros((2 * m, 3), np.int32)
    for i in range(m):
        adj[i] = [(i - 1) % m, (i + 1) % m, m + i]
        adj[m + i] = [m + (i - 1) % m, m + (i + 1) % m, i]
    return adj


@pytest.mark.parametrize("plan", [None, dict(plan=LEV, kw=4, upw=4, seg_major=0, xcd=1, direct=0),
                                  dict(plan=LEV, kw=2, upw=2, seg_major=1, xcd=1, direct=-1)])
def test_level_count_guess_mismatch(plan):
    """An evaluation launches the level the previous one ended on without writing
    its rows (counts only).  Alternating a short-diameter and a long-diameter
    graph on one Init makes that guess wrong both ways (the level turns out not
    to be the last, and rows are then written after all; or the graph ends
    earlier); every result must equal the oracle."""
    short = odp.generate_regular(200, 3, 41)
    long_ = prism_graph(100)
    ws, wl = oracle.evaluate(short, nodes=200), oracle.evaluate(long_, nodes=200)
    odp.init_aspl_general(200, 3)
    try:
        if plan:
            odp.set_plan(**plan)
        for adj, want in ((short, ws), (long_, wl), (long_, wl), (short, ws), (short, ws), (long_, wl), (short, ws)):
            assert same(odp.set_aspl(adj), want)
    finally:
        odp.finalize_aspl()


def _trimmed(adj, seed, drop):
    """adj with `drop` random edges removed (both directions): rows compacted,
    padded with -1, and their degrees."""
    rng = np.random.default_rng(seed)
    n, d = adj.shape
    rows = [list(r) for r in adj]
    for u in rng.choice(n, drop, replace=False):
        if len(rows[u]) > 1:
            v = rows[u][int(rng.integers(len(rows[u])))]
            if len(rows[v]) > 1 and u != v:
                rows[u].remove(v)
                rows[v].remove(u)
    out = np.full((n, d), -1, np.int32)
    nd = np.zeros(n, np.int32)
    for u, r in enumerate(rows):
        out[u, :len(r)] = r
        nd[u] = len(r)
    return out, nd


SPARSE_PLANS = [
    dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=-1),
    dict(plan=LEV, kw=4, upw=4, seg_major=1, xcd=0, direct=0, batch=8),
    dict(plan=LEV, kw=64, upw=2, seg_major=0, xcd=1, direct=0, batch=16),
    dict(plan=LEV, kw=2, upw=2, seg_major=0, xcd=1, group=2, direct=3),
    dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=-1, parts=4),
    dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=0, parts=2),
]


@pytest.mark.parametrize("frac", [None, "1000"])
@pytest.mark.parametrize("cap", [None, "0", "5"])
@pytest.mark.parametrize("plan", range(len(SPARSE_PLANS)))
def test_sparse_last_level(plan, cap, frac, monkeypatch):
    """An evaluation that expects to end at the level the previous one ended on
    lists, at the level before, the 16-byte chunks left unsaturated, and counts
    the last level from those alone (k_sparse); when the list is too small
    (ODP_HIP_LIST_CAP) it visits every chunk instead.  It lists only while the
    last evaluation left few chunks unsaturated (ODP_HIP_LIST_FRAC=1000: always).
    Repeated evaluations (eager, captured, replayed) of a connected graph, a
    disconnected one and one with absent edges all give the oracle's results."""
    if cap is not None:
        monkeypatch.setenv("ODP_HIP_LIST_CAP", cap)
    if frac is not None:
        monkeypatch.setenv("ODP_HIP_LIST_FRAC", frac)
    g = odp.generate_regular(1500, 5, 17)
    half = odp.generate_regular(750, 5, 18)
    split = np.concatenate([half, half + 750]).astype(np.int32)
    holes, nd = _trimmed(odp.generate_regular(1500, 5, 19), 7, 300)
    for adj, deg in ((g, None), (split, None), (holes, nd)):
        want = oracle.evaluate(adj, nodes=1500, num_degrees=deg)
        odp.init_aspl_general(1500, 5, deg)
        try:
            odp.set_plan(**SPARSE_PLANS[plan])
            for _ in range(4):
                assert same(odp.set_aspl(adj), want)
        finally:
            odp.finalize_aspl()


@pytest.mark.parametrize("n,d", [(16386, 6), (18000, 12), (20416, 4), (20418, 8), (24576, 16), (28672, 4),
                                 (30000, 12), (32768, 6), (40832, 8)])
def test_resident_past_16384_rows(n, d):
    """Single graphs of 16384 < n <= 40832 rows keep the resident plan, in kernel
    instances of their own: two-word blocks with up to 20 rows a thread up to
    n = 20416 (what LDS holds), one-word blocks with up to 28, 32 or 40 above.  The default
    plan and an explicit one match the oracle, also after a device swap."""
    adj = odp.generate_regular(n, d, 5 + n + d)
    want = oracle.evaluate(adj, nodes=n)
    moved = adj.copy()
    odp.two_opt(moved, odp.Rng(9))
    want_moved = oracle.evaluate(moved, nodes=n)
    wb = 2 if n <= 20416 else 1
    for plan in (None, dict(plan=RES, wb=wb, threads=1024)):
        odp.init_aspl_general(n, d)
        try:
            if plan is not None:
                odp.set_plan(**plan)
            assert odp.get_plan()["plan"] == RES and odp.get_plan()["wb"] == wb
            for _ in range(3):
                assert same(odp.set_aspl(adj), want)
            assert same(odp.set_aspl(moved), want_moved)
        finally:
            odp.finalize_aspl()


@pytest.mark.parametrize("n,d", [(20000, 16), (24576, 8)])
def test_dense_next_to_last_level(n, d):
    """Graphs whose next-to-last level leaves most chunks unsaturated (n = 20000,
    d = 16: diameter 5; n = 24576, d = 8: diameter 7): the first evaluation
    lists them for the sparse last level, later ones count that level with a
    plain count-only launch.  Every evaluation, and a 2-opt neighbour after it,
    gives the oracle's results."""
    adj = odp.generate_regular(n, d, 11 + n + d)
    want = oracle.evaluate(adj, nodes=n)
    moved = adj.copy()
    odp.two_opt(moved, odp.Rng(5))
    want_moved = oracle.evaluate(moved, nodes=n)
    odp.init_aspl_general(n, d)
    try:
        for _ in range(4):
            assert same(odp.set_aspl(adj), want)
        for _ in range(2):
            assert same(odp.set_aspl(moved), want_moved)
    finally:
        odp.finalize_aspl()


def test_batch_policy_autotune_and_graph_replay():
    """Every batch launch policy (block width per size class, workgroups per CU)
    gives the oracle's results; the autotuner keeps one of them; the captured
    launch replays identically with eager launches and host work in between."""
    graphs = _batch_graphs(3, 24, nmax=12000)
    graphs.append(two_triangles())
    packed = odp.pack(graphs)
    want = [oracle.evaluate(g, nodes=g.shape[0]) for g in graphs]
    ng = odp.batch_load(packed)
    try:
        for caps, pc in (([8] * 5, 2), ([1] * 5, 1), ([2, 4, 1, 2, 8], 2), ([4, 1, 8, 1, 2], 1)):
            odp.set_batch_policy(caps, pc)
            assert odp.batch_policy() == (caps, pc)
            d, sm, a = odp.batch_run(ng)
            for g in range(ng):
                assert same((int(d[g]), int(sm[g]), float(a[g])), want[g]), (caps, pc, g)
        with pytest.raises(ValueError):
            odp.set_batch_policy([3] * 5, 2)
        tried = odp.batch_autotune(1)
        assert tried >= 2
        caps, pc = odp.batch_policy()
        assert all(c in (1, 2, 4, 8) for c in caps) and pc in (1, 2)
        ref = odp.batch_run(ng)
        for _ in range(4):                     # replays, with eager work and syncs between them
            odp.stats_enable(True)
            eager = odp.batch_run(ng)
            odp.stats_enable(False)
            rep = odp.batch_run(ng)
            for x, y, z in zip(ref, eager, rep):
                assert x.tobytes() == y.tobytes() and x.tobytes() == z.tobytes()
        for g in range(ng):
            assert same((int(ref[0][g]), int(ref[1][g]), float(ref[2][g])), want[g])
    finally:
        odp.stats_enable(False)
        odp.set_batch_policy([8] * 5, 2)
        odp.batch_release()


def test_batch_large_graph_falls_back():
    graphs = [odp.generate_regular(24000, 4, 3), odp.generate_regular(500, 6, 4)]
    d, s, a = odp.batch_eval(odp.pack(graphs))
    for g, adj in enumerate(graphs):
        assert same((int(d[g]), int(s[g]), float(a[g])), oracle.evaluate(adj, nodes=adj.shape[0]))


def test_invalid_vertex_id_exits():
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import numpy as np, odp_amd as o\n"
        "a = np.array([[1, 2], [0, 2], [0, 7]], np.int32)\n"
        "o.init_aspl_general(3, 2)\n"
        "o.set_aspl(a)\n" % os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "outside [0, nodes)" in r.stderr


def test_invalid_id_in_a_small_change_exits():
    """A valid evaluation, then the same array with one entry made invalid:
    the change is too small to resend the table, but an invalid id still takes
    the full path and exits as the first evaluation would."""
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import odp_amd as o\n"
        "a = o.generate_regular(200, 4, 5)\n"
        "o.init_aspl_general(200, 4)\n"
        "o.set_aspl(a)\n"
        "print('first ok', flush=True)\n"
        "a[17, 2] = 200\n"
        "o.set_aspl(a)\n" % os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert "first ok" in r.stdout and r.returncode == 1 and "outside [0, nodes)" in r.stderr


@pytest.mark.parametrize("plan", [None, PLANS[7], PLANS[1]])
def test_graph_replay_matches_eager(plan):
    """Evaluations replay a captured hipGraph unless kernel timing is on (eager
    launches); several replays with eager swap kernels in between must equal the
    eager evaluation of the same graph every time."""
    n, d = 3000, 6
    adj = odp.generate_regular(n, d, 31)
    rng = odp.Rng(9)
    odp.init_aspl_general(n, d)
    if plan:
        odp.set_plan(**plan)
    host = adj.copy()
    odp.upload_adjacency(host)
    try:
        for _ in range(4):
            a, b, c, dd = odp.two_opt(host, rng)
            assert odp.swap_device(a, b, c, dd) == 0
            replay = [odp.set_aspl_resident() for _ in range(3)]
            odp.stats_enable(True)
            eager = odp.set_aspl_resident()
            odp.stats_enable(False)
            want = oracle.evaluate(host, nodes=n)
            assert all(same(r, eager) for r in replay) and same(eager, want)
    finally:
        odp.stats_enable(False)
        odp.finalize_aspl()


def direct_eval(adj, K, kw, group=0, **init):
    """Evaluate with the first K levels enumerated by k_direct; returns (result, K that ran)."""
    nodes = init.pop("nodes", None)
    width, height = init.pop("width", None), init.pop("height", None)
    s = init.pop("symmetries", 1)
    nd = init.pop("num_degrees", None)
    if width is None:
        (odp.init_aspl_general(nodes, adj.shape[1], nd) if s == 1
         else odp.init_aspl_general_s(nodes, adj.shape[1], nd, s))
    else:
        (odp.init_aspl_grid(width, height, adj.shape[1], nd) if s == 1
         else odp.init_aspl_grid_s(width, height, adj.shape[1], nd, s))
    try:
        odp.set_plan(plan=LEV, kw=kw, upw=4, seg_major=0, xcd=1, group=group, direct=K)
        ran = odp.get_plan()["direct"]
        return odp.set_aspl(adj), ran
    finally:
        odp.finalize_aspl()


@pytest.mark.parametrize("K,kw,group", [(2, 2, 0), (3, 16, 0), (4, 4, 3), (5, 8, 1), (6, 2, 0), (3, 32, 2),
                                        (2, 128, 0)])
def test_direct_levels(K, kw, group):
    """k_direct (levels 0..K-1 by walk enumeration) against the oracle: regular,
    symmetric (circulant and grid), non-regular, disconnected and asymmetric
    tables; every count and word must equal the gather levels'."""
    cases = []
    adj = odp.generate_regular(2400, 5, 7)
    cases.append((adj, dict(nodes=2400), oracle.evaluate(adj, nodes=2400)))
    big = odp.generate_regular(8192, 4, 11)                       # 128 words per row
    cases.append((big, dict(nodes=8192), oracle.evaluate(big, nodes=8192)))
    base = odp.generate_general_s(1200, 4, 4, 5)
    cases.append((base, dict(nodes=1200, symmetries=4),
                  oracle.evaluate(oracle.expand(base, nodes=1200, symmetries=4), nodes=1200)))
    full = odp.generate_grid(40, 40, 6, 4, 4, 3)
    gb = odp.grid_base_table(40, 40, 6, 4, full)
    cases.append((gb, dict(width=40, height=40, symmetries=4), oracle.evaluate(full, width=40, height=40)))
    pa, nd = path_graph(300)
    cases.append((pa, dict(nodes=300, num_degrees=nd), oracle.evaluate(pa, nodes=300, num_degrees=nd)))
    dis = np.concatenate([cycle_graph(500), cycle_graph(700) + 500]).astype(np.int32)
    cases.append((dis, dict(nodes=1200), (INT_MAX, LONG_MAX, DBL_MAX)))
    rng = np.random.default_rng(K * 100 + kw)
    asym = rng.integers(0, 900, size=(900, 3)).astype(np.int32)    # a directed table, rows need not agree
    asym[:, 0] = (np.arange(900) + 1) % 900                         # with a directed cycle through all rows
    cases.append((asym, dict(nodes=900), oracle.evaluate(asym, nodes=900)))
    ran_any = False
    for adj, init, want in cases:
        got, ran = direct_eval(adj, K, kw, group, **dict(init))
        assert same(got, want), (K, kw, group, init, got, want)
        ref, _ = direct_eval(adj, 0, kw, group, **dict(init))
        assert same(ref, want)
        ran_any = ran_any or ran >= 2
        assert ran == 0 or 2 <= ran <= K
    assert ran_any, "k_direct never applied"


@pytest.mark.parametrize("n,d", [(66000, 6), (70002, 4)])
def test_past_16bit_ids(n, d):
    """Graphs just past 65536 vertices: no 16-bit id table, no resident plan;
    the level plan's 32-bit tables and k_direct's 32-bit path (K = 4 and 6)."""
    adj = odp.generate_regular(n, d, 99)
    want = oracle.evaluate(adj, nodes=n)
    odp.init_aspl_general(n, d)
    try:
        assert odp.get_plan()["plan"] == LEV
        for _ in range(2):
            assert same(odp.set_aspl(adj), want)
    finally:
        odp.finalize_aspl()


@pytest.mark.parametrize("parts", [2, 4])
def test_part_phased_gather(parts):
    """k_pgather (the XCD-phased gather over row parts) against the oracle:
    graphs whose rows hold several 16-word blocks and a partial last block,
    regular and non-regular tables, a disconnected graph, MATRIX and groups,
    repeated evaluations (frontier first, then plain levels with the
    count-only, listing and sparse last levels); the plan reports the parts
    that run."""
    cases = []
    g = odp.generate_regular(3000, 12, 71)                       # 47 words: 2 full blocks + 15 words
    cases.append((g, None))
    holes, nd = _trimmed(odp.generate_regular(2200, 9, 72), 5, 200)
    cases.append((holes, nd))
    half = odp.generate_regular(1100, 6, 73)
    cases.append((np.concatenate([half, half + 1100]).astype(np.int32), None))
    for adj, deg in cases:
        n = adj.shape[0]
        want = oracle.evaluate(adj, nodes=n, num_degrees=deg)
        for group in (0, 2):
            odp.init_aspl_general(n, adj.shape[1], deg)
            try:
                odp.set_plan(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=-1, group=group, parts=parts)
                assert odp.get_plan()["parts"] == parts
                for _ in range(4):
                    assert same(odp.set_aspl(adj), want), (n, group)
            finally:
                odp.finalize_aspl()
    # rows narrower than 16 words: the plan falls back to k_level and says so
    small = odp.generate_regular(500, 4, 74)
    odp.init_aspl_general(500, 4)
    try:
        odp.set_plan(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, parts=parts)
        assert odp.get_plan()["parts"] == 0
        assert same(odp.set_aspl(small), oracle.evaluate(small, nodes=500))
    finally:
        odp.finalize_aspl()
