"""Serving engine (ODPX_Serve_*): every result is the same alone and in the mix,
and equals the oracle on the graph the request saw."""
import numpy as np
import pytest

import odp_amd as odp
import oracle

pytestmark = pytest.mark.gpu


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()


def chain_plan(seed):
    """(initial adjacency, [swap per request], [graph after each request]) per chain."""
    specs = [(512, 4), (2000, 8), (4096, 8), (6000, 12), (20000, 4), (18000, 6)]   # the last two are large
    rng = odp.Rng(seed)
    chains = []
    for i, (n, d) in enumerate(specs):
        adj = odp.generate_regular(n, d, seed * 100 + i)
        cur = adj.copy()
        swaps, states = [], []
        for _ in range(3):
            swaps.append(odp.two_opt(cur, rng))
            states.append(cur.copy())
        chains.append((adj, swaps, states))
    return chains


def run(chains, mixed):
    srv = odp.Server(small_slots=2, large_slots=2, hbm_budget_gb=16)
    try:
        ids = [srv.add_chain(c[0]) for c in chains]
        results = {i: [] for i in range(len(chains))}
        if mixed:
            order = []
            for k in range(3):
                for i, (_, swaps, _) in enumerate(chains):
                    order.append((i, srv.submit(ids[i], [swaps[k]])))
            srv.drain()
            got = {r["id"]: r for r in srv.poll()}
            for i, rid in order:
                results[i].append(got[rid])
        else:
            for i, (_, swaps, _) in enumerate(chains):
                for k in range(3):
                    results[i].append(srv.wait(srv.submit(ids[i], [swaps[k]])))
        return results
    finally:
        srv.stop()


@pytest.mark.skip(reason="a full GPU-suite run ended in a host segmentation fault inside a serving "
                         "worker thread during the mixed run (two large workers capturing hipGraphs "
                         "concurrently); the cause was not shown; the threaded engine it ran was "
                         "replaced by the single-threaded progress engine, tested below")
def test_alone_equals_mixed_and_oracle():
    chains = chain_plan(3)
    alone = run(chains, mixed=False)
    mixed = run(chains, mixed=True)
    for i, (_, _, states) in enumerate(chains):
        for k in range(3):
            a, m = alone[i][k], mixed[i][k]
            assert a["status"] == 0 and m["status"] == 0
            ra = (a["diameter"], a["sum"], a["ASPL"])
            rm = (m["diameter"], m["sum"], m["ASPL"])
            assert same(ra, rm), (i, k, ra, rm)
            want = oracle.evaluate(states[k], nodes=states[k].shape[0])
            assert same(ra, want), (i, k, ra, want)
            assert m["t_submit"] <= m["t_start"] <= m["t_done"]


def test_budget_refuses_chain():
    srv = odp.Server(small_slots=1, large_slots=1, hbm_budget_gb=0.0001)   # 100 KB < 196 KB needed
    try:
        with pytest.raises(MemoryError):
            srv.add_chain(odp.generate_regular(4096, 8, 1))
    finally:
        srv.stop()


def test_missing_edge_is_reported():
    srv = odp.Server(small_slots=1, large_slots=0, hbm_budget_gb=4)
    try:
        adj = odp.generate_regular(300, 4, 2)
        cid = srv.add_chain(adj)
        a = 0
        b = int(adj[0][0])
        c = next(v for v in range(1, 300) if v != b and v not in adj[0] and v not in adj[b])
        r = srv.wait(srv.submit(cid, [(a, b, c, c)]))     # c-c is not an edge
        assert r["status"] == -1
        # the refused swap changed nothing: an empty request sees the original graph
        r = srv.wait(srv.submit(cid, []))
        assert r["status"] == 0
        assert same((r["diameter"], r["sum"], r["ASPL"]), oracle.evaluate(adj, nodes=300))
    finally:
        srv.stop()


def progress_chains(seed):
    """Seven chains, four small (LDS-resident) and three large; per chain four
    requests carrying 0, 1, 2 and 1 swaps, and the graph after each."""
    specs = [(600, 4), (3000, 6), (4096, 8), (7000, 10), (17000, 4), (24000, 6), (20000, 8)]
    rng = odp.Rng(seed)
    chains = []
    for i, (n, d) in enumerate(specs):
        adj = odp.generate_regular(n, d, seed * 1000 + i)
        cur = adj.copy()
        reqs, states = [], []
        for nsw in (0, 1, 2, 1):
            reqs.append([odp.two_opt(cur, rng) for _ in range(nsw)])
            states.append(cur.copy())
        chains.append((adj, reqs, states))
    return chains


def test_progress_engine_mix_equals_alone_and_oracle():
    """Progress engine: every request of a mixed stream (all chains submitted,
    interleaved with polls, two small and two large slots) equals the same
    request served alone (one chain per server, one request at a time) and
    the oracle on the graph it saw; requests admitted together overlap."""
    chains = progress_chains(5)
    alone = []
    for adj, reqs, _ in chains:
        srv = odp.Server(small_slots=1, large_slots=1, hbm_budget_gb=16)
        try:
            cid = srv.add_chain(adj)
            alone.append([srv.wait(srv.submit(cid, r)) for r in reqs])
        finally:
            srv.stop()
    srv = odp.Server(small_slots=2, large_slots=2, hbm_budget_gb=16)
    try:
        cids = [srv.add_chain(c[0]) for c in chains]
        order, got = [], {}
        for k in range(4):
            for i, (_, reqs, _) in enumerate(chains):
                order.append((i, k, srv.submit(cids[i], reqs[k])))
            for r in srv.poll():
                got[r["id"]] = r
        srv.drain()
        for r in srv.poll():
            got[r["id"]] = r
    finally:
        srv.stop()
    assert len(got) == len(order)
    for i, k, rid in order:
        m, a = got[rid], alone[i][k]
        assert m["status"] == 0 and a["status"] == 0, (i, k)
        rm, ra = (m["diameter"], m["sum"], m["ASPL"]), (a["diameter"], a["sum"], a["ASPL"])
        assert same(rm, ra), (i, k, rm, ra)
        st = chains[i][2][k]
        want = oracle.evaluate(st, nodes=st.shape[0])
        assert same(rm, want), (i, k, rm, want)
        assert m["t_submit"] <= m["t_start"] <= m["t_done"]
    # the first requests of the two large slots were admitted together and ran at once
    first_large = [got[rid] for i, k, rid in order if k == 0 and chains[i][0].shape[0] > 16384][:2]
    assert first_large[0]["t_start"] < first_large[1]["t_done"] and first_large[1]["t_start"] < first_large[0]["t_done"]
