"""The benchmark configurations at their full sizes, through the C ABI.

* n=65536, d=16 random regular graph, MATRIX (and MATRIX_SAVING), several plans;
* the 256x256 4-fold symmetric grid (d=8, edge length <= 4), MATRIX_SAVING and
  MATRIX, temporal tiles and per-level plans;
* the 1024 batched 2-opt candidates (n up to 16384, d up to 24): one batch
  equals looping the single-graph path under both plans, byte for byte;
* 64 served chains (n in {4096, 16384, 65536} x d in {8, 16, 32}): every
  request of a mixed stream equals the same request served alone.

Expected values come from the reference library (oracle/_ref, run in the
container that holds the reference sources) on the same regenerated graphs,
committed in tests/golden/vectors.json by tests/golden/make_vectors.py.
"""
import numpy as np
import pytest

import bench
import odp_amd as odp
from goldens import C4_SAMPLE, C5_SAMPLE, build_case, vectors

pytestmark = pytest.mark.gpu

LEV = odp.PLAN_LEVEL


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()


def vector(**match):
    for c in vectors():
        if all(c.get(k) == v for k, v in match.items()):
            return c, (c["diameter"], c["sum"], float.fromhex(c["aspl"]))
    raise KeyError(match)


_cache = {}


def case_graph(c):
    key = repr(sorted((k, v) for k, v in c.items() if k not in ("diameter", "sum", "aspl")))
    if key not in _cache:
        _cache[key] = build_case(c)
    return _cache[key]


# plans for the n=65536 graph; "env" sets ODP_ASPL for the default plan
C2_PLANS = {
    "default": (None, None),
    "saving": ("MATRIX_SAVING", None),
    "blocks16-k_level": (None, dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=-1, parts=0)),
    "parts2-no-direct": (None, dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=0, parts=2)),
    "blocks64-gather": (None, dict(plan=LEV, kw=64, upw=8, seg_major=0, xcd=1, direct=0)),
    "whole-rows": (None, dict(plan=LEV, kw=0, upw=4, seg_major=1, xcd=0, direct=0)),
}


@pytest.mark.parametrize("name", list(C2_PLANS))
def test_random65536(name, monkeypatch):
    c, want = vector(graph="regular", nodes=65536, degree=16, seed=1)
    kw, adj = case_graph(c)
    env, plan = C2_PLANS[name]
    if env:
        monkeypatch.setenv("ODP_ASPL", env)
    else:
        monkeypatch.delenv("ODP_ASPL", raising=False)
    print(f"random65536 plan={name}", flush=True)
    odp.init_aspl_general(65536, 16)
    try:
        if plan:
            odp.set_plan(**plan)
        if name == "default":
            assert odp.get_plan()["parts"] == 4
        assert same(odp.set_aspl(adj), want)
        for _ in range(3):                      # eager, captured, replayed
            assert same(odp.set_aspl_resident(), want)
    finally:
        odp.finalize_aspl()


C3_PLANS = {
    "saving-default": ("MATRIX_SAVING", None),
    "matrix-default": ("MATRIX", None),
    "frontier-levels": ("MATRIX_SAVING", dict(plan=LEV, kw=2, upw=8, seg_major=0, xcd=1, group=256, direct=-1, tile=0)),
    "plain-blocks16": ("MATRIX", dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=0)),
}


@pytest.mark.parametrize("name", list(C3_PLANS))
def test_grid256(name, monkeypatch):
    c, want = vector(graph="grid", width=256, height=256, degree=8, length=4, symmetries=4, seed=1)
    kw, base = case_graph(c)
    env, plan = C3_PLANS[name]
    monkeypatch.setenv("ODP_ASPL", env)
    print(f"grid256 plan={name}", flush=True)
    odp.init_aspl_grid_s(256, 256, 8, None, 4)
    try:
        if plan:
            odp.set_plan(**plan)
        if name.endswith("default"):
            assert odp.get_plan()["tile"] == 1
        assert same(odp.set_aspl(base), want)
        for _ in range(2):
            assert same(odp.set_aspl_resident(), want)
    finally:
        odp.finalize_aspl()


def test_batch1024_equals_loops_and_reference():
    """One batched run of the 1024 candidates equals looping the single-graph
    path with the level plan and with the library's default plan, byte for
    byte; sampled graphs equal the reference's values."""
    graphs = bench.c4_graphs(1024)
    ng = odp.batch_load(odp.pack(graphs))
    try:
        print("batch1024: batch run", flush=True)
        got = odp.batch_run(ng)
        again = odp.batch_run(ng)                    # second run: the captured launch graph
        print("batch1024: loops", flush=True)
        lev = odp.loop_run(ng, odp.PLAN_LEVEL)
        auto = odp.loop_run(ng, odp.PLAN_AUTO)
        for other in (again, lev, auto):
            for x, y in zip(got, other):
                assert x.tobytes() == y.tobytes()
        assert (got[0] < 2**31 - 1).all()
        for g in C4_SAMPLE:
            c, want = vector(graph="batch", index=g, batch_seed=2024)
            assert same((int(got[0][g]), int(got[1][g]), float(got[2][g])), want), g
    finally:
        odp.batch_release()


def test_served_64_chains_mixed_equals_alone():
    """64 chains (every (n, d) class of the served workload), three requests of
    one 2-opt swap each.  Served as a mixed stream (two small and four large
    slots, all chains submitted between polls) and again one request at a time:
    every result is identical; the chains' final states equal the reference's
    values for one chain of each class; requests of the mix overlapped."""
    nchains, nreq = 64, 3
    starts = [bench.c5_chain(i) for i in range(nchains)]
    reqs = []
    for i in range(nchains):
        cur, rng = starts[i].copy(), odp.Rng(9000 + i)
        reqs.append([odp.two_opt(cur, rng) for _ in range(nreq)])

    def serve(mixed):
        srv = odp.Server(small_slots=2, large_slots=4, hbm_budget_gb=64)
        try:
            cids = [srv.add_chain(a) for a in starts]
            res = {}
            if mixed:
                order = []
                for k in range(nreq):
                    for i in range(nchains):
                        order.append((i, k, srv.submit(cids[i], [reqs[i][k]])))
                    for r in srv.poll():
                        res[r["id"]] = r
                srv.drain()
                for r in srv.poll():
                    res[r["id"]] = r
                return {(i, k): res[rid] for i, k, rid in order}
            out = {}
            for i in range(nchains):
                for k in range(nreq):
                    out[(i, k)] = srv.wait(srv.submit(cids[i], [reqs[i][k]]))
            return out
        finally:
            srv.stop()

    print("served64: mixed", flush=True)
    mixed = serve(True)
    print("served64: alone", flush=True)
    alone = serve(False)
    for key in alone:
        m, a = mixed[key], alone[key]
        assert m["status"] == 0 and a["status"] == 0, key
        assert same((m["diameter"], m["sum"], m["ASPL"]), (a["diameter"], a["sum"], a["ASPL"])), key
        assert m["t_submit"] <= m["t_start"] <= m["t_done"]
    for i in C5_SAMPLE:
        c, want = vector(graph="chain", chain=i, swaps=nreq, rng=9000 + i)
        r = mixed[(i, nreq - 1)]
        assert same((r["diameter"], r["sum"], r["ASPL"]), want), i
    spans = sorted((r["t_start"], r["t_done"]) for r in mixed.values())
    assert any(spans[j + 1][0] < spans[j][1] for j in range(len(spans) - 1)), "no two requests overlapped"
