"""Simulated annealing with the adjacency on the device (ODPX_Anneal) against a
host replica of the reference's annealing loop (sample/SA/general.c:149-181,
acceptance test sample/SA/common.c:15-43) that evaluates every graph with the
oracle.  Both draw the moves and the acceptance numbers from the same seeded
generator, so the whole trajectory must agree: steps, accepted moves, the best
graph and its integers bit for bit.
"""
import math

import numpy as np
import pytest

import odp_amd as odp
import oracle

pytestmark = pytest.mark.gpu

INT_MAX = 2**31 - 1


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()


def accept(nodes, cur_d, d, cur_a, a, temp, hill, prio, rng):
    if hill:
        return a <= cur_a
    if d < cur_d and not prio:
        return True
    if d > cur_d and not prio:
        return False
    if a <= cur_a:
        return True
    diff = (cur_a - a) * nodes * (nodes - 1)
    u = ((rng.next() >> 11) + 0.5) / 9007199254740992.0
    return math.exp(diff / temp) > u


def replica(adj, ncalcs, max_temp, min_temp, hill, prio, seed):
    n, deg = adj.shape
    cur = adj.copy()
    rng = odp.Rng(seed)
    r = oracle.evaluate(cur, nodes=n)
    cur_d, cur_a = r[0], r[2]
    best, best_adj = r, cur.copy()
    low_d, low_a = odp.lbounds_general(n, deg)
    steps = accepted = 0
    optimum = r[0] == low_d and r[2] == low_a
    cooling = 0.0 if hill else math.pow(min_temp / max_temp, 1.0 / ncalcs)
    temp = max_temp
    i = 0
    while i < ncalcs and not optimum:
        a, b, c, d = odp.two_opt(cur, rng)
        r = oracle.evaluate(cur, nodes=n)
        steps += 1
        if r[0] < best[0] or (r[0] == best[0] and r[2] < best[2]):
            best, best_adj = r, cur.copy()
            if r[0] == low_d and r[2] == low_a:
                optimum = True
                break
        if accept(n, cur_d, r[0], cur_a, r[2], temp, hill, prio, rng):
            cur_d, cur_a = r[0], r[2]
            accepted += 1
        else:
            odp.swap_edges(cur, a, c, b, d)
        temp *= cooling
        i += 1
    return best_adj, dict(steps=steps, accepted=accepted, best=best, diameter=cur_d, ASPL=cur_a,
                          optimum=int(optimum))


CASES = [
    # nodes, degree, ncalcs, max_temp, min_temp, hill, ASPL priority, seed, plan
    (128, 4, 300, 50.0, 0.5, False, False, 3, None),
    (100, 3, 200, 100.0, 1.0, True, False, 5, None),
    (96, 5, 250, 20.0, 0.2, False, True, 7, None),
    (600, 5, 60, 80.0, 1.0, False, False, 11, dict(plan=odp.PLAN_LEVEL, kw=4, upw=4, seg_major=0, xcd=1, direct=-1)),
    # row-part gathers: every move and undo patches the tables and the units' pair lists
    (2100, 12, 40, 80.0, 1.0, False, False, 13,
     dict(plan=odp.PLAN_LEVEL, kw=16, upw=8, seg_major=0, xcd=1, direct=-1, parts=4)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_anneal_matches_replica(case):
    n, deg, ncalcs, tmax, tmin, hill, prio, seed, plan = CASES[case]
    adj = odp.generate_regular(n, deg, seed)
    want_adj, want = replica(adj, ncalcs, tmax, tmin, hill, prio, seed)
    odp.init_aspl_general(n, deg)
    try:
        if plan:
            odp.set_plan(**plan)
        got_adj, got = odp.anneal(adj, ncalcs, tmax, tmin, hill, prio, seed)
        # the library's state afterwards is the chain's current graph; its best is returned
        assert same(odp.set_aspl(got_adj), want["best"])
    finally:
        odp.finalize_aspl()
    assert got["steps"] == want["steps"] and got["accepted"] == want["accepted"], (got, want)
    assert same((got["best_diameter"], got["best_sum"], got["best_ASPL"]), want["best"])
    assert got["diameter"] == want["diameter"] and got["ASPL"] == want["ASPL"]
    assert got["optimum"] == want["optimum"]
    assert np.array_equal(got_adj, want_adj)
    assert got["best_diameter"] < INT_MAX


def test_device_loop_many_launches(monkeypatch):
    """n=1024, d=6 (32 blocks of 32 sources, one workgroup each): 3000 steps in
    launches of 700 (the chain's state carried between launches) against the
    replica, and against the host-stepped loop (ODP_HIP_ANNEAL=host)."""
    n, deg, ncalcs, seed = 1024, 6, 3000, 21
    adj = odp.generate_regular(n, deg, seed)
    want_adj, want = replica(adj, ncalcs, 30.0, 0.5, False, False, seed)
    runs = []
    for env in ({"ODP_HIP_ANNEAL_CHUNK": "700"}, {"ODP_HIP_ANNEAL": "host"}):
        for k, val in env.items():
            monkeypatch.setenv(k, val)
        odp.init_aspl_general(n, deg)
        try:
            got_adj, got = odp.anneal(adj, ncalcs, 30.0, 0.5, False, False, seed)
            cur = odp.set_aspl_resident()          # the device copy holds the chain's current graph
        finally:
            odp.finalize_aspl()
        for k in env:
            monkeypatch.delenv(k)
        assert got["steps"] == want["steps"] and got["accepted"] == want["accepted"], (env, got, want)
        assert same((got["best_diameter"], got["best_sum"], got["best_ASPL"]), want["best"])
        assert got["diameter"] == want["diameter"] and got["ASPL"] == want["ASPL"]
        assert cur[0] == want["diameter"] and cur[2] == want["ASPL"]
        assert np.array_equal(got_adj, want_adj)
        runs.append(got)
    assert runs[0]["steps"] == ncalcs


def grid_length(base, w, h, s):
    bh = h // 2
    length = 0
    for i, row in enumerate(base):
        x = (i // bh) * h + (i % bh) if s == 4 else i
        for u in row:
            if u >= 0:
                length = max(length, abs(int(u) // h - x // h) + abs(int(u) % h - x % h))
    return length


def replica_s(base, ncalcs, max_temp, min_temp, hill, prio, seed, nodes=None, width=None, height=None,
              symmetries=1):
    """The loops of sample/SA/general_s.c and grid(_s).c on the oracle, with the
    library's seeded moves (ODPX_Mutate) and accept_s."""
    kw = dict(nodes=nodes) if width is None else dict(width=width, height=height)
    n = nodes if width is None else width * height
    s = symmetries
    cur = base.copy()
    rng = odp.Rng(seed)
    length = grid_length(cur, width, height, s) if width is not None else 0
    r = oracle.evaluate(cur, symmetries=s, **kw)
    cur_d, cur_a = r[0], r[2]
    best, best_adj = r, cur.copy()
    low_d, low_a = (odp.lbounds_grid(width, height, base.shape[1], length) if width is not None
                    else odp.lbounds_general(n, base.shape[1]))
    steps = accepted = 0
    optimum = r[0] == low_d and r[2] == low_a
    cooling = 0.0 if hill else math.pow(min_temp / max_temp, 1.0 / ncalcs)
    temp = max_temp
    i = 0
    while i < ncalcs and not optimum:
        rec = odp.mutate(cur, rng, length=length, symmetries=s, **kw)
        r = oracle.evaluate(cur, symmetries=s, **kw)
        steps += 1
        if r[0] < best[0] or (r[0] == best[0] and r[2] < best[2]):
            best, best_adj = r, cur.copy()
            if r[0] == low_d and r[2] == low_a:
                optimum = True
                break
        ok = True
        if hill:
            ok = r[2] <= cur_a
        elif r[0] < cur_d and not prio:
            ok = True
        elif r[0] > cur_d and not prio:
            ok = False
        elif r[2] <= cur_a:
            ok = True
        else:
            diff = ((cur_a - r[2]) * n * (n - 1)) / s
            u = ((rng.next() >> 11) + 0.5) / 9007199254740992.0
            ok = math.exp(diff / temp) > u
        if ok:
            cur_d, cur_a = r[0], r[2]
            accepted += 1
        else:
            (odp.restore_grid if width is not None else odp.restore_general)(rec, cur)
        temp *= cooling
        i += 1
    return best_adj, dict(steps=steps, accepted=accepted, best=best, diameter=cur_d, ASPL=cur_a,
                          optimum=int(optimum))


SYM_CASES = [
    # kind, init kwargs, degree, ncalcs, max_temp, min_temp, hill, ASPL priority, seed
    ("general_s", dict(nodes=96, symmetries=4), 4, 200, 40.0, 0.3, False, False, 3),
    ("general_s", dict(nodes=120, symmetries=2), 3, 150, 20.0, 0.2, False, True, 4),
    ("grid_s", dict(width=12, height=12, symmetries=4), 4, 150, 30.0, 0.3, False, False, 5),
    ("grid_s", dict(width=10, height=8, symmetries=2), 3, 150, 30.0, 0.3, True, False, 6),
    ("grid", dict(width=10, height=10, symmetries=1), 4, 150, 30.0, 0.3, False, False, 7),
]


@pytest.mark.parametrize("case", range(len(SYM_CASES)))
def test_anneal_symmetric_matches_replica(case):
    kind, kw, deg, ncalcs, tmax, tmin, hill, prio, seed = SYM_CASES[case]
    s = kw["symmetries"]
    if kind == "general_s":
        base = odp.generate_general_s(kw["nodes"], deg, s, seed)
        odp.init_aspl_general_s(kw["nodes"], deg, None, s)
        okw = dict(nodes=kw["nodes"])
    else:
        w, h = kw["width"], kw["height"]
        full = odp.generate_grid(w, h, deg, 3, s, seed)
        base = odp.grid_base_table(w, h, deg, s, full) if s > 1 else full
        (odp.init_aspl_grid(w, h, deg) if s == 1 else odp.init_aspl_grid_s(w, h, deg, None, s))
        okw = dict(width=w, height=h)
    try:
        got_adj, got = odp.anneal(base, ncalcs, tmax, tmin, hill, prio, seed)
    finally:
        odp.finalize_aspl()
    want_adj, want = replica_s(base, ncalcs, tmax, tmin, hill, prio, seed, symmetries=s, **okw)
    assert got["steps"] == want["steps"] and got["accepted"] == want["accepted"], (got, want)
    assert same((got["best_diameter"], got["best_sum"], got["best_ASPL"]), want["best"])
    assert got["diameter"] == want["diameter"] and got["ASPL"] == want["ASPL"]
    assert np.array_equal(got_adj, want_adj)
