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


def test_anneal_rejects_symmetric_init():
    odp.init_aspl_general_s(64, 4, None, 2)
    try:
        with pytest.raises(ValueError):
            odp.anneal(odp.generate_regular(32, 4, 1), 10)
    finally:
        odp.finalize_aspl()
