"""Temporal tiles (k_tile) on grid graphs vs the oracle, through the C ABI.

A tile launch runs k = min(8, 16 // lmax) levels of a 32 x 32 tile of cells in
LDS (lmax: the longest table entry in x or y).  Every result must equal the
oracle's bit for bit, for every k, table width, symmetry, grid shape, group
size and repeated (captured, replayed) evaluation.
"""
import os

import numpy as np
import pytest

import odp_amd as odp
import oracle

pytestmark = pytest.mark.gpu

LEV = odp.PLAN_LEVEL
TILE_PLAN = dict(plan=LEV, kw=2, upw=8, seg_major=0, xcd=1, direct=-1, tile=1)

# (width, height, degree, length, symmetries): lmax = length here, k = min(8, 16 // lmax)
TILE_GRIDS = [
    (64, 64, 8, 4, 4),      # k = 4
    (96, 40, 6, 3, 2),      # k = 5; height not a multiple of 32
    (70, 70, 4, 2, 1),      # k = 8; 3 x 3 tiles, partial ones at the edges
    (34, 34, 8, 6, 2),      # k = 2
    (64, 64, 6, 12, 4),     # k = 1
    (48, 48, 12, 5, 1),     # table width 12
    (40, 40, 16, 6, 1),     # table width 16
    (130, 130, 6, 5, 2),    # 5 x 5 tiles, k = 3
]


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()


def lmax_of(full, h):
    n, d = full.shape
    v = np.repeat(np.arange(n), d)
    u = full.reshape(-1)
    m = u >= 0
    return int(np.max(np.maximum(np.abs(u[m] // h - v[m] // h), np.abs(u[m] % h - v[m] % h))))


def init(w, h, d, s, nd=None):
    if s == 1:
        odp.init_aspl_grid(w, h, d, nd)
    else:
        odp.init_aspl_grid_s(w, h, d, nd, s)


def run(base, w, h, d, s, plan, want, nd=None, repeats=3):
    """Evaluations on one Init (eager, captured, replayed); returns the number of
    gather launches of a timed evaluation."""
    init(w, h, d, s, nd)
    try:
        odp.set_plan(**plan)
        for _ in range(repeats):
            assert same(odp.set_aspl(base), want), plan
        odp.stats_enable(True)
        odp.stats_reset()
        assert same(odp.set_aspl(base), want), plan
        st = odp.stats()
        odp.stats_enable(False)
        return st["level_launches"] + st["idle_launches"]
    finally:
        odp.stats_enable(False)
        odp.finalize_aspl()


@pytest.mark.parametrize("case", range(len(TILE_GRIDS)))
def test_tile_levels(case):
    w, h, d, length, s = TILE_GRIDS[case]
    full = odp.generate_grid(w, h, d, length, s, 5)
    base = full if s == 1 else odp.grid_base_table(w, h, d, s, full)
    want = oracle.evaluate(base, width=w, height=h, symmetries=s)
    side = int(os.environ.get("ODP_HIP_TILE_SIDE", "32"))      # the library's tile side (halo = (64 - side) / 2)
    k = min(8, ((64 - side) // 2) // lmax_of(full, h))
    for plan in (TILE_PLAN, dict(TILE_PLAN, direct=0), dict(TILE_PLAN, group=3, xcd=0)):
        launches = run(base, w, h, d, s, plan, want)
        # the gather levels ran as launches of k levels each (plus one launch
        # past the end at most) per group of blocks, not one launch per level
        blocks = -(-(-(-(w * h // s) // 64)) // 2)
        groups = -(-blocks // plan["group"]) if plan.get("group") else 1
        if k == 0:          # entries longer than the halo: one level per launch
            assert launches >= want[0] - 6, (plan, launches)
        else:
            assert launches <= groups * ((want[0] - 1) // k + 2), (plan, launches, k)


def lattice(w, h, cut=None):
    """The 4-neighbour lattice (lmax 1): rows padded with -1, and num_degrees.
    With cut, no edge joins columns cut - 1 and cut (two components)."""
    n = w * h
    adj = np.full((n, 4), -1, np.int32)
    nd = np.zeros(n, np.int32)
    for x in range(w):
        for y in range(h):
            nb = []
            for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                a, b = x + dx, y + dy
                if 0 <= a < w and 0 <= b < h and not (cut is not None and {x, a} == {cut - 1, cut}):
                    nb.append(a * h + b)
            adj[x * h + y, :len(nb)] = nb
            nd[x * h + y] = len(nb)
    return adj, nd


@pytest.mark.parametrize("cut", [None, 20])
def test_tile_lattice_nonregular(cut):
    """lmax 1 (k = 8), rows of degree 2..4 through num_degrees, 63 x 45 cells;
    cut into two components the result is the sentinel triple."""
    w, h = 63, 45
    adj, nd = lattice(w, h, cut)
    want = oracle.evaluate(adj, width=w, height=h, num_degrees=nd)
    if cut is not None:
        assert want[0] == 2**31 - 1
    run(adj, w, h, 4, 1, TILE_PLAN, want, nd=nd)


def test_tile_long_entries_fall_back():
    """Entries longer than the halo (lmax > 16) turn the tiles off for that
    adjacency: the plan stays valid and the gather levels run one per launch."""
    w = h = 64
    full = odp.generate_grid(w, h, 6, 24, 4, 7)
    assert lmax_of(full, h) > 16
    base = odp.grid_base_table(w, h, 6, 4, full)
    want = oracle.evaluate(base, width=w, height=h, symmetries=4)
    launches = run(base, w, h, 6, 4, TILE_PLAN, want)
    assert launches >= want[0] - 4


def test_tile_swaps_lengthen_halo():
    """2-opt swaps on the device copy of an s = 1 grid: a swap that creates a
    longer edge lowers k for the next evaluation; every result equals the
    oracle on the swapped table."""
    w, h = 64, 48
    adj = odp.generate_grid(w, h, 4, 2, 1, 11)
    init(w, h, 4, 1)
    try:
        odp.set_plan(**TILE_PLAN)
        assert same(odp.set_aspl(adj), oracle.evaluate(adj, width=w, height=h))
        rng = np.random.default_rng(3)
        cur = adj.copy()
        done = 0
        while done < 6:
            a = int(rng.integers(w * h))
            b = int(cur[a][int(rng.integers(4))])
            c = int(rng.integers(w * h))
            dd = int(cur[c][int(rng.integers(4))])
            if len({a, b, c, dd}) < 4 or c in cur[a] or dd in cur[b]:
                continue
            if odp.swap_device(a, b, c, dd) != 0:
                continue
            for row, old, new in ((a, b, c), (b, a, dd), (c, dd, a), (dd, c, b)):
                cur[row][list(cur[row]).index(old)] = new
            done += 1
            assert same(odp.set_aspl_resident(), oracle.evaluate(cur, width=w, height=h)), done
    finally:
        odp.finalize_aspl()


def test_tile_graph_follows_a_longer_edge():
    """A device swap that makes the longest edge longer without changing the
    tiles' levels per launch (lmax 9 -> 10..16: one level per launch either
    way) must not replay the launch sequence captured for the shorter edges:
    its offsets and window were built for them.  Three evaluations capture and
    replay the sequence; then swaps add longer edges, each result against the
    oracle."""
    w, h, L = 64, 64, 9
    adj = odp.generate_grid(w, h, 4, L, 1, 21)
    cheb = lambda p, q: max(abs(p // h - q // h), abs(p % h - q % h))
    lmax = max(cheb(v, int(u)) for v in range(w * h) for u in adj[v])
    assert lmax <= L
    init(w, h, 4, 1)
    try:
        odp.set_plan(**TILE_PLAN)
        for _ in range(3):                              # eager, captured, replayed
            assert same(odp.set_aspl(adj), oracle.evaluate(adj, width=w, height=h))
        rng = np.random.default_rng(5)
        cur = adj.copy()
        done = 0
        while done < 3:
            a = int(rng.integers(w * h))
            b = int(cur[a][int(rng.integers(4))])
            c = int(rng.integers(w * h))
            dd = int(cur[c][int(rng.integers(4))])
            if len({a, b, c, dd}) < 4 or c in cur[a] or dd in cur[b]:
                continue
            longest = max(cheb(a, c), cheb(b, dd))
            if not (lmax < longest <= 16):
                continue
            assert odp.swap_device(a, b, c, dd) == 0
            for row, old, new in ((a, b, c), (b, a, dd), (c, dd, a), (dd, c, b)):
                cur[row][list(cur[row]).index(old)] = new
            lmax = longest
            done += 1
            for _ in range(2):
                assert same(odp.set_aspl_resident(), oracle.evaluate(cur, width=w, height=h)), (done, lmax)
    finally:
        odp.finalize_aspl()


def test_tile_autotune_includes_tiles():
    """The autotuner times the tile plan beside the others and checks that all
    agree (it exits otherwise)."""
    w, h, d, length, s = 64, 64, 8, 4, 4
    full = odp.generate_grid(w, h, d, length, s, 3)
    base = odp.grid_base_table(w, h, d, s, full)
    want = oracle.evaluate(base, width=w, height=h, symmetries=s)
    init(w, h, d, s)
    try:
        odp.set_plan(**TILE_PLAN)
        assert odp.autotune(base, 1) >= 2
        assert same(odp.set_aspl(base), want)
    finally:
        odp.finalize_aspl()
