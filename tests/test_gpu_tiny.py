"""Tiny and odd-sized graphs through the C ABI against the oracle.

Sizes at the edges of the kernels' tiling: n = 2 (a single edge), graphs
smaller than a wave, a block or a 32-row group, one row past a power of two,
degree 1, non-regular paths (num_degrees), and grids from 2 x 2 to 8 x 8 --
each under the default plan, a resident plan and a level plan, evaluated twice
(the second evaluation replays or reuses what the first set up).
"""
import numpy as np
import pytest

import odp_amd as odp
import oracle
from goldens import complete_graph, cycle_graph, path_graph

pytestmark = pytest.mark.gpu

PLANS = [None, dict(plan=odp.PLAN_RESIDENT, wb=1, threads=256),
         dict(plan=odp.PLAN_LEVEL, kw=1, upw=1, seg_major=0, xcd=1)]


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()


def run(adj, want, init, label):
    for p in PLANS:
        init()
        try:
            if p:
                try:
                    odp.set_plan(**p)
                except ValueError:
                    continue                      # the plan does not apply to this graph
            for rep in range(2):
                got = odp.set_aspl(adj)
                assert same(got, want), (label, p, rep, got, want)
        finally:
            odp.finalize_aspl()


@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 9, 16, 31, 33, 64, 65])
def test_tiny_general(n):
    cases = [complete_graph(n)]
    if n >= 3:
        cases.append(cycle_graph(n))
    for d in (2, 3):
        if d < n and (n * d) % 2 == 0:
            cases.append(odp.generate_regular(n, d, 11 * n + d))
    for adj in cases:
        adj = np.ascontiguousarray(adj, np.int32)
        run(adj, oracle.evaluate(adj, nodes=n), lambda: odp.init_aspl_general(n, adj.shape[1]), f"n={n}")
    pa, nd = path_graph(n)
    run(pa, oracle.evaluate(pa, nodes=n, num_degrees=nd), lambda: odp.init_aspl_general(n, pa.shape[1], nd),
        f"path n={n}")


def reference_grid(w, h, d, length, seed):
    """A random grid from the reference-compatible generator (its lattice start
    and 2-opt moves: self-loops and parallel edges can remain)."""
    import ctypes
    lib = odp.load()
    P, I = ctypes.c_void_p, ctypes.c_int
    lib.ODP_Srand.argtypes = [ctypes.c_uint]
    lib.ODP_Generate_random_grid.argtypes = [I, I, I, I, P]
    lib.ODP_Conv_edge2adjacency_grid.argtypes = [I, I, I, I, P, P]
    lines = w * h * d // 2
    edge = np.zeros((lines, 2), np.int32)
    adj = np.zeros((w * h, d), np.int32)
    lib.ODP_Srand(seed)
    lib.ODP_Generate_random_grid(w, h, d, length, edge.ctypes.data)
    lib.ODP_Conv_edge2adjacency_grid(w, h, lines, d, edge.ctypes.data, adj.ctypes.data)
    return adj


# (width, height, degree, length); every one of them generates in milliseconds
# (some other tiny shapes, e.g. 3 x 5 with degree 2 and length 1, have no
# connected graph and the generator, like the reference's, would not return)
GRIDS = [(2, 2, 2, 1), (2, 2, 4, 3), (2, 3, 2, 1), (2, 3, 3, 2), (2, 3, 4, 3), (3, 5, 4, 3), (5, 3, 4, 3),
         (3, 4, 3, 2), (4, 4, 3, 2), (6, 6, 2, 2), (7, 2, 2, 1), (7, 2, 4, 3), (8, 8, 2, 1), (8, 8, 3, 2),
         (8, 8, 4, 3)]


@pytest.mark.parametrize("w,h,d,length", GRIDS)
def test_tiny_grid(w, h, d, length):
    adj = reference_grid(w, h, d, length, 31 * w + 7 * h + d)
    run(adj, oracle.evaluate(adj, width=w, height=h), lambda: odp.init_aspl_grid(w, h, d, None),
        f"grid {w}x{h} d={d}")
