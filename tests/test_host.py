"""Host-side checks that need no GPU: the C-ABI library loads and exports every
declared symbol, the seeded generators produce what they promise, the graph
utilities reproduce the reference's outputs, and invalid parameters end the
process the way the reference does."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import odp_amd as odp
from goldens import GOLDEN, TEST_GENERAL, TEST_GRID, parse_aspl_txt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("odp.h", "odpx.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b((?:ODP|ODPX)_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = odp.load()
    names = declared_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_cpu_fallback_in_product():
    # the product path never loads the oracle
    for dirpath, _, files in os.walk(os.path.join(ROOT, "odp_amd")):
        for f in files:
            if f.endswith((".py", ".c", ".hip", ".hpp", ".h")):
                assert "oracle" not in open(os.path.join(dirpath, f)).read().lower().replace("oracle/", ""), f


def simple_regular(adj, n, d):
    assert adj.shape == (n, d)
    assert (adj >= 0).all() and (adj < n).all()
    for i in range(n):
        row = adj[i]
        assert len(set(row.tolist())) == d and i not in row
    # symmetric
    pairs = set((i, int(j)) for i in range(n) for j in adj[i])
    assert all((j, i) in pairs for (i, j) in pairs)


def connected(adj):
    n = adj.shape[0]
    seen = np.zeros(n, bool)
    seen[0] = True
    front = [0]
    while front:
        nxt = []
        for v in front:
            for u in adj[v]:
                if u >= 0 and not seen[u]:
                    seen[u] = True
                    nxt.append(int(u))
        front = nxt
    return seen.all()


@pytest.mark.parametrize("n,d,seed", [(10, 3, 1), (100, 4, 2), (1024, 6, 1), (3000, 16, 3), (500, 24, 4)])
def test_generate_regular(n, d, seed):
    a = odp.generate_regular(n, d, seed)
    simple_regular(a, n, d)
    assert connected(a)
    assert (odp.generate_regular(n, d, seed) == a).all()
    assert not (odp.generate_regular(n, d, seed + 1) == a).all()


def test_generate_regular_rejects_odd():
    with pytest.raises(ValueError):
        odp.generate_regular(11, 3, 1)


def rot90(v, w, h):
    x, y = divmod(v, h)
    return y * h + (h - x - 1)


@pytest.mark.parametrize("w,d,L,s,seed", [(16, 4, 3, 4, 1), (24, 8, 4, 4, 2), (20, 6, 4, 2, 3), (12, 3, 2, 1, 4)])
def test_generate_grid(w, d, L, s, seed):
    a = odp.generate_grid(w, w, d, L, s, seed)
    n = w * w
    simple_regular(a, n, d)
    assert connected(a)
    for v in range(n):
        for u in a[v]:
            x0, y0 = divmod(v, w)
            x1, y1 = divmod(int(u), w)
            assert abs(x0 - x1) + abs(y0 - y1) <= L
    edges = set((min(v, int(u)), max(v, int(u))) for v in range(n) for u in a[v])
    if s == 4:
        for (p, q) in edges:
            r = (min(rot90(p, w, w), rot90(q, w, w)), max(rot90(p, w, w), rot90(q, w, w)))
            assert r in edges
    if s == 2:
        for (p, q) in edges:
            r = (min(n - 1 - p, n - 1 - q), max(n - 1 - p, n - 1 - q))
            assert r in edges


def test_generate_general_s():
    import oracle
    for (n, d, s, seed) in ((600, 4, 3, 1), (1024, 6, 4, 2), (960, 5, 6, 3)):
        base = odp.generate_general_s(n, d, s, seed)
        full = oracle.expand(base, nodes=n, symmetries=s)
        simple_regular(full, n, d)


def test_two_opt_keeps_simple_regular():
    a = odp.generate_regular(200, 5, 7)
    rng = odp.Rng(3)
    before = a.copy()
    for _ in range(50):
        sw = odp.two_opt(a, rng)
        assert len(set(sw)) == 4
    simple_regular(a, 200, 5)
    assert (a != before).any()


def test_lbounds_match_reference_outputs():
    exp = parse_aspl_txt(os.path.join(GOLDEN, "general", "ASPL.txt"))
    for name, e in zip(TEST_GENERAL, exp):
        edge, n, d = odp.read_edges_general(os.path.join(GOLDEN, "general", name + ".edges"))
        ld, la = odp.lbounds_general(n, d)
        assert ld == e["low_diameter"] and f"{la:.10f}" == e["low_aspl10"]
    exp = parse_aspl_txt(os.path.join(GOLDEN, "grid", "ASPL.txt"))
    for name, e in zip(TEST_GRID, exp):
        edge, w, h, d, length = odp.read_edges_grid(os.path.join(GOLDEN, "grid", name + ".edges"))
        ld, la = odp.lbounds_grid(w, h, d, length)
        assert ld == e["low_diameter"] and f"{la:.10f}" == e["low_aspl10"]


def test_edge_adjacency_round_trip():
    lib = odp.load()
    edge, n, d = odp.read_edges_general(os.path.join(GOLDEN, "general", "n100d10.edges"))
    adj = odp.edges_to_adjacency_general(n, d, edge)
    back = np.zeros_like(edge)
    lib.ODP_Conv_adjacency2edge_general(n, d, None, adj.ctypes.data, back.ctypes.data)
    assert set(map(tuple, np.sort(back, 1).tolist())) == set(map(tuple, np.sort(edge, 1).tolist()))
    # symmetric grid round trip: base table -> edges -> base table
    edge, w, h, d, _ = odp.read_edges_grid(os.path.join(GOLDEN, "grid", "w6h6d3r2g4.edges"))
    base = odp.edges_to_adjacency_grid(w, h, d, edge, 4)
    back = np.zeros_like(edge)
    lib.ODP_Conv_adjacency2edge_grid_s(w, h, d, None, base.ctypes.data, 4, back.ctypes.data)
    assert set(map(tuple, np.sort(back, 1).tolist())) == set(map(tuple, np.sort(edge, 1).tolist()))


def _run_expect_exit(code, env=None):
    full = dict(os.environ)
    full.update(env or {})
    prog = f"import sys; sys.path.insert(0, {ROOT!r})\nimport odp_amd as o\n" + code
    return subprocess.run([sys.executable, "-c", prog], capture_output=True, text=True, env=full, timeout=120)


def test_error_nodes_not_divisible():
    r = _run_expect_exit("o.init_aspl_general_s(10, 3, None, 4)")
    assert r.returncode == 1 and "must be divisible by symmetries" in r.stderr


def test_bfs_kind_accepted(monkeypatch):
    """ODP_ASPL=BFS selects the BFS kind (3, common.h:41) instead of exiting as
    the reference's GPU flavour does (utils.c:1256-1259); its memory estimate is
    the reference's BFS formula (utils.c:1220-1224, no OpenMP term)."""
    monkeypatch.setenv("ODP_ASPL", "BFS")
    assert odp.get_kind(1000, 4) == 3
    assert odp.get_mem_usage(3, 1000, 4) == pytest.approx((1000 * 13 + 8 * 448 + 4 * 1000 * 4) / 2**20)


def test_error_unknown_kind():
    r = _run_expect_exit("o.init_aspl_general(10, 3)", env={"ODP_ASPL": "FOO"})
    assert r.returncode == 1 and "Unknown ASPL value (FOO)" in r.stderr


def test_error_grid_s_shape():
    r = _run_expect_exit("o.init_aspl_grid_s(6, 8, 3, None, 4)")
    assert r.returncode == 1 and "Must be the same as width" in r.stderr


def test_set_before_init_exits():
    r = _run_expect_exit("import numpy as np; o.set_aspl(np.zeros((2, 1), np.int32))")
    assert r.returncode == 1 and "before ODP_Init_aspl" in r.stderr


REF_KIND = os.path.join(ROOT, "oracle", "_ref", "odp_ref_kind")


def _ref_kinds(shapes, env=None):
    full = dict(os.environ)
    full.pop("ODP_ASPL", None)
    full.update(env or {})
    text = "".join(f"{n} {d} {nd} {s} {g}\n" for (n, d, nd, s, g) in shapes)
    r = subprocess.run([REF_KIND], input=text, capture_output=True, text=True, env=full, timeout=120)
    assert r.returncode == 0, r.stderr
    return [(int(k), float.fromhex(m1), float.fromhex(m2)) for k, m1, m2 in (ln.split() for ln in r.stdout.splitlines())]


@pytest.mark.skipif(not os.path.exists(REF_KIND), reason="reference kind checker not built (needs the reference sources)")
def test_kind_rule_matches_reference(monkeypatch):
    """The kind Init picks without ODP_ASPL, and the memory estimate behind it,
    equal the reference's own rule (utils.c:1206-1266, GPU branch) on both sides
    of the 2048 MB threshold, for every argument the estimate depends on."""
    monkeypatch.delenv("ODP_ASPL", raising=False)
    shapes = []
    for s, grid_s in ((1, 0), (2, 0), (2, 1), (3, 0), (4, 0), (4, 1)):
        nb = int((8 * 2**30 * s) ** 0.5)             # n^2 / (4 s) bytes = 2048 MB
        for d in (3, 16, 64):
            for nd in (0, 1):
                for n in range(nb - 40 * s, nb + 40 * s, s):
                    shapes.append((n - n % s, d, nd, s, grid_s))
    ref = _ref_kinds(shapes)
    kinds = set()
    for (n, d, nd, s, g), (k, m1, m2) in zip(shapes, ref):
        assert odp.get_kind(n, d, nd, s, g) == k, (n, d, nd, s, g)
        assert odp.get_mem_usage(1, n, d, nd, s, g) == m1 and odp.get_mem_usage(2, n, d, nd, s, g) == m2
        kinds.add(k)
    assert kinds == {1, 2}                          # both sides of the threshold were exercised
    # ODP_ASPL overrides the threshold on both sides
    for val, want in (("MATRIX", 1), ("MATRIX_SAVING", 2)):
        monkeypatch.setenv("ODP_ASPL", val)
        few = [shapes[0], shapes[-1]]
        assert [k for k, _, _ in _ref_kinds(few, {"ODP_ASPL": val})] == [want, want]
        assert [odp.get_kind(*sh) for sh in few] == [want, want]
