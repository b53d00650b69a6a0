"""ODP_Srand / ODP_Generate_random_* / ODP_Mutate_adjacency_* / ODP_Restore_adjacency_*
against the reference library (utils.c, built into oracle/_ref/odp_ref_gen).

Both sides seed libc rand() and run the same sequence -- generate, convert to
the (base) table, then moves with every third one restored -- and must print
the same edges, undo records, table checksums and final table, byte for byte.
Where the reference build is absent the committed fixture (the same runs'
SHA-256, tests/golden/mutate_vectors.json) is the check.  Invariants that hold
for any seed (degrees kept, symmetry kept, edge lengths within the limit,
a restore undoing its move) are checked beside them.
"""
import ctypes
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import odp_amd as odp

HERE = os.path.dirname(os.path.abspath(__file__))
REF_GEN = os.path.join(HERE, "..", "oracle", "_ref", "odp_ref_gen")
VECTORS = os.path.join(HERE, "golden", "mutate_vectors.json")

# (family, args) as odp_ref_gen takes them: general N D S SEED MOVES | grid W H D LEN S SEED MOVES [grid_s]
CASES = [
    ("general", (24, 3, 1, 5, 30)),
    ("general", (40, 4, 1, 11, 60)),
    ("general", (36, 5, 1, 2, 40)),          # odd degree: the start graph's cross matching
    ("general", (48, 4, 2, 3, 60)),          # self-symmetric (diameter) edges
    ("general", (60, 3, 3, 4, 60)),
    ("general", (64, 6, 4, 9, 80)),
    ("general", (72, 4, 6, 1, 80)),          # s > 4: the turned-edge loop
    ("general", (90, 4, 9, 6, 60)),
    ("grid", (6, 6, 3, 2, 1, 7, 40)),
    ("grid", (7, 5, 4, 3, 1, 3, 40)),        # odd width: vertical pairs, odd grid self-loops
    ("grid", (8, 8, 4, 3, 1, 2, 40, 1)),     # s = 1 moves through ODP_Mutate_adjacency_grid_s
    ("grid", (8, 6, 4, 3, 2, 5, 60)),
    ("grid", (10, 6, 3, 4, 2, 8, 60)),
    ("grid", (10, 10, 4, 4, 4, 1, 80)),
    ("grid", (10, 10, 3, 4, 4, 3, 80)),      # odd base region and odd degree (s = 4)
    ("grid", (6, 5, 3, 3, 2, 4, 60)),        # odd base region and odd degree (s = 2)
    ("grid", (12, 12, 6, 5, 4, 12, 100)),
    # small symmetric graphs, long runs: both edges self-symmetric, two images
    # of one edge (the turned-edge move, and its redraw loop for s > 4)
    ("general", (12, 3, 2, 7, 300)),
    ("general", (20, 3, 2, 8, 300)),
    ("general", (16, 4, 4, 3, 300)),
    ("general", (18, 4, 6, 5, 300)),
    ("general", (24, 4, 8, 2, 300)),
    ("general", (14, 4, 7, 1, 300)),
    ("grid", (4, 4, 3, 3, 2, 9, 300)),
    ("grid", (6, 6, 4, 4, 4, 5, 300)),
    ("grid", (6, 6, 3, 3, 4, 2, 300)),
]


def _lib():
    lib = odp.load()
    P = ctypes.c_void_p
    I = ctypes.c_int
    lib.ODP_Srand.argtypes = [ctypes.c_uint]
    lib.ODP_Generate_random_general_s.argtypes = [I, I, I, P]
    lib.ODP_Generate_random_grid_s.argtypes = [I, I, I, I, I, P]
    lib.ODP_Conv_edge2adjacency_general_s.argtypes = [I, I, I, P, I, P]
    lib.ODP_Conv_edge2adjacency_grid_s.argtypes = [I, I, I, I, P, I, P]
    lib.ODP_Mutate_adjacency_general_s.argtypes = [I, I, P, I, P, P]
    lib.ODP_Mutate_adjacency_grid.argtypes = [I, I, I, P, I, P, P]
    lib.ODP_Mutate_adjacency_grid_s.argtypes = [I, I, I, P, I, I, P, P]
    lib.ODP_Restore_adjacency_general.argtypes = [Restore, P]
    lib.ODP_Restore_adjacency_grid.argtypes = [Restore, P]
    for f in ("ODP_Srand", "ODP_Generate_random_general_s", "ODP_Generate_random_grid_s",
              "ODP_Mutate_adjacency_general_s", "ODP_Mutate_adjacency_grid", "ODP_Mutate_adjacency_grid_s",
              "ODP_Restore_adjacency_general", "ODP_Restore_adjacency_grid"):
        getattr(lib, f).restype = None
    return lib


class Restore(ctypes.Structure):
    _fields_ = [("u", ctypes.c_int * 2), ("v", ctypes.c_int * 2), ("u_d", ctypes.c_int * 2), ("v_d", ctypes.c_int * 2)]


def fnv(a):
    h = 1469598103934665603
    for b in np.ascontiguousarray(a, dtype=np.int32).tobytes():
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def run_ours(family, args, record=None):
    """The sequence of odp_ref_gen on libodp_hip; returns its text (and fills
    `record` with (table before, undo record, table after) per move)."""
    lib = _lib()
    if family == "general":
        nodes, d, s, seed, moves = args
        use_s = 0
    else:
        w, h, d, length, s, seed, moves = args[:7]
        use_s = args[7] if len(args) > 7 else 0
        nodes = w * h
    lines, rows = nodes * d // 2, nodes // s
    edge = np.zeros((lines, 2), np.int32)
    adj = np.zeros((rows, d), np.int32)
    lib.ODP_Srand(seed)
    if family == "general":
        lib.ODP_Generate_random_general_s(nodes, d, s, edge.ctypes.data)
    else:
        lib.ODP_Generate_random_grid_s(w, h, d, length, s, edge.ctypes.data)
    out = [f"E {lines}"] + [f"{a} {b}" for a, b in edge.tolist()]
    if family == "general":
        lib.ODP_Conv_edge2adjacency_general_s(nodes, lines, d, edge.ctypes.data, s, adj.ctypes.data)
    else:
        lib.ODP_Conv_edge2adjacency_grid_s(w, h, lines, d, edge.ctypes.data, s, adj.ctypes.data)
    out.append(f"H {fnv(adj)}")
    for m in range(moves):
        r = Restore()
        before = adj.copy()
        if family == "general":
            lib.ODP_Mutate_adjacency_general_s(nodes, d, None, s, ctypes.byref(r), adj.ctypes.data)
        elif s == 1 and not use_s:
            lib.ODP_Mutate_adjacency_grid(w, h, d, None, length, ctypes.byref(r), adj.ctypes.data)
        else:
            lib.ODP_Mutate_adjacency_grid_s(w, h, d, None, length, s, ctypes.byref(r), adj.ctypes.data)
        out.append("M " + " ".join(str(x) for x in (*r.u, *r.v, *r.u_d, *r.v_d)))
        if record is not None:
            record.append((before, (list(r.u), list(r.v), list(r.u_d), list(r.v_d)), adj.copy()))
        if m % 3 == 2:
            if family == "general":
                lib.ODP_Restore_adjacency_general(r, adj.ctypes.data)
            else:
                lib.ODP_Restore_adjacency_grid(r, adj.ctypes.data)
            out.append("R")
        out.append(f"H {fnv(adj)}")
    out.append("A")
    out += [" ".join(str(x) for x in row) for row in adj.tolist()]
    return "\n".join(out) + "\n"


def case_id(c):
    return c[0] + "-" + "-".join(str(x) for x in c[1])


def ref_text(family, args):
    return subprocess.run([REF_GEN, family, *map(str, args)], capture_output=True, text=True, check=True,
                          timeout=120).stdout


@pytest.mark.parametrize("case", CASES, ids=case_id)
def test_matches_reference(case):
    family, args = case
    ours = run_ours(family, args)
    if os.path.exists(REF_GEN):
        ref = ref_text(family, args)
        if ours != ref:
            a, b = ours.splitlines(), ref.splitlines()
            k = next(i for i in range(min(len(a), len(b))) if a[i] != b[i]) if a[:len(b)] != b[:len(a)] else min(len(a), len(b))
            pytest.fail(f"first difference at line {k}: ours {a[k:k+2]} reference {b[k:k+2]}")
    with open(VECTORS) as f:
        want = json.load(f)[case_id(case)]
    assert hashlib.sha256(ours.encode()).hexdigest() == want


def _expand_general(adj, nodes, s):
    b = nodes // s
    full = np.zeros((nodes, adj.shape[1]), np.int64)
    for p in range(s):
        full[p * b:(p + 1) * b] = (adj + p * b) % nodes
    return full


@pytest.mark.parametrize("case", [c for c in CASES if c[0] == "general"], ids=case_id)
def test_general_invariants(case):
    """Every table along the run describes a d-regular symmetric multigraph:
    u lists v as often as v lists u; each move changes at most 4 base entries;
    a restore returns the table to the state before its move."""
    _, args = case
    nodes, d, s = args[:3]
    rec = []
    run_ours("general", args, rec)
    for m, (before, r, after) in enumerate(rec):
        full = _expand_general(after, nodes, s)
        cnt = {}
        for v in range(nodes):
            for u in full[v]:
                cnt[(v, int(u))] = cnt.get((v, int(u)), 0) + 1
        for (v, u), c in cnt.items():
            assert cnt.get((u, v), 0) == c or u == v, (m, v, u)
        assert int((before != after).sum()) <= 4
        if m % 3 == 2 and m + 1 < len(rec):
            assert np.array_equal(rec[m + 1][0], before), m      # restored


def _rot(v, w, h, s, a):
    x, y = divmod(v, h)
    if s == 2:
        return (w - x - 1) * h + (h - y - 1)
    return {90: y * h + (h - x - 1), 180: (h - x - 1) * h + (h - y - 1), 270: (h - y - 1) * h + x}[a]


@pytest.mark.parametrize("case", [c for c in CASES if c[0] == "grid" and c[1][4] > 1], ids=case_id)
def test_grid_s_invariants(case):
    """Symmetric grids: the expanded edge set is closed under the symmetry's
    turns and no edge is longer than the length limit."""
    _, args = case
    w, h, d, length, s, seed = args[:6]
    lib = _lib()
    nodes = w * h
    edge = np.zeros((nodes * d // 2, 2), np.int32)
    lib.ODP_Srand(seed)
    lib.ODP_Generate_random_grid_s(w, h, d, length, s, edge.ctypes.data)
    es = {tuple(sorted(e)) for e in edge.tolist()}
    for a, b in es:
        assert abs(a // h - b // h) + abs(a % h - b % h) <= length
        for ang in ((180,) if s == 2 else (90, 180, 270)):
            assert tuple(sorted((_rot(a, w, h, s, ang), _rot(b, w, h, s, ang)))) in es
