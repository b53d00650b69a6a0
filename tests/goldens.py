"""Readers for the golden fixtures (the reference's own sample graphs and
expected outputs under tests/golden/{general,grid}, plus the vectors written by
tests/golden/make_vectors.py) and the graph constructions the tests share."""
import json
import os
import re

import numpy as np

import odp_amd as odp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# the graphs sample/Makefile:31-32 feeds to `make test`, in ASPL.txt order
TEST_GENERAL = ["n12d3", "n16d4", "n72d4", "n100d10", "n256d5", "n256d10"]
TEST_GRID = ["w4h4d3r2", "w5h5d4r2", "w10h10d6r3", "w10h10d8r4"]


def parse_aspl_txt(path):
    """ASPL.txt -> list of dicts (diameter, low_diameter, aspl10, sum, low_aspl10)."""
    out = []
    cur = {}
    for line in open(path):
        if line.startswith("Diameter     ="):
            cur["diameter"] = int(line.split("=")[1])
        elif line.startswith("Diameter Gap"):
            cur["low_diameter"] = int(re.search(r"- (\d+)\)", line).group(1))
        elif line.startswith("ASPL         ="):
            m = re.search(r"= ([0-9.]+) \((\d+)/(\d+)\)", line)
            cur["aspl10"], cur["sum"] = m.group(1), int(m.group(2))
        elif line.startswith("ASPL Gap"):
            cur["low_aspl10"] = re.search(r"- ([0-9.]+)\)", line).group(1)
            out.append(cur)
            cur = {}
    return out


def general_fixture(name, symmetries=1):
    edge, n, d = odp.read_edges_general(os.path.join(GOLDEN, "general", name + ".edges"))
    return n, d, odp.edges_to_adjacency_general(n, d, edge, symmetries)


def grid_fixture(name, symmetries=1):
    edge, w, h, d, length = odp.read_edges_grid(os.path.join(GOLDEN, "grid", name + ".edges"))
    return w, h, d, length, odp.edges_to_adjacency_grid(w, h, d, edge, symmetries)


def symmetric_fixtures():
    return json.load(open(os.path.join(GOLDEN, "symmetric_fixtures.json")))


def vectors():
    path = os.path.join(GOLDEN, "vectors.json")
    return json.load(open(path))["cases"] if os.path.exists(path) else []


def build_case(c):
    """Regenerates the graph of one vectors.json case -> (kwargs for evaluate, adjacency)."""
    k = c["graph"]
    if k == "regular":
        adj = odp.generate_regular(c["nodes"], c["degree"], c["seed"])
        rng = odp.Rng(c["seed"] + 1000003)
        for _ in range(c.get("two_opt", 0)):
            odp.two_opt(adj, rng)
        return dict(nodes=c["nodes"]), adj
    if k == "general_s":
        adj = odp.generate_general_s(c["nodes"], c["degree"], c["symmetries"], c["seed"])
        return dict(nodes=c["nodes"], symmetries=c["symmetries"]), adj
    if k == "grid":
        w, h, s = c["width"], c["height"], c["symmetries"]
        full = odp.generate_grid(w, h, c["degree"], c["length"], s, c["seed"])
        return dict(width=w, height=h, symmetries=s), odp.grid_base_table(w, h, c["degree"], s, full)
    if k == "batch":
        # graph `index` of the batched workload (bench.c4_params / c4_graph)
        import bench
        n, d = bench.c4_params(c["index"] + 1, c["batch_seed"])[c["index"]]
        return dict(nodes=n), bench.c4_graph(c["index"], n, d, c["batch_seed"])
    if k == "chain":
        # served chain `chain` (bench.c5_chain) after `swaps` 2-opt moves drawn from Rng(rng)
        import bench
        adj = bench.c5_chain(c["chain"])
        rng = odp.Rng(c["rng"])
        for _ in range(c["swaps"]):
            odp.two_opt(adj, rng)
        return dict(nodes=adj.shape[0]), adj
    raise ValueError(k)


def case_nodes(c):
    """Vertex count of a vectors.json case."""
    if c["graph"] == "batch":
        import bench
        return bench.c4_params(c["index"] + 1, c["batch_seed"])[c["index"]][0]
    if c["graph"] == "chain":
        import bench
        return bench.C5_NODES[c["chain"] % 3]
    return c.get("nodes") or c["width"] * c["height"]


C4_SAMPLE = [0, 59, 63, 341, 618, 771, 837, 868]   # n from 1036 to 16332, every d of the workload
C5_SAMPLE = list(range(9))                          # chains 0..8: every (n, d) class once


# ---- closed-form graphs ------------------------------------------------------

def complete_graph(n):
    return np.array([[j for j in range(n) if j != i] for i in range(n)], np.int32)


def cycle_graph(n):
    return np.array([[(i - 1) % n, (i + 1) % n] for i in range(n)], np.int32)


def petersen():
    outer = [[(i + 1) % 5, (i - 1) % 5, i + 5] for i in range(5)]
    inner = [[5 + (i + 2) % 5, 5 + (i - 2) % 5, i] for i in range(5)]
    return np.array(outer + inner, np.int32)


def hypercube(k):
    n = 1 << k
    return np.array([[i ^ (1 << b) for b in range(k)] for i in range(n)], np.int32)


def path_graph(n):
    """non-regular: degree-2 table padded with -1, plus num_degrees"""
    adj = np.full((n, 2), -1, np.int32)
    nd = np.zeros(n, np.int32)
    for i in range(n):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < n]
        adj[i, :len(nb)] = nb
        nd[i] = len(nb)
    return adj, nd


def two_triangles():
    return np.array([[1, 2], [0, 2], [0, 1], [4, 5], [3, 5], [3, 4]], np.int32)


def cycle_sum(n):
    """sum over unordered pairs of cycle distances"""
    return n * sum(min(k, n - k) for k in range(1, n)) // 2
