"""One process's share of an evaluation (ODPX_Set_aspl_shard), as the
reference's MPI+CUDA flavour divides the work (aspl_mpi_cuda.cu:44-123): every
rank of P = 2, 3 and 8 runs one after another on the one GPU, each rank's
partial equals the oracle's restatement of that rank's loop, and the partials
combined as aspl_mpi_cuda.cu:82-87 does equal the whole evaluation and the
reference's values."""
import json
import os

import numpy as np
import pytest

import odp_amd as odp
import oracle
from goldens import GOLDEN, build_case, cycle_graph, vectors

pytestmark = pytest.mark.gpu

LEV, RES = odp.PLAN_LEVEL, odp.PLAN_RESIDENT


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()


def vector(**match):
    for c in vectors():
        if all(c.get(k) == v for k, v in match.items()):
            return c, (c["diameter"], c["sum"], float.fromhex(c["aspl"]))
    raise KeyError(match)


def init(adj, nodes=None, width=None, height=None, symmetries=1, num_degrees=None):
    if width is None:
        (odp.init_aspl_general(nodes, adj.shape[1], num_degrees) if symmetries == 1
         else odp.init_aspl_general_s(nodes, adj.shape[1], num_degrees, symmetries))
    else:
        (odp.init_aspl_grid(width, height, adj.shape[1], num_degrees) if symmetries == 1
         else odp.init_aspl_grid_s(width, height, adj.shape[1], num_degrees, symmetries))


@pytest.mark.parametrize("kind", ["MATRIX", "MATRIX_SAVING"])
def test_random65536_shards(kind, monkeypatch):
    monkeypatch.setenv("ODP_ASPL", kind)
    c, want = vector(graph="regular", nodes=65536, degree=16, seed=1)
    _, adj = build_case(c)
    ref = json.load(open(os.path.join(GOLDEN, "shard_vectors.json")))["cases"]
    odp.init_aspl_general(65536, 16)
    try:
        assert same(odp.set_aspl(adj), want)
        for procs in (2, 3, 8):
            print(f"random65536 {kind} procs={procs}", flush=True)
            parts = [odp.set_aspl_shard(r, procs) for r in range(procs)]
            exp = next(x["ranks"] for x in ref if x["kind"] == kind and x["procs"] == procs)
            assert [[p[0], p[1]] for p in parts] == exp
            assert all(p[2] == 0 for p in parts)
            assert same(odp.combine_shards(65536, parts), want)
        assert same(odp.set_aspl_resident(), want)      # the whole evaluation is unchanged afterwards
    finally:
        odp.finalize_aspl()


def test_grid256_shards(monkeypatch):
    monkeypatch.setenv("ODP_ASPL", "MATRIX_SAVING")
    c, want = vector(graph="grid", width=256, height=256, degree=8, length=4, symmetries=4, seed=1)
    _, base = build_case(c)
    odp.init_aspl_grid_s(256, 256, 8, None, 4)
    try:
        assert same(odp.set_aspl(base), want)
        for procs in (2, 3, 8):
            print(f"grid256 procs={procs}", flush=True)
            parts = [odp.set_aspl_shard(r, procs) for r in range(procs)]
            assert same(odp.combine_shards(65536, parts), want)
        assert same(odp.set_aspl_resident(), want)
    finally:
        odp.finalize_aspl()


def _cases():
    out = []
    g = odp.generate_regular(5000, 5, 3)
    out.append((g, dict(nodes=5000)))
    out.append((odp.generate_general_s(4096, 6, 4, 5), dict(nodes=4096, symmetries=4)))
    for w, s in ((64, 4), (40, 2)):
        full = odp.generate_grid(w, w, 6, 3, s, 9)
        out.append((odp.grid_base_table(w, w, 6, s, full), dict(width=w, height=w, symmetries=s)))
    dis = np.concatenate([cycle_graph(300), cycle_graph(200) + 300]).astype(np.int32)
    out.append((dis, dict(nodes=500)))
    return out


_oracle_parts = {}


def oracle_part(ci, adj, r, procs, okind, kw):
    key = (ci, r, procs, okind)
    if key not in _oracle_parts:
        _oracle_parts[key] = oracle.evaluate_shard(adj, r, procs, kind=okind, **kw)
    return _oracle_parts[key]


PLANS = [None, dict(plan=LEV, kw=4, upw=4, seg_major=0, xcd=1, direct=0),
         dict(plan=LEV, kw=16, upw=8, seg_major=0, xcd=1, direct=-1, parts=4), dict(plan=RES, wb=2, threads=1024)]


@pytest.mark.parametrize("kind", ["MATRIX", "MATRIX_SAVING"])
@pytest.mark.parametrize("plan", range(len(PLANS)))
def test_shards_equal_oracle_ranks(kind, plan, monkeypatch):
    """Every rank's partial equals the oracle's restatement of that rank's loop
    (general, circulant-symmetric, 4- and 2-fold symmetric grids, and a
    disconnected graph), and the partials combine to the whole evaluation."""
    monkeypatch.setenv("ODP_ASPL", kind)
    okind = oracle.MATRIX if kind == "MATRIX" else oracle.MATRIX_SAVING
    for ci, (adj, kw) in enumerate(_cases()):
        n = kw.get("nodes") or kw["width"] * kw["height"]
        init(adj, **kw)
        try:
            if PLANS[plan]:
                odp.set_plan(**PLANS[plan])
            whole = odp.set_aspl(adj)
            for procs in (1, 2, 3, 8):
                parts = [odp.set_aspl_shard(r, procs) for r in range(procs)]
                for r, p in enumerate(parts):
                    ex, dia = oracle_part(ci, adj, r, procs, okind, kw)
                    if dia > n:                        # a block that never completes
                        assert p[1] == n + 1 and p[2] == 1, (kw, procs, r)
                    else:
                        assert (p[0], p[1], p[2]) == (ex, dia, 0), (kw, procs, r)
                assert same(odp.combine_shards(n, parts), whole), (kw, procs)
            assert same(odp.set_aspl_resident(), whole)
        finally:
            odp.finalize_aspl()


def test_shard_rank_beyond_the_blocks():
    """More ranks than MATRIX blocks: the ranks without a block hold the
    reference's starting values (excess 0, diameter 1); invalid ranks refuse."""
    adj = odp.generate_regular(200, 4, 1)           # 4 words: at most 4 MATRIX blocks
    os.environ["ODP_ASPL"] = "MATRIX"
    try:
        odp.init_aspl_general(200, 4)
        whole = odp.set_aspl(adj)
        parts = [odp.set_aspl_shard(r, 8) for r in range(8)]
        assert parts[4:] == [(0, 1, 0)] * 4
        assert same(odp.combine_shards(200, parts), whole)
        for bad in ((8, 8), (-1, 2), (0, 0)):
            with pytest.raises(ValueError):
                odp.set_aspl_shard(*bad)
    finally:
        os.environ.pop("ODP_ASPL", None)
        odp.finalize_aspl()
