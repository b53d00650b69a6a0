"""odp_amd -- Python host mirror of libodp_hip, the MI355X (gfx950) drop-in for
mnakao/ODP's ASPL/diameter evaluator.

The product is the C-ABI library ``libodp_hip.so`` (include/odp.h,
include/odpx.h).  This module binds it with ctypes and mirrors the reference's
C API names (ODP_Init_aspl_general -> init_aspl_general, ...), argument
meaning and error behaviour (invalid parameters end the process, as the
reference's ERROR() does, common.h:32).  There is no CPU fallback: importing
works without a GPU, but every evaluation runs the HIP kernels, and a missing
library raises at load time.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libodp_hip.so")

PLAN_AUTO, PLAN_RESIDENT, PLAN_LEVEL = 0, 1, 2


class Plan(ctypes.Structure):
    """ODPX_Plan (include/odpx.h)."""
    _fields_ = [("plan", ctypes.c_int), ("wb", ctypes.c_int), ("threads", ctypes.c_int),
                ("kw", ctypes.c_int), ("upw", ctypes.c_int), ("seg_major", ctypes.c_int),
                ("xcd", ctypes.c_int), ("group", ctypes.c_int), ("direct", ctypes.c_int),
                ("batch", ctypes.c_int), ("tile", ctypes.c_int), ("parts", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class AnnealResult(ctypes.Structure):
    """ODPX_AnnealResult (include/odpx.h)."""
    _fields_ = [("steps", ctypes.c_longlong), ("accepted", ctypes.c_longlong),
                ("best_diameter", ctypes.c_int), ("best_sum", ctypes.c_long), ("best_ASPL", ctypes.c_double),
                ("diameter", ctypes.c_int), ("ASPL", ctypes.c_double), ("optimum", ctypes.c_int),
                ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Restore(ctypes.Structure):
    """ODP_Restore (include/odp.h): the undo record of one 2-edge swap."""
    _fields_ = [("u", ctypes.c_int * 2), ("v", ctypes.c_int * 2), ("u_d", ctypes.c_int * 2), ("v_d", ctypes.c_int * 2)]


class Stats(ctypes.Structure):
    """ODPX_Stats (include/odpx.h)."""
    _fields_ = [("evals", ctypes.c_longlong), ("eval_ms", ctypes.c_double),
                ("level_launches", ctypes.c_longlong), ("level_ms", ctypes.c_double),
                ("level_bytes", ctypes.c_double), ("level_min_bytes", ctypes.c_double),
                ("idle_launches", ctypes.c_longlong), ("idle_ms", ctypes.c_double),
                ("resident_launches", ctypes.c_longlong), ("resident_ms", ctypes.c_double),
                ("resident_bytes", ctypes.c_double),
                ("other_launches", ctypes.c_longlong), ("other_ms", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Result(ctypes.Structure):
    """ODPX_Result (include/odpx.h)."""
    _fields_ = [("id", ctypes.c_longlong), ("chain", ctypes.c_int), ("status", ctypes.c_int),
                ("diameter", ctypes.c_int), ("sum", ctypes.c_long), ("ASPL", ctypes.c_double),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None

_I = ctypes.c_int
_P = ctypes.c_void_p
_L = ctypes.c_long
_D = ctypes.c_double
_U64 = ctypes.c_uint64


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def load():
    """Loads libodp_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libodp_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    ip = ctypes.POINTER(ctypes.c_int)
    _sig(lib, "ODP_Init_aspl_general", None, [_I, _I, _P])
    _sig(lib, "ODP_Init_aspl_general_s", None, [_I, _I, _P, _I])
    _sig(lib, "ODP_Init_aspl_grid", None, [_I, _I, _I, _P])
    _sig(lib, "ODP_Init_aspl_grid_s", None, [_I, _I, _I, _P, _I])
    _sig(lib, "ODP_Set_aspl", None, [_P, ip, ctypes.POINTER(_L), ctypes.POINTER(_D)])
    _sig(lib, "ODP_Finalize_aspl", None, [])
    _sig(lib, "ODP_Set_lbounds_general", None, [_I, _I, ip, ctypes.POINTER(_D)])
    _sig(lib, "ODP_Set_lbounds_grid", None, [_I, _I, _I, _I, ip, ctypes.POINTER(_D)])
    _sig(lib, "ODP_Conv_edge2adjacency_general", None, [_I, _I, _I, _P, _P])
    _sig(lib, "ODP_Conv_edge2adjacency_grid", None, [_I, _I, _I, _I, _P, _P])
    _sig(lib, "ODP_Conv_edge2adjacency_general_s", None, [_I, _I, _I, _P, _I, _P])
    _sig(lib, "ODP_Conv_edge2adjacency_grid_s", None, [_I, _I, _I, _I, _P, _I, _P])
    _sig(lib, "ODP_Conv_adjacency2edge_general", None, [_I, _I, _P, _P, _P])
    _sig(lib, "ODP_Conv_adjacency2edge_general_s", None, [_I, _I, _P, _P, _I, _P])
    _sig(lib, "ODP_Conv_adjacency2edge_grid_s", None, [_I, _I, _I, _P, _P, _I, _P])
    _sig(lib, "ODP_Get_lines", _I, [ctypes.c_char_p])
    _sig(lib, "ODP_Get_nodes", _I, [_I, _P])
    _sig(lib, "ODP_Get_degree", _I, [_I, _I, _P])
    _sig(lib, "ODP_Get_length", _I, [_I, _I, _P])
    _sig(lib, "ODP_Read_edge_general", None, [ctypes.c_char_p, _P])
    _sig(lib, "ODP_Read_edge_grid", None, [ctypes.c_char_p, ip, ip, _P])
    _sig(lib, "ODP_Set_degrees", None, [_I, _I, _P, _P])
    _sig(lib, "ODPX_Rng_seed", None, [_U64, _P])
    _sig(lib, "ODPX_Rng_next", _U64, [_P])
    _sig(lib, "ODPX_Generate_regular", _I, [_I, _I, _U64, _P])
    _sig(lib, "ODPX_Generate_grid", _I, [_I, _I, _I, _I, _I, _U64, _I, _P])
    _sig(lib, "ODPX_Grid_base_table", _I, [_I, _I, _I, _I, _P, _P])
    _sig(lib, "ODPX_Generate_general_s", _I, [_I, _I, _I, _U64, _P])
    _sig(lib, "ODPX_Swap_edges", _I, [_I, _I, _P, _I, _I, _I, _I])
    _sig(lib, "ODPX_Two_opt", _I, [_I, _I, _P, _P, _P])
    _sig(lib, "ODPX_Get_plan", _I, [ctypes.POINTER(Plan)])
    _sig(lib, "ODPX_Set_plan", _I, [ctypes.POINTER(Plan)])
    _sig(lib, "ODPX_Autotune", _I, [_P, _I])
    _sig(lib, "ODPX_Upload_adjacency", None, [_P])
    _sig(lib, "ODPX_Set_aspl_resident", None, [ip, ctypes.POINTER(_L), ctypes.POINTER(_D)])
    _sig(lib, "ODPX_Swap_device", _I, [_I, _I, _I, _I])
    _sig(lib, "ODPX_Set_aspl_shard", _I, [_I, _I, ctypes.POINTER(ctypes.c_longlong), ip, ip])
    _sig(lib, "ODPX_Set_aspl_noalloc", _I, [_I, _I, _P, ip, ctypes.POINTER(_L), ctypes.POINTER(_D)])
    _sig(lib, "ODPX_Get_kind", _I, [_I, _I, _I, _I, _I])
    _sig(lib, "ODPX_Get_mem_usage", _D, [_I, _I, _I, _I, _I, _I])
    _sig(lib, "ODPX_Anneal", _I, [_P, ctypes.c_longlong, _D, _D, _I, _I, _U64, ctypes.POINTER(AnnealResult)])
    _sig(lib, "ODPX_Mutate", _I, [_I, _I, _I, _I, _I, _I, _P, _P, ctypes.POINTER(Restore)])
    _sig(lib, "ODP_Restore_adjacency_general", None, [Restore, _P])
    _sig(lib, "ODP_Restore_adjacency_grid", None, [Restore, _P])
    _sig(lib, "ODPX_Stats_enable", None, [_I])
    _sig(lib, "ODPX_Stats_reset", None, [])
    _sig(lib, "ODPX_Stats_get", None, [ctypes.POINTER(Stats)])
    batch_args = [_I, _P, _P, _P, _P, _P, _P, _P]
    _sig(lib, "ODPX_Batch_eval", _I, batch_args)
    _sig(lib, "ODPX_Batch_load", _I, [_I, _P, _P, _P, _P])
    _sig(lib, "ODPX_Batch_run", _I, [_P, _P, _P])
    _sig(lib, "ODPX_Loop_run", _I, [_I, _P, _P, _P])
    _sig(lib, "ODPX_Loop_eval", _I, [_I, _P, _P, _P, _P, _I, _P, _P, _P])
    _sig(lib, "ODPX_Batch_release", None, [])
    _sig(lib, "ODPX_Batch_get_policy", None, [_P, _P])
    _sig(lib, "ODPX_Batch_set_policy", _I, [_P, _I])
    _sig(lib, "ODPX_Batch_autotune", _I, [_I])
    _sig(lib, "ODPX_Serve_start", _I, [_I, _I, _D])
    _sig(lib, "ODPX_Serve_add_chain", _I, [_I, _I, _P])
    _sig(lib, "ODPX_Serve_submit", ctypes.c_longlong, [_I, _I, _P])
    _sig(lib, "ODPX_Serve_poll", _I, [ctypes.POINTER(Result), _I])
    _sig(lib, "ODPX_Serve_wait", _I, [ctypes.c_longlong, ctypes.POINTER(Result)])
    _sig(lib, "ODPX_Serve_drain", None, [])
    _sig(lib, "ODPX_Serve_stop", None, [])
    _lib = lib
    return lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---- the ODP_* lifecycle ---------------------------------------------------

class _Keep:
    adjacency = None


def init_aspl_general(nodes, degree, num_degrees=None):
    """ODP_Init_aspl_general (aspl_cuda.cu:159-162)."""
    nd = None if num_degrees is None else _i32(num_degrees)
    load().ODP_Init_aspl_general(nodes, degree, _ptr(nd))


def init_aspl_general_s(nodes, degree, num_degrees, symmetries):
    """ODP_Init_aspl_general_s (aspl_cuda.cu:164-179); num_degrees has nodes/symmetries entries."""
    nd = None if num_degrees is None else _i32(num_degrees)
    load().ODP_Init_aspl_general_s(nodes, degree, _ptr(nd), symmetries)


def init_aspl_grid(width, height, degree, num_degrees=None):
    """ODP_Init_aspl_grid (aspl_cuda.cu:181-186)."""
    nd = None if num_degrees is None else _i32(num_degrees)
    load().ODP_Init_aspl_grid(width, height, degree, _ptr(nd))


def init_aspl_grid_s(width, height, degree, num_degrees, symmetries):
    """ODP_Init_aspl_grid_s (aspl_cuda.cu:188-224)."""
    nd = None if num_degrees is None else _i32(num_degrees)
    load().ODP_Init_aspl_grid_s(width, height, degree, _ptr(nd), symmetries)


def set_aspl(adjacency):
    """ODP_Set_aspl (aspl_cuda.cu:242-263) -> (diameter, sum, ASPL)."""
    a = _i32(adjacency)
    d, s, x = ctypes.c_int(), ctypes.c_long(), ctypes.c_double()
    load().ODP_Set_aspl(_ptr(a), ctypes.byref(d), ctypes.byref(s), ctypes.byref(x))
    return d.value, s.value, x.value


def finalize_aspl():
    """ODP_Finalize_aspl (aspl_cuda.cu:226-240)."""
    load().ODP_Finalize_aspl()


def upload_adjacency(adjacency):
    a = _i32(adjacency)
    load().ODPX_Upload_adjacency(_ptr(a))


def set_aspl_resident():
    """Evaluates the adjacency already on the device."""
    d, s, x = ctypes.c_int(), ctypes.c_long(), ctypes.c_double()
    load().ODPX_Set_aspl_resident(ctypes.byref(d), ctypes.byref(s), ctypes.byref(x))
    return d.value, s.value, x.value


def swap_device(a, b, c, d):
    return load().ODPX_Swap_device(a, b, c, d)


def set_aspl_shard(rank, nranks):
    """ODPX_Set_aspl_shard: one rank's share of the resident adjacency's evaluation
    -> (excess, diameter, disconnected), as the reference's MPI+CUDA rank holds it."""
    ex, d, dc = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_int()
    if load().ODPX_Set_aspl_shard(rank, nranks, ctypes.byref(ex), ctypes.byref(d), ctypes.byref(dc)) != 0:
        raise ValueError(f"invalid shard {rank} of {nranks}")
    return ex.value, d.value, dc.value


def combine_shards(nodes, partials):
    """The reference's combine (aspl_mpi_cuda.cu:82-87) of (excess, diameter, _)
    partials -> (diameter, sum, ASPL) with the sentinels of aspl_cuda.cu:256-260."""
    d = max(p[1] for p in partials)
    total = sum(p[0] for p in partials) + nodes * (nodes - 1)
    if d > nodes:
        return 2**31 - 1, 2**63 - 1, float.fromhex("0x1.fffffffffffffp+1023")
    return d, int(total / 2.0), total / ((nodes - 1.0) * nodes)


def get_kind(nodes, degree, has_num_degrees=False, symmetries=1, grid_s=False):
    """The algorithm kind ODP_Init_aspl_* picks (1 MATRIX, 2 MATRIX_SAVING, 3 BFS); host only."""
    return load().ODPX_Get_kind(nodes, degree, int(bool(has_num_degrees)), symmetries, int(bool(grid_s)))


def get_mem_usage(kind, nodes, degree, has_num_degrees=False, symmetries=1, grid_s=False):
    """The memory estimate (MB) behind that choice and the ODP_PROFILE report; host only."""
    return load().ODPX_Get_mem_usage(kind, nodes, degree, int(bool(has_num_degrees)), symmetries, int(bool(grid_s)))


def get_plan():
    p = Plan()
    if load().ODPX_Get_plan(ctypes.byref(p)) != 0:
        raise RuntimeError("no active evaluator")
    return p.as_dict()


def set_plan(**kw):
    p = Plan(**{k: int(v) for k, v in kw.items()})
    if load().ODPX_Set_plan(ctypes.byref(p)) != 0:
        raise ValueError(f"invalid plan {kw}")


def autotune(adjacency, repeats=2):
    a = _i32(adjacency)
    return load().ODPX_Autotune(_ptr(a), repeats)


def stats_enable(on=True, live=False):
    """Kernel timing (HIP events around every launch; timed evaluations launch
    eagerly): live=False runs every level whole, live=True keeps the launch
    sequence of untimed evaluations (count-only, listing and sparse last levels)."""
    load().ODPX_Stats_enable((2 if live else 1) if on else 0)


def stats_reset():
    load().ODPX_Stats_reset()


def stats():
    s = Stats()
    load().ODPX_Stats_get(ctypes.byref(s))
    return s.as_dict()


# ---- batched evaluation -----------------------------------------------------

def pack(adjacencies):
    """Packs [int[n_g][d_g]] tables -> (nodes, degrees, offsets, packed)."""
    nodes = np.array([a.shape[0] for a in adjacencies], dtype=np.int32)
    degrees = np.array([a.shape[1] for a in adjacencies], dtype=np.int32)
    sizes = nodes.astype(np.int64) * degrees
    offsets = np.zeros(len(adjacencies), dtype=np.int64)
    if len(adjacencies) > 1:
        offsets[1:] = np.cumsum(sizes)[:-1]
    packed = np.concatenate([_i32(a).ravel() for a in adjacencies]) if adjacencies else np.zeros(0, np.int32)
    return nodes, degrees, offsets, packed


def _batch(fn, packed_graphs, *extra):
    nodes, degrees, offsets, packed = packed_graphs
    g = len(nodes)
    d = np.zeros(g, np.int32)
    s = np.zeros(g, np.int64)
    x = np.zeros(g, np.float64)
    fn(g, _ptr(nodes), _ptr(degrees), _ptr(offsets), _ptr(packed), *extra, _ptr(d), _ptr(s), _ptr(x))
    return d, s, x


def batch_eval(packed_graphs):
    """ODPX_Batch_eval: every graph in one persistent launch where LDS allows."""
    return _batch(load().ODPX_Batch_eval, packed_graphs)


def loop_eval(packed_graphs, plan=PLAN_LEVEL):
    """ODPX_Loop_eval: one single-graph evaluation per graph."""
    return _batch(load().ODPX_Loop_eval, packed_graphs, plan)


def batch_load(packed_graphs):
    """ODPX_Batch_load: uploads the graphs and builds their tables (kept on the device)."""
    nodes, degrees, offsets, packed = packed_graphs
    load().ODPX_Batch_load(len(nodes), _ptr(nodes), _ptr(degrees), _ptr(offsets), _ptr(packed))
    return len(nodes)


def batch_run(ngraphs):
    """ODPX_Batch_run on the loaded graphs -> (diameter[], sum[], ASPL[])."""
    d = np.zeros(ngraphs, np.int32)
    s = np.zeros(ngraphs, np.int64)
    x = np.zeros(ngraphs, np.float64)
    load().ODPX_Batch_run(_ptr(d), _ptr(s), _ptr(x))
    return d, s, x


def loop_run(ngraphs, plan=PLAN_LEVEL):
    """ODPX_Loop_run on the loaded graphs (one single-graph evaluation each)."""
    d = np.zeros(ngraphs, np.int32)
    s = np.zeros(ngraphs, np.int64)
    x = np.zeros(ngraphs, np.float64)
    load().ODPX_Loop_run(plan, _ptr(d), _ptr(s), _ptr(x))
    return d, s, x


def batch_release():
    load().ODPX_Batch_release()


def batch_policy():
    """ODPX_Batch_get_policy -> (widest block per size class [5], workgroups per CU)."""
    caps = np.zeros(5, np.int32)
    pc = np.zeros(1, np.int32)
    load().ODPX_Batch_get_policy(_ptr(caps), _ptr(pc))
    return [int(c) for c in caps], int(pc[0])


def set_batch_policy(wbcap, per_cu=2):
    caps = np.ascontiguousarray(wbcap, dtype=np.int32)
    if caps.shape != (5,) or load().ODPX_Batch_set_policy(_ptr(caps), int(per_cu)) != 0:
        raise ValueError(f"invalid batch policy {wbcap} {per_cu}")


def batch_autotune(repeats=2):
    """ODPX_Batch_autotune on the loaded batch; returns the number of candidates timed."""
    return load().ODPX_Batch_autotune(int(repeats))


# ---- seeded graphs ------------------------------------------------------------

def generate_regular(nodes, degree, seed):
    adj = np.empty((nodes, degree), np.int32)
    rc = load().ODPX_Generate_regular(nodes, degree, seed, _ptr(adj))
    if rc != 0:
        raise ValueError(f"ODPX_Generate_regular({nodes}, {degree}, {seed}) failed: {rc}")
    return adj


def generate_grid(width, height, degree, length, symmetries, seed, iters_per_edge=10):
    adj = np.empty((width * height, degree), np.int32)
    rc = load().ODPX_Generate_grid(width, height, degree, length, symmetries, seed, iters_per_edge, _ptr(adj))
    if rc != 0:
        raise ValueError(f"ODPX_Generate_grid failed: {rc}")
    return adj


def grid_base_table(width, height, degree, symmetries, full):
    full = _i32(full)
    base = np.empty((width * height // symmetries, degree), np.int32)
    if load().ODPX_Grid_base_table(width, height, degree, symmetries, _ptr(full), _ptr(base)) != 0:
        raise ValueError("ODPX_Grid_base_table failed")
    return base


def generate_general_s(nodes, degree, symmetries, seed):
    base = np.empty((nodes // symmetries, degree), np.int32)
    rc = load().ODPX_Generate_general_s(nodes, degree, symmetries, seed, _ptr(base))
    if rc != 0:
        raise ValueError(f"ODPX_Generate_general_s failed: {rc}")
    return base


class Rng:
    """splitmix64-seeded xoshiro256** state shared with the C generators."""

    def __init__(self, seed):
        self.state = np.zeros(4, np.uint64)
        load().ODPX_Rng_seed(seed, _ptr(self.state))

    def next(self):
        return int(load().ODPX_Rng_next(_ptr(self.state)))


def swap_edges(adjacency, a, b, c, d):
    """a-b, c-d -> a-c, b-d in place on a full int[n][d] table (ODPX_Swap_edges)."""
    assert adjacency.dtype == np.int32 and adjacency.flags.c_contiguous
    n, deg = adjacency.shape
    if load().ODPX_Swap_edges(n, deg, _ptr(adjacency), a, b, c, d) != 0:
        raise ValueError("ODPX_Swap_edges: an edge is missing")


def two_opt(adjacency, rng):
    """One valid 2-opt move in place on a full int[n][d] table; returns (a, b, c, d)."""
    assert adjacency.dtype == np.int32 and adjacency.flags.c_contiguous
    sw = np.zeros(4, np.int32)
    n, d = adjacency.shape
    if load().ODPX_Two_opt(n, d, _ptr(adjacency), _ptr(rng.state), _ptr(sw)) != 0:
        raise ValueError("ODPX_Two_opt found no valid move")
    return tuple(int(v) for v in sw)


def mutate(adjacency, rng, nodes=None, width=None, height=None, length=0, symmetries=1):
    """ODPX_Mutate in place on a base table: one 2-opt move of the reference's kind
    drawn from rng -> the ODP_Restore record."""
    assert adjacency.dtype == np.int32 and adjacency.flags.c_contiguous
    r = Restore()
    fam, nw = (0, nodes) if width is None else (1, width)
    if load().ODPX_Mutate(fam, nw, height or 0, adjacency.shape[1], length, symmetries, _ptr(adjacency),
                          _ptr(rng.state), ctypes.byref(r)) != 0:
        raise ValueError("ODPX_Mutate: invalid parameters")
    return r


def restore_general(r, adjacency):
    load().ODP_Restore_adjacency_general(r, _ptr(adjacency))


def restore_grid(r, adjacency):
    load().ODP_Restore_adjacency_grid(r, _ptr(adjacency))


def anneal(adjacency, ncalcs, max_temp=100.0, min_temp=0.217147, hill_climbing=False, aspl_priority=False, seed=1):
    """ODPX_Anneal on the active general Init: returns (best adjacency, result dict)."""
    a = np.array(adjacency, dtype=np.int32, order="C", copy=True)
    r = AnnealResult()
    if load().ODPX_Anneal(_ptr(a), int(ncalcs), float(max_temp), float(min_temp), int(bool(hill_climbing)),
                          int(bool(aspl_priority)), int(seed), ctypes.byref(r)) != 0:
        raise ValueError("ODPX_Anneal: invalid arguments or no active Init")
    return a, r.as_dict()


def lbounds_general(nodes, degree):
    d, a = ctypes.c_int(), ctypes.c_double()
    load().ODP_Set_lbounds_general(nodes, degree, ctypes.byref(d), ctypes.byref(a))
    return d.value, a.value


def lbounds_grid(width, height, degree, length):
    d, a = ctypes.c_int(), ctypes.c_double()
    load().ODP_Set_lbounds_grid(width, height, degree, length, ctypes.byref(d), ctypes.byref(a))
    return d.value, a.value


def edges_to_adjacency_general(nodes, degree, edges, symmetries=1):
    edges = _i32(edges).reshape(-1, 2)
    adj = np.empty((nodes // symmetries, degree), np.int32)
    load().ODP_Conv_edge2adjacency_general_s(nodes, len(edges), degree, _ptr(edges), symmetries, _ptr(adj))
    return adj


def edges_to_adjacency_grid(width, height, degree, edges, symmetries=1):
    edges = _i32(edges).reshape(-1, 2)
    adj = np.empty((width * height // symmetries, degree), np.int32)
    load().ODP_Conv_edge2adjacency_grid_s(width, height, len(edges), degree, _ptr(edges), symmetries, _ptr(adj))
    return adj


def read_edges_general(path):
    lib = load()
    lines = lib.ODP_Get_lines(path.encode())
    edge = np.zeros((lines, 2), np.int32)
    lib.ODP_Read_edge_general(path.encode(), _ptr(edge))
    nodes = lib.ODP_Get_nodes(lines, _ptr(edge))
    degree = lib.ODP_Get_degree(nodes, lines, _ptr(edge))
    return edge, nodes, degree


def read_edges_grid(path):
    lib = load()
    lines = lib.ODP_Get_lines(path.encode())
    edge = np.zeros((lines, 2), np.int32)
    w, h = ctypes.c_int(), ctypes.c_int()
    lib.ODP_Read_edge_grid(path.encode(), ctypes.byref(w), ctypes.byref(h), _ptr(edge))
    nodes = lib.ODP_Get_nodes(lines, _ptr(edge))
    degree = lib.ODP_Get_degree(nodes, lines, _ptr(edge))
    length = lib.ODP_Get_length(lines, h.value, _ptr(edge))
    return edge, w.value, h.value, degree, length


# ---- serving ------------------------------------------------------------------

class Server:
    """ODPX_Serve_*: continuous batching of 2-opt evaluation requests over
    chains whose graphs stay resident in HBM."""

    def __init__(self, small_slots=2, large_slots=4, hbm_budget_gb=64.0):
        if load().ODPX_Serve_start(small_slots, large_slots, hbm_budget_gb) != 0:
            raise RuntimeError("a server is already running")

    def add_chain(self, adjacency):
        a = _i32(adjacency)
        cid = load().ODPX_Serve_add_chain(a.shape[0], a.shape[1], _ptr(a))
        if cid < 0:
            raise MemoryError("chain refused (HBM budget or invalid graph)")
        return cid

    def submit(self, chain, swaps=()):
        sw = _i32(np.array(swaps, dtype=np.int32).reshape(-1, 4)) if len(swaps) else np.zeros((0, 4), np.int32)
        rid = load().ODPX_Serve_submit(chain, len(sw), _ptr(sw))
        if rid < 0:
            raise ValueError("request refused")
        return rid

    def poll(self, max_results=4096):
        buf = (Result * max_results)()
        k = load().ODPX_Serve_poll(buf, max_results)
        return [buf[i].as_dict() for i in range(k)]

    def wait(self, rid):
        r = Result()
        if load().ODPX_Serve_wait(rid, ctypes.byref(r)) != 0:
            raise KeyError(rid)
        return r.as_dict()

    def drain(self):
        load().ODPX_Serve_drain()

    def stop(self):
        load().ODPX_Serve_stop()
