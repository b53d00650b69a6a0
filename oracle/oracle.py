"""The oracle -- TEST INFRASTRUCTURE ONLY.

Wraps (1) liboracle.so, the C restatement of the reference's ASPL method
(odp_oracle.c), and (2) oracle/_ref/odp_ref_run[_threads], the reference's own
serial / OpenMP library (aspl.c + utils.c built from the reference sources by
oracle/Makefile) behind ref_driver.c.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg use this module, and only as the checker.
"""
import ctypes
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_SERIAL = os.path.join(HERE, "_ref", "odp_ref_run")
REF_THREADS = os.path.join(HERE, "_ref", "odp_ref_run_threads")

MATRIX, MATRIX_SAVING, BFS = 1, 2, 3
GENERAL, GRID = 0, 1

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle oracle)")
        _lib = ctypes.CDLL(LIB)
        p = ctypes.c_void_p
        _lib.oracle_eval.restype = ctypes.c_int
        _lib.oracle_eval.argtypes = [ctypes.c_int] * 4 + [p, ctypes.c_int, p, ctypes.c_int, ctypes.c_int,
                                                           ctypes.POINTER(ctypes.c_int),
                                                           ctypes.POINTER(ctypes.c_long),
                                                           ctypes.POINTER(ctypes.c_double)]
        _lib.oracle_expand.restype = None
        _lib.oracle_expand.argtypes = [ctypes.c_int] * 4 + [p, ctypes.c_int, p, p]
        _lib.oracle_eval_shard.restype = ctypes.c_int
        _lib.oracle_eval_shard.argtypes = [ctypes.c_int] * 4 + [p, ctypes.c_int, p, ctypes.c_int, ctypes.c_int,
                                                                 ctypes.c_int, ctypes.POINTER(ctypes.c_long),
                                                                 ctypes.POINTER(ctypes.c_int)]
        _lib.oracle_set_fault.restype = None
        _lib.oracle_set_fault.argtypes = [ctypes.c_int]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def evaluate(adjacency, nodes=None, width=None, height=None, symmetries=1, num_degrees=None,
             kind=MATRIX_SAVING, chunk=24):
    """One Init/Set/Finalize cycle on the restatement -> (diameter, sum, ASPL).

    General graphs pass nodes=; grid graphs pass width= and height=.
    adjacency: int[nodes/symmetries][degree]; num_degrees: None or the array the
    matching ODP_Init_aspl_* entry point takes.
    """
    adj = np.ascontiguousarray(adjacency, dtype=np.int32)
    nd = None if num_degrees is None else np.ascontiguousarray(num_degrees, dtype=np.int32)
    fam = GRID if width is not None else GENERAL
    nw = width if fam == GRID else nodes
    h = height if fam == GRID else -1
    d, s, x = ctypes.c_int(), ctypes.c_long(), ctypes.c_double()
    rc = lib().oracle_eval(fam, nw, h, adj.shape[1], _p(nd), symmetries, _p(adj), kind, chunk,
                           ctypes.byref(d), ctypes.byref(s), ctypes.byref(x))
    if rc != 0:
        raise ValueError("invalid parameters")
    return d.value, s.value, x.value


def evaluate_shard(adjacency, rank, procs, nodes=None, width=None, height=None, symmetries=1, num_degrees=None,
                   kind=MATRIX):
    """One rank's share of the MPI+CUDA flavour (aspl_mpi_cuda.cu:44-123) -> (excess, diameter):
    the rank-local sum over the n(n-1) base and the rank-local diameter, before the combine."""
    adj = np.ascontiguousarray(adjacency, dtype=np.int32)
    nd = None if num_degrees is None else np.ascontiguousarray(num_degrees, dtype=np.int32)
    fam = GRID if width is not None else GENERAL
    nw = width if fam == GRID else nodes
    h = height if fam == GRID else -1
    s, d = ctypes.c_long(), ctypes.c_int()
    rc = lib().oracle_eval_shard(fam, nw, h, adj.shape[1], _p(nd), symmetries, _p(adj), kind, rank, procs,
                                 ctypes.byref(s), ctypes.byref(d))
    if rc != 0:
        raise ValueError("invalid parameters")
    return s.value, d.value


def expand(adjacency, nodes=None, width=None, height=None, symmetries=1, num_degrees=None):
    """Symmetric base table -> full int[n][degree] table with global ids (-1 = absent)."""
    adj = np.ascontiguousarray(adjacency, dtype=np.int32)
    nd = None if num_degrees is None else np.ascontiguousarray(num_degrees, dtype=np.int32)
    fam = GRID if width is not None else GENERAL
    n = width * height if fam == GRID else nodes
    out = np.empty((n, adj.shape[1]), np.int32)
    lib().oracle_expand(fam, width if fam == GRID else nodes, height if fam == GRID else -1,
                        adj.shape[1], _p(nd), symmetries, _p(adj), _p(out))
    return out


def set_fault(code):
    lib().oracle_set_fault(code)


def ref_available(threads=False):
    return os.path.exists(REF_THREADS if threads else REF_SERIAL)


def write_graph(path, adjacency, nodes=None, width=None, height=None, symmetries=1, num_degrees=None, repeat=0):
    adj = np.ascontiguousarray(adjacency, dtype=np.int32)
    fam = GRID if width is not None else GENERAL
    nw = width if fam == GRID else nodes
    h = height if fam == GRID else 0
    with open(path, "wb") as f:
        f.write(struct.pack("<8i", 0x3150444F, fam, nw, h, adj.shape[1], symmetries,
                            0 if num_degrees is None else 1, repeat))
        if num_degrees is not None:
            f.write(np.ascontiguousarray(num_degrees, dtype=np.int32).tobytes())
        f.write(adj.tobytes())


def run_reference(adjacency, nodes=None, width=None, height=None, symmetries=1, num_degrees=None,
                  kind=None, threads=False, repeat=0, omp_threads=None, timeout=600):
    """Runs the reference library on one graph -> ((diameter, sum, ASPL), [seconds per timed call])."""
    exe = REF_THREADS if threads else REF_SERIAL
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built (make -C oracle ref)")
    env = dict(os.environ)
    env.pop("ODP_ASPL", None)
    if kind is not None:
        env["ODP_ASPL"] = {MATRIX: "MATRIX", MATRIX_SAVING: "MATRIX_SAVING", BFS: "BFS"}[kind]
    if omp_threads is not None:
        env["OMP_NUM_THREADS"] = str(omp_threads)
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as tf:
        path = tf.name
    try:
        write_graph(path, adjacency, nodes, width, height, symmetries, num_degrees, repeat)
        out = subprocess.run([exe, path], env=env, capture_output=True, text=True, timeout=timeout)
    finally:
        os.unlink(path)
    if out.returncode != 0:
        raise RuntimeError(f"reference run failed ({out.returncode}): {out.stderr.strip()}")
    lines = out.stdout.strip().splitlines()
    d, s, a = lines[0].split()
    times = [float(ln.split()[1]) for ln in lines[1:] if ln.startswith("t ")]
    return (int(d), int(s), float.fromhex(a)), times
