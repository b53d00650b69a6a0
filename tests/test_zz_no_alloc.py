"""The allocation-free entry point ODPX_Set_aspl_noalloc against ODP_Set_aspl,
through the C ABI: the same graph gives the same integers and the same ASPL
double; two calls in a row, the second on a smaller graph (the shared device
array keeps the first call's contents); sizes past the stated maximum are
refused without touching the device."""
import ctypes

import numpy as np
import pytest

import odp_amd as odp
import oracle

pytestmark = pytest.mark.gpu


def noalloc(adj):
    f = odp.load().ODPX_Set_aspl_noalloc
    a = np.ascontiguousarray(adj, dtype=np.int32)
    d, s, x = ctypes.c_int(), ctypes.c_long(), ctypes.c_double()
    rc = f(a.shape[0], a.shape[1], a.ctypes.data_as(ctypes.c_void_p), ctypes.byref(d), ctypes.byref(s),
           ctypes.byref(x))
    return rc, (d.value, s.value, x.value)


def drop_in(adj):
    n, deg = adj.shape
    odp.init_aspl_general(n, deg)
    try:
        return odp.set_aspl(adj)
    finally:
        odp.finalize_aspl()


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()


def test_noalloc_equals_set_aspl_twice_smaller_second():
    big = odp.generate_regular(16384, 24, 11)          # the largest size and a wide table
    small = odp.generate_regular(1000, 5, 12)
    for adj in (big, small, big, odp.generate_regular(3000, 8, 13)):
        rc, got = noalloc(adj)
        assert rc == 0
        assert same(got, drop_in(adj)), adj.shape
    assert same(noalloc(small)[1], oracle.evaluate(small, nodes=1000))


def test_noalloc_disconnected_and_absent_edges():
    two = np.concatenate([odp.generate_regular(300, 4, 3), odp.generate_regular(200, 4, 4) + 300]).astype(np.int32)
    rc, got = noalloc(two)
    assert rc == 0 and got == (2**31 - 1, 2**63 - 1, float.fromhex("0x1.fffffffffffffp+1023"))
    path = np.full((50, 2), -1, np.int32)                # a path with absent entries (-1)
    for i in range(50):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < 50]
        path[i, :len(nb)] = nb
    rc, got = noalloc(path)
    assert rc == 0 and got[:2] == (49, 49 * 50 * 51 // 6)


def test_noalloc_refuses_sizes_past_the_maximum():
    assert noalloc(np.zeros((16386, 2), np.int32))[0] == -1
    assert noalloc(np.zeros((100, 33), np.int32))[0] == -1
    assert noalloc(np.zeros((0, 4), np.int32))[0] == -1
