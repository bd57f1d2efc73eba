"""User HIP kernels over the device-side header include/nlhip_device.h.

The reference's "SortedCellList in user kernels" test (test/test_gpu.jl:94-122)
counts neighbours per atom in a custom kernel that calls for_each_neighbour
and compares the counts with a materialised pair list.  Here the user kernel
(tests/user_kernel/user_kernel.hip, built with the package) also sums R and S
per atom in traversal order, and the sums are compared bit for bit with the
same sums over the oracle's pair list.
"""
import ctypes as C
import os

import numpy as np
import pytest

import nlhip
import oracle
from nlhip import _lib, workloads

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "user_kernel", "libnlh_user_kernel.so")


def _user():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is missing: run the package build first")
    _lib.lib()   # the HIP runtime libnlhip.so uses, loaded first
    L = C.CDLL(LIB)
    L.user_count_neighbours.argtypes = [C.POINTER(_lib.CellListView), C.POINTER(_lib.Geometry), C.c_void_p,
                                        C.c_void_p, C.c_void_p]
    L.user_count_neighbours.restype = C.c_int
    return L


def _run(X, rc, cell, pbc, int_type=np.int32):
    cl = nlhip.build_cell_list(X, rc, cell, pbc, int_type=int_type)
    v = cl.device_view()
    g = cl.geometry
    n = X.shape[0]
    cnt = np.zeros(n, dtype=np.int32)
    Rs = np.zeros((n, 3))
    Ss = np.zeros((n, 3), dtype=np.int64)
    _lib.check(_user().user_count_neighbours(C.byref(v), C.byref(g), cnt.ctypes.data, Rs.ctypes.data,
                                             Ss.ctypes.data), "user kernel")
    return cnt, Rs, Ss


def _expected(ref, n):
    cnt = np.diff(ref.first).astype(np.int32)
    Rs = np.zeros((n, 3))
    Ss = np.zeros((n, 3), dtype=np.int64)
    for a in range(n):
        lo, hi = ref.first[a] - 1, ref.first[a + 1] - 1
        for k in range(lo, hi):                 # traversal order, one add at a time
            Rs[a] += ref.R[k]
            Ss[a] += ref.S[k]
    return cnt, Rs, Ss


@pytest.mark.parametrize("pbc", [(True, True, True), (True, False, True), (False, False, False)])
def test_user_kernel_counts_and_sums(pbc):
    s = workloads.random_cubic(50, seed=71)          # rand_config(50), cutoff L/3 as in the reference test
    L = s.cell[0, 0]
    print(f"user kernel pbc={pbc}", flush=True)
    cnt, Rs, Ss = _run(s.X, L / 3, s.cell, pbc)
    ref = oracle.neighbour_list(s.X, L / 3, s.cell, pbc, want_R=True)
    ec, eR, eS = _expected(ref, 50)
    assert np.array_equal(cnt, ec)
    assert np.array_equal(Ss, eS)
    assert np.array_equal(Rs.view(np.int64), eR.view(np.int64))


def test_user_kernel_triclinic_multi_image_int64():
    C3 = np.array([[10.0, 2.0, 1.0], [0.0, 9.0, 1.5], [0.0, 0.0, 8.0]])
    s = workloads.random_in_cell(60, C3, seed=72)
    cnt, Rs, Ss = _run(s.X, 6.5, C3, (True, True, True), np.int64)
    ref = oracle.neighbour_list(s.X, 6.5, C3, (True, True, True), ti_bits=64, want_R=True)
    ec, eR, eS = _expected(ref, 60)
    assert np.array_equal(cnt, ec) and np.array_equal(Ss, eS)
    assert np.array_equal(Rs.view(np.int64), eR.view(np.int64))
    # the count agrees with the materialised GPU list too (test_gpu.jl:114-121)
    nl = nlhip.neighbour_list(s.X, 6.5, C3, (True, True, True), int_type=np.int64)
    assert np.array_equal(np.bincount(nl.i - 1, minlength=60), cnt)
