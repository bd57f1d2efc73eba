"""CPU oracle for the sort-based neighbour-list build -- TEST INFRASTRUCTURE.

A C restatement (``nl_oracle.c``) of NeighbourLists.jl v0.6.2:

* sort-based path: ``src/cell_list.jl:152-170, 632-748`` and
  ``src/gpu_kernels.jl:20-101, 132-180, 211-219, 299-364``;
* legacy linked-list PairList: ``src/cell_list.jl:185-505``;
* an independent brute force over all (i, j, S).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker / reported
baseline.  The product library (``libnlhip.so``) never loads it.

All index outputs are 1-based, as in the reference.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

ORC_OK, ORC_EINVAL, ORC_EOVERFLOW = 0, 1, 2


class OracleError(RuntimeError):
    pass


class _Geom(C.Structure):
    _fields_ = [
        ("cell", C.c_double * 9),
        ("inv", C.c_double * 9),
        ("lens", C.c_double * 3),
        ("cutoff", C.c_double),
        ("cutoff_sq", C.c_double),
        ("ncells", C.c_int64 * 3),
        ("nxyz", C.c_int64 * 3),
        ("ncells_total", C.c_int64),
        ("pbc", C.c_int32 * 3),
        ("zero_volume", C.c_int32),
    ]


class _Result(C.Structure):
    _fields_ = [
        ("nat", C.c_int64),
        ("npairs", C.c_int64),
        ("M", C.c_int64),
        ("perm", C.POINTER(C.c_int64)),
        ("cell_id", C.POINTER(C.c_int64)),
        ("cell_offsets", C.POINTER(C.c_int64)),
        ("wind", C.POINTER(C.c_int64)),
        ("i", C.POINTER(C.c_int64)),
        ("j", C.POINTER(C.c_int64)),
        ("S", C.POINTER(C.c_int64)),
        ("first", C.POINTER(C.c_int64)),
        ("R", C.POINTER(C.c_double)),
        ("g", _Geom),
    ]


_lib = None


def build(quiet: bool = True) -> str:
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    out = subprocess.run(["make", "-C", _HERE, "liboracle.so"], capture_output=True, text=True)
    if out.returncode != 0:
        raise OracleError("oracle build failed:\n" + out.stdout + out.stderr)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp, ip, i32p = C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int32)
        L.orc_analyze.argtypes = [dp, C.c_double, i32p, C.c_int, C.POINTER(_Geom)]
        L.orc_analyze.restype = C.c_int
        L.orc_inv3.argtypes = [dp, dp]
        L.orc_det3.argtypes = [dp]
        L.orc_det3.restype = C.c_double
        L.orc_lengths.argtypes = [dp, dp]
        L.orc_neighbour_list.argtypes = [dp, C.c_int64, dp, C.c_double, i32p, C.c_int, C.c_int,
                                         C.c_int, C.POINTER(_Result)]
        L.orc_neighbour_list.restype = C.c_int
        L.orc_legacy_pairlist.argtypes = [dp, C.c_int64, dp, C.c_double, i32p, C.c_int, C.c_int,
                                          dp, dp, C.POINTER(_Result)]
        L.orc_legacy_pairlist.restype = C.c_int
        L.orc_bruteforce.argtypes = [dp, C.c_int64, dp, C.c_double, i32p, C.POINTER(_Result)]
        L.orc_bruteforce.restype = C.c_int
        L.orc_neighbours_of.argtypes = [dp, C.c_int64, C.POINTER(_Geom), ip, ip, C.c_int64,
                                        C.c_int64, ip, ip, dp]
        L.orc_neighbours_of.restype = C.c_int64
        L.orc_cell_ids.argtypes = [dp, C.c_int64, C.POINTER(_Geom), ip, ip]
        L.orc_free.argtypes = [C.POINTER(_Result)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _cell_cm(cell) -> np.ndarray:
    """3x3 cell (rows = lattice vectors) -> column-major 9-vector."""
    c = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    return np.ascontiguousarray(c.T.reshape(9))


def _pbc(pbc) -> np.ndarray:
    p = np.asarray(pbc, dtype=bool).reshape(3)
    return p.astype(np.int32)


def _X(X) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(X, dtype=np.float64).reshape(-1, 3))


@dataclass
class Geometry:
    inv_cell: np.ndarray  # 3x3 (row-major numpy view of the same matrix)
    lens: np.ndarray
    ncells: np.ndarray
    nxyz: np.ndarray
    ncells_total: int
    cutoff_sq: float
    zero_volume: bool


def _geom_py(g: _Geom) -> Geometry:
    inv = np.array(g.inv[:], dtype=np.float64).reshape(3, 3).T.copy()
    return Geometry(inv, np.array(g.lens[:]), np.array(g.ncells[:], dtype=np.int64),
                    np.array(g.nxyz[:], dtype=np.int64), int(g.ncells_total), float(g.cutoff_sq),
                    bool(g.zero_volume))


def _check(st: int, what: str):
    if st == ORC_EOVERFLOW:
        raise OverflowError(f"{what}: ratio of simulation cell size to cutoff is too large for the index type")
    if st != ORC_OK:
        raise ValueError(f"{what}: invalid input (status {st})")


def analyze(cell, cutoff: float, pbc=(True, True, True), ti_bits: int = 32) -> Geometry:
    g = _Geom()
    _check(lib().orc_analyze(_dp(_cell_cm(cell)), float(cutoff), _pbc(pbc).ctypes.data_as(
        C.POINTER(C.c_int32)), ti_bits, C.byref(g)), "analyze_cell")
    return _geom_py(g)


def inv3(cell) -> np.ndarray:
    out = np.zeros(9)
    lib().orc_inv3(_dp(_cell_cm(cell)), _dp(out))
    return out.reshape(3, 3).T.copy()


def lengths(cell) -> np.ndarray:
    out = np.zeros(3)
    lib().orc_lengths(_dp(_cell_cm(cell)), _dp(out))
    return out


@dataclass
class PairSet:
    """Materialised pair list (1-based), oracle side."""
    i: np.ndarray
    j: np.ndarray
    S: np.ndarray  # (P, 3)
    first: np.ndarray | None = None
    R: np.ndarray | None = None
    perm: np.ndarray | None = None
    cell_id: np.ndarray | None = None
    cell_offsets: np.ndarray | None = None
    wind: np.ndarray | None = None
    geometry: Geometry | None = None
    X: np.ndarray | None = None
    cell: np.ndarray | None = None

    @property
    def npairs(self) -> int:
        return int(self.i.shape[0])


def _arr(ptr, n, dtype=np.int64, shape=None):
    if n == 0:
        return np.zeros(shape if shape else (0,), dtype=dtype)
    a = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)
    return a.reshape(shape) if shape else a


def neighbour_list(X, cutoff, cell, pbc=(True, True, True), ti_bits: int = 32,
                   want_R: bool = True, nthreads: int = 1) -> PairSet:
    """Sort-based build_cell_list + materialize_pairlist (cell_list.jl:897-916)."""
    X = _X(X)
    nat = X.shape[0]
    r = _Result()
    st = lib().orc_neighbour_list(_dp(X), nat, _dp(_cell_cm(cell)), float(cutoff),
                                  _pbc(pbc).ctypes.data_as(C.POINTER(C.c_int32)), ti_bits,
                                  1 if want_R else 0, int(nthreads), C.byref(r))
    try:
        _check(st, "neighbour_list")
        P = r.npairs
        M = r.M
        return PairSet(
            i=_arr(r.i, P), j=_arr(r.j, P), S=_arr(r.S, 3 * P, shape=(P, 3)),
            first=_arr(r.first, nat + 1),
            R=_arr(r.R, 3 * P, np.float64, (P, 3)) if want_R else None,
            perm=_arr(r.perm, nat), cell_id=_arr(r.cell_id, nat),
            cell_offsets=_arr(r.cell_offsets, M + 1), wind=_arr(r.wind, 3 * nat, shape=(nat, 3)),
            geometry=_geom_py(r.g), X=X, cell=np.asarray(cell, dtype=np.float64).reshape(3, 3))
    finally:
        lib().orc_free(C.byref(r))


def time_neighbour_list(X, cutoff, cell, pbc, nthreads: int, ti_bits: int = 32) -> int:
    """Run the full CPU build once without copying results out; returns npairs."""
    X = _X(X)
    r = _Result()
    st = lib().orc_neighbour_list(_dp(X), X.shape[0], _dp(_cell_cm(cell)), float(cutoff),
                                  _pbc(pbc).ctypes.data_as(C.POINTER(C.c_int32)), ti_bits, 0,
                                  int(nthreads), C.byref(r))
    P = r.npairs
    lib().orc_free(C.byref(r))
    _check(st, "neighbour_list")
    return int(P)


def legacy_pairlist(X, cutoff, cell, pbc=(True, True, True), ti_bits: int = 32,
                    fixcell: bool = True) -> PairSet:
    """Legacy linked-list PairList (cell_list.jl:369-384)."""
    X = _X(X)
    nat = X.shape[0]
    Xf = np.zeros_like(X)
    Cf = np.zeros(9)
    r = _Result()
    st = lib().orc_legacy_pairlist(_dp(X), nat, _dp(_cell_cm(cell)), float(cutoff),
                                   _pbc(pbc).ctypes.data_as(C.POINTER(C.c_int32)), ti_bits,
                                   1 if fixcell else 0, _dp(Xf), _dp(Cf), C.byref(r))
    try:
        _check(st, "legacy PairList")
        P = r.npairs
        return PairSet(i=_arr(r.i, P), j=_arr(r.j, P), S=_arr(r.S, 3 * P, shape=(P, 3)),
                       first=_arr(r.first, nat + 1), X=Xf, cell=Cf.reshape(3, 3).T.copy())
    finally:
        lib().orc_free(C.byref(r))


def bruteforce(X, cutoff, cell, pbc=(True, True, True)) -> PairSet:
    X = _X(X)
    r = _Result()
    st = lib().orc_bruteforce(_dp(X), X.shape[0], _dp(_cell_cm(cell)), float(cutoff),
                              _pbc(pbc).ctypes.data_as(C.POINTER(C.c_int32)), C.byref(r))
    try:
        _check(st, "bruteforce")
        P = r.npairs
        return PairSet(i=_arr(r.i, P), j=_arr(r.j, P), S=_arr(r.S, 3 * P, shape=(P, 3)))
    finally:
        lib().orc_free(C.byref(r))


def cell_ids(X, cutoff, cell, pbc=(True, True, True), ti_bits: int = 32):
    """(_compute_cell_ids, windings) for every atom; ids 1-based."""
    X = _X(X)
    g = _Geom()
    _check(lib().orc_analyze(_dp(_cell_cm(cell)), float(cutoff), _pbc(pbc).ctypes.data_as(
        C.POINTER(C.c_int32)), ti_bits, C.byref(g)), "analyze_cell")
    n = X.shape[0]
    ids = np.zeros(max(n, 1), dtype=np.int64)
    w = np.zeros((max(n, 1), 3), dtype=np.int64)
    lib().orc_cell_ids(_dp(X), n, C.byref(g), _ip(ids), _ip(w))
    return ids[:n], w[:n]


def canonical(i, j, S, R=None):
    """Sort pairs canonically by (i, j, S) -- the reference's get_sorted_pairs
    (test/test_utils.jl:68-70)."""
    i = np.asarray(i, dtype=np.int64)
    j = np.asarray(j, dtype=np.int64)
    S = np.asarray(S, dtype=np.int64).reshape(-1, 3)
    order = np.lexsort((S[:, 2], S[:, 1], S[:, 0], j, i))
    out = (i[order], j[order], S[order])
    if R is not None:
        out = out + (np.asarray(R).reshape(-1, 3)[order],)
    return out
