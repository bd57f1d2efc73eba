"""Host-side mirror of NeighbourLists.jl's sort-based API over libnlhip.

Names, argument meaning, 1-based indices and error behaviour follow the
reference (v0.6.2):

=========================================  ===================================
this module                                reference
=========================================  ===================================
``neighbour_list(X, cutoff, cell, pbc)``   ``neighbour_list`` cell_list.jl:897-916
``build_cell_list``                        cell_list.jl:632-642
``materialize_pairlist``                   gpu_kernels.jl:299-364
``PairList(X, cutoff, cell, pbc)``         cell_list.jl:848-855
``PairList`` / ``SortedCellList`` fields   types.jl:34-43 / 70-82
``npairs``, ``nsites``, ``cutoff``         cell_list.jl:25-27, 760-767
``nneigs``, ``maxneigs``/``max_neighbours`` cell_list.jl:513-523, 611
``neigs``, ``neigss``, ``neighbours``      cell_list.jl:549-606, 821-833
``for_each_neighbour``, ``count_neighbours`` cell_list.jl:779-814
``num_neighbours``                         cell_list.jl:926-927
``pairs``, ``sites``                       iterators.jl:9-40
=========================================  ===================================

Positions ``X`` are an (nat, 3) float64 array: a numpy array (host in,
host out) or a torch CUDA tensor (device in, device out -- the analogue of
passing a ROCArray).  Every build runs on the GPU through libnlhip; there is
no CPU fallback.  ``cell`` is 3x3 with ROWS = lattice vectors.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable

import numpy as np

from . import _lib
from ._lib import NLH_WANT_R, NLH_X_DEVICE, Geometry, Handle, check, geometry, lib, ti_bytes_of, ti_dtype

Int32, Int64 = np.int32, np.int64


def _is_torch_cuda(X) -> bool:
    try:
        import torch
    except ImportError:
        return False
    return isinstance(X, torch.Tensor) and X.is_cuda


def _prep_positions(X):
    """-> (pointer, nat, flags, keepalive, device_tensor_or_None)"""
    if _is_torch_cuda(X):
        import torch

        if X.dtype != torch.float64 or X.dim() != 2 or X.shape[1] != 3:
            raise ValueError("X must be an (nat, 3) float64 tensor")
        Xc = X.contiguous()
        torch.cuda.current_stream(Xc.device).synchronize()
        return Xc.data_ptr(), Xc.shape[0], NLH_X_DEVICE, Xc, Xc
    Xn = np.ascontiguousarray(np.asarray(X, dtype=np.float64))
    if Xn.ndim == 1 and Xn.size == 0:
        Xn = Xn.reshape(0, 3)
    if Xn.ndim != 2 or Xn.shape[1] != 3:
        raise ValueError("X must be an (nat, 3) array of float64 positions")
    return Xn.ctypes.data, Xn.shape[0], 0, Xn, None


def _empty(shape, dtype, like_device):
    if like_device is not None:
        import torch

        tdt = {np.int32: torch.int32, np.int64: torch.int64, np.float64: torch.float64}[dtype]
        return torch.empty(shape, dtype=tdt, device=like_device.device)
    return np.empty(shape, dtype=dtype)


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data if a.size else None
    return a.data_ptr() if a.numel() else None


def _host(a):
    """numpy view of an output (copies device tensors to the host)."""
    if a is None or isinstance(a, np.ndarray):
        return a
    return a.cpu().numpy()


class SortedCellList:
    """Sort-based cell list (src/types.jl:70-82).

    Fields mirror the reference: ``X`` (positions sorted by cell),
    ``X_orig``, ``perm`` (sorted index -> original index, 1-based),
    ``cell_id`` (1-based, sorted order), ``cell_offsets`` (1-based CSR,
    ncells_total + 1 entries), ``cell``, ``inv_cell``, ``pbc``, ``cutoff``,
    ``ncells``, ``ncells_total``.  The device state stays resident in the
    owning handle for lazy queries and materialisation.
    """

    def __init__(self, handle: Handle, geom: Geometry, X_orig, device_like):
        self._handle = handle
        self._geom = geom
        self._device_like = device_like
        self.X_orig = X_orig
        self.cell = geom.cell_matrix
        self.inv_cell = geom.inv_matrix
        self.pbc = np.array([bool(b) for b in geom.pbc], dtype=bool)
        self.cutoff = float(geom.cutoff)
        self.ncells = np.array(geom.ncells[:], dtype=ti_dtype(geom.ti_bytes))
        self.ncells_total = int(geom.ncells_total)
        self.int_type = ti_dtype(geom.ti_bytes)
        self._fields = None

    def _load(self):
        if self._fields is None:
            nat, M, dt = self.nsites(), self.ncells_total, self.int_type
            Xs = _empty((nat, 3), np.float64, self._device_like)
            perm = _empty((nat,), dt, self._device_like)
            cid = _empty((nat,), dt, self._device_like)
            coff = _empty((M + 1,), dt, self._device_like)
            check(lib().nlh_copy_celllist(self._handle.ptr, _ptr(Xs), _ptr(perm), _ptr(cid), _ptr(coff)),
                  "copy cell list")
            self._fields = (Xs, perm, cid, coff)
        return self._fields

    @property
    def X(self):
        return self._load()[0]

    @property
    def perm(self):
        return self._load()[1]

    @property
    def cell_id(self):
        return self._load()[2]

    @property
    def cell_offsets(self):
        return self._load()[3]

    def nsites(self) -> int:
        return int(self.X_orig.shape[0])

    def device_view(self):
        """Device pointers (nlh_celllist_view) for user kernels."""
        v = _lib.CellListView()
        check(lib().nlh_celllist_device(self._handle.ptr, C.byref(v)))
        return v

    @property
    def geometry(self) -> Geometry:
        return self._geom


class PairList:
    """Materialised neighbour list (src/types.jl:34-43).

    Fields: ``X`` (the input positions, aliased), ``C`` (cell), ``cutoff``,
    ``i``, ``j`` (1-based), ``S`` (nP x 3 cell shifts), ``first`` (1-based
    CSR: pairs of atom m are ``first[m-1]:first[m]-1`` in 1-based terms).
    There is no R field; R is recomputed on access as in the reference
    (``_getR``, cell_list.jl:525-531).  ``R`` is filled only when the build
    was asked for it (``want_R=True``).

    ``PairList(X, cutoff, cell, pbc; int_type)`` builds one with the
    sort-based path, like the reference's new-API constructor
    (cell_list.jl:848-855).
    """

    def __init__(self, X, cutoff=None, cell=None, pbc=None, *, int_type=Int32, want_R: bool = False,
                 _fields=None):
        if _fields is not None:
            (self.X, self.C, self.cutoff, self.i, self.j, self.S, self.first, self.R) = _fields
            return
        if cutoff is None or cell is None or pbc is None:
            raise TypeError("PairList(X, cutoff, cell, pbc; int_type)")
        nl = neighbour_list(X, cutoff, cell, pbc, int_type=int_type, want_R=want_R)
        (self.X, self.C, self.cutoff, self.i, self.j, self.S, self.first, self.R) = (
            nl.X, nl.C, nl.cutoff, nl.i, nl.j, nl.S, nl.first, nl.R)

    @classmethod
    def _make(cls, *fields):
        return cls(None, _fields=fields)

    def npairs(self) -> int:
        return int(self.i.shape[0])

    def nsites(self) -> int:
        return int(self.first.shape[0]) - 1


# ---------------------------------------------------------------------------
# builds
# ---------------------------------------------------------------------------

def build_cell_list(X, cutoff, cell, pbc, *, int_type=Int32, handle: Handle | None = None) -> SortedCellList:
    """Sort-based cell list on the GPU (cell_list.jl:632-679)."""
    ptr, nat, flags, keep, dev = _prep_positions(X)
    g = geometry(cell, cutoff, pbc, ti_bytes_of(int_type))
    h = handle or Handle()
    check(lib().nlh_build_cell_list(h.ptr, ptr, nat, C.byref(g), flags), "build_cell_list")
    cl = SortedCellList(h, g, keep if dev is None else dev, dev)
    cl._keepalive = keep
    return cl


def materialize_pairlist(clist: SortedCellList, *, want_R: bool = False) -> PairList:
    """Two-pass pair materialisation (gpu_kernels.jl:299-364)."""
    n = C.c_int64()
    flags = NLH_WANT_R if want_R else 0
    check(lib().nlh_materialize(clist._handle.ptr, flags, C.byref(n)), "materialize_pairlist")
    return _collect_pairs(clist._handle, clist._geom, clist.X_orig, n.value, want_R, clist._device_like)


def _collect_pairs(h: Handle, g: Geometry, X, P: int, want_R: bool, dev) -> PairList:
    dt = ti_dtype(g.ti_bytes)
    nat = int(X.shape[0])
    i = _empty((P,), dt, dev)
    j = _empty((P,), dt, dev)
    S = _empty((P, 3), dt, dev)
    first = _empty((nat + 1,), dt, dev)
    R = _empty((P, 3), np.float64, dev) if want_R else None
    check(lib().nlh_copy_pairlist(h.ptr, _ptr(i), _ptr(j), _ptr(S), _ptr(first), _ptr(R)), "copy pair list")
    return PairList._make(X, g.cell_matrix, float(g.cutoff), i, j, S, first, R)


def neighbour_list(X, cutoff, cell, pbc, *, lazy: bool = False, int_type=Int32, want_R: bool = False,
                   handle: Handle | None = None, no_graph: bool = False):
    """Unified entry point (cell_list.jl:897-916): a PairList, or with
    ``lazy=True`` a SortedCellList.  Repeated calls of one shape on a handle
    replay a captured HIP graph unless ``no_graph``."""
    if lazy:
        return build_cell_list(X, cutoff, cell, pbc, int_type=int_type, handle=handle)
    ptr, nat, flags, keep, dev = _prep_positions(X)
    g = geometry(cell, cutoff, pbc, ti_bytes_of(int_type))
    h = handle or Handle()
    n = C.c_int64()
    if want_R:
        flags |= NLH_WANT_R
    if no_graph:
        flags |= _lib.NLH_NO_GRAPH
    check(lib().nlh_neighbour_list(h.ptr, ptr, nat, C.byref(g), flags, C.byref(n)), "neighbour_list")
    return _collect_pairs(h, g, keep if dev is None else dev, n.value, want_R, dev)


# ---------------------------------------------------------------------------
# accessors (1-based atom indices, as in the reference)
# ---------------------------------------------------------------------------

def npairs(nlist: PairList) -> int:
    return nlist.npairs()


def nsites(obj) -> int:
    return obj.nsites()


def cutoff(obj) -> float:
    return float(obj.cutoff)


def nneigs(nlist: PairList, i0: int) -> int:
    f = _host(nlist.first)
    return int(f[i0] - f[i0 - 1])


def maxneigs(nlist: PairList) -> int:
    f = _host(nlist.first).astype(np.int64)
    if f.shape[0] <= 1:
        # the reference's maximum over an empty range throws; an empty
        # system has no neighbours
        return 0
    return int(np.max(np.diff(f)))


max_neighbours = maxneigs
max_neigs = maxneigs


def _getR(X, j, i, S, C_):
    """R = (X[j] - X[i]) + C' * S in the reference's evaluation order
    (cell_list.jl:525-531): (C'S)_k = (C[0,k] S0 + C[1,k] S1) + C[2,k] S2."""
    X = _host(X)
    Sd = np.asarray(S, dtype=np.float64).reshape(-1, 3)
    dX = X[np.asarray(j) - 1] - X[np.asarray(i) - 1]
    cs = (C_[0][None, :] * Sd[:, 0:1] + C_[1][None, :] * Sd[:, 1:2]) + C_[2][None, :] * Sd[:, 2:3]
    return dX + cs


def neigss(nlist: PairList, i0: int):
    """(j, R, S) of atom i0 (cell_list.jl:583-598)."""
    f = _host(nlist.first)
    a, b = int(f[i0 - 1]) - 1, int(f[i0]) - 1
    J = _host(nlist.j)[a:b]
    S = _host(nlist.S)[a:b]
    R = _getR(nlist.X, J, np.full(J.shape, i0), S, np.asarray(nlist.C))
    return J, R, S


def neigs(nlist: PairList, i0: int):
    """(j, R) of atom i0 (cell_list.jl:549-557, 594-595)."""
    J, R, _ = neigss(nlist, i0)
    return J, R


def neighbours(obj, i: int):
    """PairList: alias of neigss (cell_list.jl:606).  SortedCellList: lazy
    query on the device (cell_list.jl:821-833)."""
    if isinstance(obj, PairList):
        return neigss(obj, i)
    return _lazy_neighbours(obj, i)


def neighbours_range(clist: SortedCellList, i_first: int, n: int):
    """Neighbours of atoms i_first .. i_first + n - 1 (1-based) in one pass
    on the device (nlh_neighbours_range): (offsets, j, R, S), where atom
    i_first + a owns entries offsets[a] .. offsets[a + 1] - 1 of j, R, S --
    exactly what neighbours(clist, i) returns for each of them, in the
    reference order (cell_list.jl:779-833)."""
    h = clist._handle
    dt = clist.int_type
    n = int(n)
    offs = np.zeros(n + 1, dtype=np.int64)
    op = offs.ctypes.data_as(C.POINTER(C.c_int64))
    check(lib().nlh_neighbours_range(h.ptr, 0, int(i_first), n, 0, op, None, None, None), "neighbours")
    tot = int(offs[-1])
    js = np.empty(tot, dtype=dt)
    Rs = np.empty((tot, 3), dtype=np.float64)
    Ss = np.empty((tot, 3), dtype=dt)
    if tot:   # the results are still on the device: this call only copies them
        check(lib().nlh_neighbours_range(h.ptr, 0, int(i_first), n, tot, op, js.ctypes.data, Rs.ctypes.data,
                                         Ss.ctypes.data), "neighbours")
    return offs, js, Rs, Ss


_LAZY_BLOCK = 1 << 16    # atoms per cached range query
_LAZY_BLOCKS_KEPT = 4


def _lazy_block(clist: SortedCellList, i: int):
    """(first atom, offsets, j, R, S) of the cached block of atoms holding
    atom i: lazy queries run over blocks of 65,536 atoms at a time, so a
    sweep over the atoms costs a few device passes, not one per atom."""
    blocks = clist.__dict__.setdefault("_lazy_blocks", {})
    b = (int(i) - 1) // _LAZY_BLOCK
    hit = blocks.get(b)
    if hit is None:
        a0 = b * _LAZY_BLOCK + 1
        n = min(_LAZY_BLOCK, clist.nsites() - a0 + 1)
        hit = (a0,) + neighbours_range(clist, a0, n)
        if len(blocks) >= _LAZY_BLOCKS_KEPT:
            blocks.pop(next(iter(blocks)))
        blocks[b] = hit
    return hit


def _lazy_neighbours(clist: SortedCellList, i: int):
    i = int(i)
    if i < 1 or i > clist.nsites():
        raise _lib.NLHInvalidError(_lib.NLH_EINVAL, "atom index out of range (1-based)")
    a0, offs, js, Rs, Ss = _lazy_block(clist, i)
    lo, hi = int(offs[i - a0]), int(offs[i - a0 + 1])
    return js[lo:hi], Rs[lo:hi], Ss[lo:hi]


def for_each_neighbour(f: Callable, clist: SortedCellList, i: int) -> None:
    """Call f(j, R, S) for every neighbour j of atom i (cell_list.jl:779-801);
    the neighbours are found on the device, in the reference order."""
    js, Rs, Ss = _lazy_neighbours(clist, i)
    for q in range(js.shape[0]):
        f(js[q], Rs[q], Ss[q])


def count_neighbours(clist: SortedCellList, i: int) -> int:
    """cell_list.jl:809-814"""
    return int(_lazy_neighbours(clist, i)[0].shape[0])


def num_neighbours(obj, i: int) -> int:
    """cell_list.jl:926-927"""
    if isinstance(obj, PairList):
        return nneigs(obj, i)
    return count_neighbours(obj, i)


# ---------------------------------------------------------------------------
# iterators (iterators.jl)
# ---------------------------------------------------------------------------

def pairs(nlist: PairList):
    """Iterate (i, j, R) over all pairs (iterators.jl:9-23)."""
    I, J, S = _host(nlist.i), _host(nlist.j), _host(nlist.S)
    R = _getR(nlist.X, J, I, S, np.asarray(nlist.C))
    for q in range(I.shape[0]):
        yield I[q], J[q], R[q]


def sites(nlist: PairList):
    """Iterate (i, j, R) per site (iterators.jl:27-40)."""
    for i in range(1, nlist.nsites() + 1):
        J, R = neigs(nlist, i)
        yield i, J, R
