"""Batched neighbour lists: many independent systems in one pipeline.

The reference builds one system per call (src/cell_list.jl:897-916); here a
batch of systems -- each with its own positions, cell, pbc and cutoff --
goes through a single launch per pipeline stage (nlh_batch_neighbour_list).
Every system's result is exactly what a single-system build returns.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (NLH_NO_GRAPH, NLH_WANT_R, NLH_X_DEVICE, Geometry, Handle, Tuning, check, geometry, lib,
                   ti_bytes_of, ti_dtype)
from .api import Int32, PairList


class BatchResult:
    """Concatenated pair arrays of a batch plus per-system slicing.

    System s owns pairs [pair_offsets[s], pair_offsets[s+1]) and atoms
    [atom_offsets[s], atom_offsets[s+1]); its j values are local and 1-based.
    """

    def __init__(self, handle, geoms, atom_offsets, pair_offsets, X, i, j, S, first, R):
        self.handle = handle
        self.geoms = geoms
        self.atom_offsets = atom_offsets
        self.pair_offsets = pair_offsets
        self.X, self.i, self.j, self.S, self.first, self.R = X, i, j, S, first, R

    @property
    def nsys(self) -> int:
        return len(self.geoms)

    @property
    def npairs(self) -> int:
        return int(self.pair_offsets[-1])

    def pairlist(self, s: int) -> PairList:
        a, b = int(self.pair_offsets[s]), int(self.pair_offsets[s + 1])
        xa, xb = int(self.atom_offsets[s]), int(self.atom_offsets[s + 1])
        g = self.geoms[s]
        first = self.first[xa + s: xb + s + 1] if self.first is not None else None
        R = self.R[a:b] if self.R is not None else None
        return PairList._make(self.X[xa:xb], g.cell_matrix, float(g.cutoff), self.i[a:b], self.j[a:b],
                              self.S[a:b], first, R)


def _device_ptr(X_device) -> int:
    """Device pointer of a DeviceArray or a torch CUDA tensor."""
    return X_device.ptr if hasattr(X_device, "ptr") else X_device.data_ptr()


def pack_systems(systems):
    """-> (X concatenated (nat,3), atom_offsets (nsys+1,))"""
    nat = np.array([s.X.shape[0] for s in systems], dtype=np.int64)
    offs = np.zeros(len(systems) + 1, dtype=np.int64)
    np.cumsum(nat, out=offs[1:])
    X = np.ascontiguousarray(np.concatenate([np.asarray(s.X, dtype=np.float64).reshape(-1, 3) for s in systems])
                             if len(systems) else np.zeros((0, 3)))
    return X, offs


class BatchBuilder:
    """Reusable batched builder: keeps one handle (device buffers and the
    pipeline) across calls."""

    def __init__(self, int_type=Int32):
        self.handle = Handle()
        self.ti = ti_bytes_of(int_type)
        self.geoms = None
        self.atom_offsets = None

    def set_systems(self, systems):
        geoms = [geometry(s.cell, s.cutoff, s.pbc, self.ti) for s in systems]
        arr = (Geometry * len(geoms))(*geoms)
        self.geoms = geoms
        self._garr = arr
        return arr

    def run(self, systems, *, want_R: bool = False, X_device=None, copy: bool = True) -> BatchResult | int:
        """Build all systems.  With X_device (a DeviceArray or torch CUDA
        float64 tensor of the concatenated positions) the positions are read
        from HBM.  With
        copy=False only the pair count is returned (results stay on the
        device)."""
        X, offs = pack_systems(systems)
        garr = self.set_systems(systems)
        flags = NLH_WANT_R if want_R else 0
        ptr = X.ctypes.data
        if X_device is not None:
            flags |= NLH_X_DEVICE
            ptr = _device_ptr(X_device)
        n = C.c_int64()
        check(lib().nlh_batch_neighbour_list(self.handle.ptr, len(systems), ptr,
                                             offs.ctypes.data_as(C.POINTER(C.c_int64)), garr, flags, C.byref(n)),
              "batch_neighbour_list")
        self.atom_offsets = offs
        if not copy:
            return int(n.value)
        return self.collect(X, want_R)

    def rebuild(self, X, *, want_R: bool = False, X_device=None, copy: bool = True, graph: bool = True):
        """New positions for the same systems (same atom counts and cells).
        The pipeline is captured once as a HIP graph and replayed
        (graph=False runs the stages one by one)."""
        flags = NLH_WANT_R if want_R else 0
        if not graph:
            flags |= NLH_NO_GRAPH
        X = np.ascontiguousarray(np.asarray(X, dtype=np.float64))
        ptr = X.ctypes.data
        if X_device is not None:
            flags |= NLH_X_DEVICE
            ptr = _device_ptr(X_device)
        n = C.c_int64()
        check(lib().nlh_batch_rebuild(self.handle.ptr, ptr, flags, C.byref(n)), "batch_rebuild")
        if not copy:
            return int(n.value)
        return self.collect(X, want_R)

    def autotune(self, X, *, X_device=None, reps: int = 2) -> Tuning:
        """Pick the fastest launch configuration for the current batch
        (nlh_autotune); results are identical under every configuration."""
        X = np.ascontiguousarray(np.asarray(X, dtype=np.float64))
        ptr, flags = X.ctypes.data, 0
        if X_device is not None:
            flags |= NLH_X_DEVICE
            ptr = _device_ptr(X_device)
        best = Tuning()
        check(lib().nlh_autotune(self.handle.ptr, ptr, flags, int(reps), C.byref(best)), "autotune")
        return best

    def collect(self, X, want_R: bool) -> BatchResult:
        nsys = len(self.geoms)
        po = np.zeros(nsys + 1, dtype=np.int64)
        check(lib().nlh_batch_pair_offsets(self.handle.ptr, po.ctypes.data_as(C.POINTER(C.c_int64))))
        P = int(po[-1])
        dt = ti_dtype(self.ti)
        nat = int(self.atom_offsets[-1])
        i = np.empty(P, dtype=dt)
        j = np.empty(P, dtype=dt)
        S = np.empty((P, 3), dtype=dt)
        first = np.empty(nat + nsys, dtype=dt)
        R = np.empty((P, 3), dtype=np.float64) if want_R else None
        check(lib().nlh_copy_pairlist(self.handle.ptr, i.ctypes.data if P else None,
                                      j.ctypes.data if P else None, S.ctypes.data if P else None,
                                      first.ctypes.data, R.ctypes.data if (want_R and P) else None),
              "copy batch")
        return BatchResult(self.handle, self.geoms, self.atom_offsets, po, X, i, j, S, first, R)


def batch_neighbour_list(systems, *, int_type=Int32, want_R: bool = False) -> BatchResult:
    """One batched build of independent systems (workloads.System objects or
    anything with X, cell, pbc, cutoff attributes)."""
    return BatchBuilder(int_type).run(systems, want_R=want_R)
