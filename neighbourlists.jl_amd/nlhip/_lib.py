"""ctypes binding of libnlhip.so (include/nlhip.h).

The library is the product path: nothing here computes neighbours on the
CPU.  If the shared library is missing or no HIP device is present, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libnlhip.so")

NLH_OK, NLH_EINVAL, NLH_EOVERFLOW, NLH_ENOMEM, NLH_EDEVICE, NLH_ESTATE = range(6)
NLH_X_DEVICE, NLH_WANT_R, NLH_NO_SYNC, NLH_NO_GRAPH = 0x1, 0x2, 0x4, 0x8


class NLHError(RuntimeError):
    """Error reported by libnlhip (status code in ``code``)."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class NLHInvalidError(NLHError, ValueError):
    pass


class NLHOverflowError(NLHError, OverflowError):
    """The reference raises `error(...)` here (src/cell_list.jl:655-659)."""


class NLHMemoryError(NLHError, MemoryError):
    pass


class Geometry(C.Structure):
    """nlh_geometry (include/nlhip.h)."""

    _fields_ = [
        ("cell", C.c_double * 9),
        ("inv_cell", C.c_double * 9),
        ("lens", C.c_double * 3),
        ("cutoff", C.c_double),
        ("cutoff_sq", C.c_double),
        ("ncells", C.c_int64 * 3),
        ("nxyz", C.c_int64 * 3),
        ("ncells_total", C.c_int64),
        ("pbc", C.c_uint8 * 3),
        ("zero_volume", C.c_uint8),
        ("ti_bytes", C.c_int32),
    ]

    # numpy views (row-major 3x3 of the same matrices)
    @property
    def cell_matrix(self) -> np.ndarray:
        return np.array(self.cell[:]).reshape(3, 3).T.copy()

    @property
    def inv_matrix(self) -> np.ndarray:
        return np.array(self.inv_cell[:]).reshape(3, 3).T.copy()


class Tuning(C.Structure):
    """nlh_tuning: launch configuration (identical results for every value)."""
    _fields_ = [("brick_atoms", C.c_int32), ("fill_block", C.c_int32), ("fill_pairs", C.c_int32),
                ("count_block", C.c_int32), ("cell_sort", C.c_int32), ("fused_small", C.c_int32)]

    def __repr__(self):
        return (f"Tuning(brick_atoms={self.brick_atoms}, fill_block={self.fill_block}, "
                f"fill_pairs={self.fill_pairs}, count_block={self.count_block}, cell_sort={self.cell_sort}, "
                f"fused_small={self.fused_small})")


class PairListView(C.Structure):
    _fields_ = [
        ("nat", C.c_int64),
        ("npairs", C.c_int64),
        ("i", C.c_void_p),
        ("j", C.c_void_p),
        ("S", C.c_void_p),
        ("first", C.c_void_p),
        ("R", C.c_void_p),
        ("ti_bytes", C.c_int32),
    ]


class CellListView(C.Structure):
    _fields_ = [
        ("nat", C.c_int64),
        ("ncells_total", C.c_int64),
        ("X_orig", C.c_void_p),
        ("perm", C.c_void_p),
        ("cell_id", C.c_void_p),
        ("cell_offsets", C.c_void_p),
        ("ti_bytes", C.c_int32),
    ]


_lib = None
_lock = threading.Lock()

EXPORTS = [
    "nlh_init", "nlh_create", "nlh_destroy", "nlh_last_error", "nlh_synchronize", "nlh_get_stream",
    "nlh_geometry_init", "nlh_build_cell_list", "nlh_materialize", "nlh_neighbour_list",
    "nlh_pairlist_device", "nlh_celllist_device", "nlh_copy_pairlist", "nlh_copy_celllist",
    "nlh_neighbours", "nlh_neighbours_range", "nlh_batch_neighbour_list", "nlh_batch_pair_offsets", "nlh_batch_rebuild",
    "nlh_neighbour_list_noalloc", "nlh_set_tuning", "nlh_get_tuning", "nlh_autotune",
    "nlh_set_profiling", "nlh_phase_times", "nlh_graph_launches", "nlh_device_malloc", "nlh_device_free", "nlh_memcpy", "nlh_device_synchronize",
    "nlh_version",
]


def _declare(L):
    vp, dp, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)
    hp = C.c_void_p
    L.nlh_init.argtypes = [C.c_int]
    L.nlh_create.argtypes = [C.POINTER(C.c_void_p)]
    L.nlh_destroy.argtypes = [hp]
    L.nlh_last_error.restype = C.c_char_p
    L.nlh_synchronize.argtypes = [hp]
    L.nlh_get_stream.argtypes = [hp, C.POINTER(C.c_void_p)]
    L.nlh_geometry_init.argtypes = [dp, C.c_double, C.POINTER(C.c_uint8), C.c_int32, C.POINTER(Geometry)]
    L.nlh_build_cell_list.argtypes = [hp, vp, C.c_int64, C.POINTER(Geometry), C.c_uint32]
    L.nlh_materialize.argtypes = [hp, C.c_uint32, i64p]
    L.nlh_neighbour_list.argtypes = [hp, vp, C.c_int64, C.POINTER(Geometry), C.c_uint32, i64p]
    L.nlh_pairlist_device.argtypes = [hp, C.POINTER(PairListView)]
    L.nlh_celllist_device.argtypes = [hp, C.POINTER(CellListView)]
    L.nlh_copy_pairlist.argtypes = [hp, vp, vp, vp, vp, vp]
    L.nlh_copy_celllist.argtypes = [hp, vp, vp, vp, vp]
    L.nlh_neighbours.argtypes = [hp, C.c_int32, C.c_int64, C.c_int64, vp, vp, vp, i64p]
    L.nlh_neighbours_range.argtypes = [hp, C.c_int32, C.c_int64, C.c_int64, C.c_int64, i64p, vp, vp, vp]
    L.nlh_batch_neighbour_list.argtypes = [hp, C.c_int32, vp, i64p, C.POINTER(Geometry), C.c_uint32, i64p]
    L.nlh_batch_pair_offsets.argtypes = [hp, i64p]
    L.nlh_batch_rebuild.argtypes = [hp, vp, C.c_uint32, i64p]
    L.nlh_neighbour_list_noalloc.argtypes = [vp, vp, C.c_int64, C.POINTER(Geometry), C.c_uint32, C.c_int64, vp, vp,
                                             vp, vp, vp, i64p]
    L.nlh_set_tuning.argtypes = [hp, C.POINTER(Tuning)]
    L.nlh_get_tuning.argtypes = [hp, C.POINTER(Tuning)]
    L.nlh_autotune.argtypes = [hp, vp, C.c_uint32, C.c_int32, C.POINTER(Tuning)]
    L.nlh_set_profiling.argtypes = [hp, C.c_int]
    L.nlh_phase_times.argtypes = [hp, C.POINTER(C.c_float)]
    L.nlh_graph_launches.argtypes = [hp, C.POINTER(C.c_int64)]
    L.nlh_device_malloc.argtypes = [C.c_uint64, C.POINTER(C.c_void_p)]
    L.nlh_device_free.argtypes = [C.c_void_p]
    L.nlh_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
    L.nlh_device_synchronize.argtypes = []
    L.nlh_version.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    for name in EXPORTS:
        if name != "nlh_last_error":
            getattr(L, name).restype = C.c_int


def lib():
    """Load libnlhip.so (built in-tree).  Raises if it is missing."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                # torch ships its own HIP/HSA runtimes: load them first so a
                # later torch.cuda initialisation still sees the device
                try:
                    import torch  # noqa: F401
                except ImportError:
                    pass
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"libnlhip.so not found at {LIB_PATH}; build it first "
                        "(python -c 'import __graft_entry__ as g; g.build()')")
                L = C.CDLL(LIB_PATH)
                _declare(L)
                _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status == NLH_OK:
        return
    msg = lib().nlh_last_error().decode(errors="replace")
    msg = f"{what}: {msg}" if what else msg
    if status == NLH_EINVAL:
        raise NLHInvalidError(status, msg)
    if status == NLH_EOVERFLOW:
        raise NLHOverflowError(status, msg)
    if status == NLH_ENOMEM:
        raise NLHMemoryError(status, msg)
    raise NLHError(status, msg)


def ti_bytes_of(int_type) -> int:
    dt = np.dtype(int_type)
    if dt == np.int32:
        return 4
    if dt == np.int64:
        return 8
    raise ValueError("int_type must be Int32 (np.int32) or Int64 (np.int64)")


def ti_dtype(ti_bytes: int):
    return np.int32 if ti_bytes == 4 else np.int64


def cell_colmajor(cell) -> np.ndarray:
    """3x3 cell with ROWS = lattice vectors -> column-major 9 doubles
    (Julia SMatrix memory order)."""
    c = np.asarray(cell, dtype=np.float64)
    if c.shape != (3, 3):
        raise ValueError("cell must be 3x3 (rows are lattice vectors)")
    return np.ascontiguousarray(c.T.reshape(9))


def pbc_bytes(pbc) -> np.ndarray:
    if isinstance(pbc, (bool, np.bool_)):
        pbc = (pbc, pbc, pbc)
    p = np.asarray(pbc, dtype=bool).reshape(-1)
    if p.shape != (3,):
        raise ValueError("pbc must have 3 entries")
    return p.astype(np.uint8)


_Cell9 = C.c_double * 9
_Pbc3 = C.c_uint8 * 3


def geometry(cell, cutoff: float, pbc, ti_bytes: int = 4) -> Geometry:
    """nlh_geometry_init: analyze_cell + overflow check + nxyz.  (The arrays
    go over as ctypes copies of the numpy buffers: `ndarray.ctypes.data_as`
    costs several times the call itself, and this runs once per request.)"""
    g = Geometry()
    c = cell_colmajor(cell)
    p = pbc_bytes(pbc)
    check(lib().nlh_geometry_init(_Cell9.from_buffer_copy(c), float(cutoff), _Pbc3.from_buffer_copy(p),
                                  int(ti_bytes), C.byref(g)),
          "geometry")
    return g


_inited = False


def init(device: int = 0) -> None:
    global _inited
    check(lib().nlh_init(int(device)), "nlh_init")
    _inited = True


class Handle:
    """Owns one nlh_handle (a HIP stream + device buffers)."""

    def __init__(self):
        if not _inited:
            init(0)
        h = C.c_void_p()
        check(lib().nlh_create(C.byref(h)), "nlh_create")
        self._h = h

    @property
    def ptr(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().nlh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream(self) -> int:
        s = C.c_void_p()
        check(lib().nlh_get_stream(self._h, C.byref(s)))
        return s.value or 0

    def synchronize(self):
        check(lib().nlh_synchronize(self._h), "synchronize")

    def set_tuning(self, brick_atoms: int, fill_block: int, fill_pairs: int = 256, count_block: int = 0,
                   cell_sort: int = 0, fused_small: int = 0):
        t = Tuning(int(brick_atoms), int(fill_block), int(fill_pairs), int(count_block), int(cell_sort),
                   int(fused_small))
        check(lib().nlh_set_tuning(self._h, C.byref(t)), "set_tuning")

    def get_tuning(self) -> "Tuning":
        t = Tuning()
        check(lib().nlh_get_tuning(self._h, C.byref(t)), "get_tuning")
        return t

    def set_profiling(self, enable: bool = True):
        check(lib().nlh_set_profiling(self._h, 1 if enable else 0), "set_profiling")

    def graph_launches(self) -> int:
        """Builds on this handle that replayed the captured HIP graph."""
        n = C.c_int64()
        check(lib().nlh_graph_launches(self._h, C.byref(n)), "graph_launches")
        return int(n.value)

    def phase_times(self) -> dict:
        """ms of the last build: cells, count, scan, fill, total (HIP events)."""
        ms = (C.c_float * 5)()
        check(lib().nlh_phase_times(self._h, ms), "phase_times")
        return dict(zip(("cells", "count", "scan", "fill", "total"), [float(v) for v in ms]))


class DeviceArray:
    """Device memory owned through libnlhip's own HIP runtime (float64 or
    integer arrays kept resident in HBM)."""

    def __init__(self, host: np.ndarray):
        if not _inited:
            init(0)
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = int(host.nbytes)
        p = C.c_void_p()
        check(lib().nlh_device_malloc(max(self.nbytes, 16), C.byref(p)), "device_malloc")
        self.ptr = p.value
        self.upload(host)

    def upload(self, host: np.ndarray):
        h = np.ascontiguousarray(host, dtype=self.dtype)
        if h.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        if self.nbytes:
            check(lib().nlh_memcpy(self.ptr, h.ctypes.data, self.nbytes), "memcpy")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(lib().nlh_memcpy(out.ctypes.data, self.ptr, self.nbytes), "memcpy")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().nlh_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
