"""Host-side checks that need no GPU.

* libnlhip.so loads and exports every function include/nlhip.h declares;
* nlh_geometry_init (the host restatement of analyze_cell, the index-type
  overflow check and the neighbour-cell reach; cell_list.jl:152-170,
  655-659, gpu_kernels.jl:315-317) equals the oracle bit for bit;
* argument errors map to the reference's exceptions;
* the seeded generators reproduce published SplitMix64 values;
* without a device the product path raises instead of computing on the CPU;
* bench.py refuses more than one GPU.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from nlhip import _lib, workloads
from nlhip.batch import pack_systems

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "nlhip.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nlh_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_abi():
    names = _declared()
    for must in ("nlh_geometry_init", "nlh_build_cell_list", "nlh_materialize", "nlh_neighbour_list",
                 "nlh_copy_pairlist", "nlh_copy_celllist", "nlh_batch_neighbour_list", "nlh_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(_declared()) == set(_lib.EXPORTS)


def test_version():
    a, b = C.c_int32(), C.c_int32()
    assert _lib.lib().nlh_version(C.byref(a), C.byref(b)) == 0
    assert a.value >= 0 and b.value >= 0


CELLS = [
    np.eye(3) * 125.99210498948732,
    np.array([[10.0, 2.0, 1.0], [0.0, 9.0, 1.5], [0.0, 0.0, 8.0]]),
    np.diag([3.84666089, 3.84666089, -5.44]),                            # left-handed
    np.array([[5.71757, -1.81834e-15, 9.74255e-41], [-2.85879, 4.95156, 4.93924e-25],
              [4.56368e-40, 9.05692e-25, 9.05629]]),                     # near-degenerate
    np.array([[227.43, 0.0, 0.0], [68.59, 227.43, 0.0], [45.486, 0.0, 257.43]]),
]


@pytest.mark.parametrize("k", range(len(CELLS)))
@pytest.mark.parametrize("rc", [0.7, 3.0, 5.0, 40.0])
def test_geometry_matches_oracle(k, rc):
    cell = CELLS[k]
    for ti in (4, 8):
        g = _lib.geometry(cell, rc, (True, False, True), ti)
        o = oracle.analyze(cell, rc, (True, False, True), ti * 8)
        inv = np.array(g.inv_cell[:]).reshape(3, 3).T
        assert np.array_equal(inv.view(np.int64), o.inv_cell.view(np.int64))
        assert np.array_equal(np.array(g.lens[:]).view(np.int64), o.lens.view(np.int64))
        assert list(g.ncells) == list(o.ncells) and list(g.nxyz) == list(o.nxyz)
        assert g.ncells_total == o.ncells_total
        assert g.cutoff_sq == o.cutoff_sq
        assert list(g.pbc) == [1, 0, 1] and g.ti_bytes == ti
        assert np.array_equal(np.array(g.cell[:]), cell.T.reshape(9))    # column-major, SMatrix order


def test_geometry_errors():
    with pytest.raises(_lib.NLHOverflowError, match="Int64"):
        _lib.geometry(np.eye(3) * 1e4, 0.01, (True,) * 3, 4)
    assert _lib.geometry(np.eye(3) * 1e4, 0.01, (True,) * 3, 8).ncells_total == 10 ** 18
    with pytest.raises(_lib.NLHInvalidError):
        _lib.geometry(np.eye(3), -1.0, (True,) * 3, 4)
    with pytest.raises(_lib.NLHInvalidError):
        _lib.geometry(np.eye(3), 1.0, (True,) * 3, 2)
    with pytest.raises(_lib.NLHInvalidError):
        _lib.geometry(np.full((3, 3), np.nan), 1.0, (True,) * 3, 4)
    with pytest.raises(ValueError):
        _lib.geometry(np.eye(2), 1.0, (True,) * 3, 4)


def test_null_arguments_are_rejected():
    L = _lib.lib()
    assert L.nlh_geometry_init(None, 1.0, None, 4, None) == _lib.NLH_EINVAL
    assert L.nlh_materialize(None, 0, None) == _lib.NLH_EINVAL
    assert L.nlh_destroy(None) in (0, _lib.NLH_EINVAL)
    assert len(L.nlh_last_error()) > 0


def test_splitmix64_published_values():
    # SplitMix64 (Steele, Lea, Flood 2014), seed 1234567
    ref = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
           16408922859458223821]
    assert [int(x) for x in workloads.splitmix64(1234567, 5)] == ref
    u = workloads.uniform(1, 100000)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


def test_workload_shapes():
    s = workloads.random_cubic(100_000, density=0.05, seed=1)
    assert abs(s.cell[0, 0] - 125.99210498948732) < 1e-9 and s.X.shape == (100_000, 3)
    slab = workloads.triclinic_fcc_slab(n=4)
    assert slab.X.shape == (4 * 4 ** 3, 3) and slab.pbc == (True, True, False)
    systems = workloads.batch_systems(64, seed=3, nmax=400)
    X, offs = pack_systems(systems)
    assert offs[0] == 0 and offs[-1] == X.shape[0] and np.all(np.diff(offs) >= 1)
    tiny = [s for s in systems if s.nat <= 4]
    assert tiny, "the batch mix includes tiny multi-image systems"


def test_no_cpu_fallback_without_device():
    """The product path must raise when no HIP device is usable."""
    try:
        _lib.init(0)
    except _lib.NLHError:
        pass
    else:
        pytest.skip("a HIP device is present")
    import nlhip
    with pytest.raises(_lib.NLHError):
        nlhip.neighbour_list(np.zeros((2, 3)), 1.0, np.eye(3) * 5, (True,) * 3)


def _header_arity():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for name, args in re.findall(r"\b(?:int|const char \*)\s*(nlh_[a-z0-9_]+)\s*\(([^;]*?)\);", src, flags=re.S):
        a = args.strip()
        out[name] = 0 if a in ("", "void") else a.count(",") + 1
    return out


def _julia_ccalls(src):
    """(symbol, number of argument types) of every ccall in Julia source."""
    out = []
    for m in re.finditer(r"ccall\(\(:(nlh_[a-z0-9_]+), libnlhip\),\s*\w+,\s*\(", src):
        depth, k = 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(src[k], 0)
            k += 1
        types = src[m.end():k - 1].strip()
        n = 0 if not types else len([t for t in re.split(r",(?![^{]*\})", types) if t.strip()])
        out.append((m.group(1), n))
    return out


def test_julia_ccalls_match_the_header():
    """Every ccall in the Julia sources binds a declared entry point with the
    header's number of arguments."""
    arity = _header_arity()
    jdir = os.path.join(REPO, "neighbourlists.jl_amd", "julia")
    for f in ("NeighbourListsHIP.jl", "nlhip_abi.jl", os.path.join("ext", "NeighbourListsAMDGPUExt.jl")):
        calls = _julia_ccalls(open(os.path.join(jdir, f)).read())
        assert calls, f
        for name, n in calls:
            assert name in arity, (f, name)
            assert n == arity[name], (f, name, n, arity[name])


def test_julia_extension_plugs_into_the_reference_dispatch():
    """The AMDGPU package extension adds methods at the reference's own
    dispatch point (_build_sorted_celllist on a ROCVector with ROCBackend,
    materialize_pairlist on a SortedCellList of ROCVectors), returns the
    reference's types over device arrays made by similar(X, ...), and the
    PairList aliases X (src/gpu_kernels.jl:347-362)."""
    ext = open(os.path.join(REPO, "neighbourlists.jl_amd", "julia", "ext", "NeighbourListsAMDGPUExt.jl")).read()
    assert re.search(r"function NeighbourLists\._build_sorted_celllist\(X::ROCVector\{SVec\{Float64\}\}, "
                     r"cell::SMat\{Float64\},\s*pbc::SVec\{Bool\}, cutoff::Float64, ::Type\{TI\},\s*"
                     r"backend::ROCBackend\)", ext)
    assert re.search(r"function NeighbourLists\.materialize_pairlist\(clist::SortedCellList\{Float64,TI,AT,VI\};",
                     ext)
    assert "where {TI, AT <: ROCVector, VI}" in ext
    called = {n for n, _ in _julia_ccalls(ext)}
    assert {"nlh_build_cell_list", "nlh_copy_celllist", "nlh_materialize", "nlh_copy_pairlist",
            "nlh_neighbours_range"} <= called
    assert "NLH_X_DEVICE" in ext and "pointer(X)" in ext
    for arr in ("i_arr = similar(clist.X_orig, TI, P)", "S_arr = similar(clist.X_orig, SVec{TI}, P)",
                "perm = similar(X, TI, nat)"):
        assert arr in ext
    assert re.search(r"PairList\{Float64, TI, AT, typeof\(i_arr\), typeof\(S_arr\)\}\(clist\.X_orig,", ext)
    assert "analyze_cell(cell, cutoff, TI)" in ext


def test_julia_shim_binds_declared_symbols():
    """The Julia ccall shim binds only entry points the header declares, and
    its geometry struct has the C layout (no Julia here: checked as text)."""
    jdir = os.path.join(REPO, "neighbourlists.jl_amd", "julia")
    jl = open(os.path.join(jdir, "NeighbourListsHIP.jl")).read() + open(os.path.join(jdir, "nlhip_abi.jl")).read()
    called = set(re.findall(r"ccall\(\(:(nlh_[a-z0-9_]+),", jl))
    assert called and called <= set(_declared())
    assert {"nlh_geometry_init", "nlh_build_cell_list", "nlh_materialize", "nlh_copy_pairlist",
            "nlh_copy_celllist", "nlh_neighbours", "nlh_batch_neighbour_list"} <= called
    sizes = {"Float64": 8, "Int64": 8, "UInt8": 1, "Int32": 4}
    body = jl[jl.index("struct NLHGeometry"):]
    body = body[:body.index("\nend")]
    total = 0
    for name, typ in re.findall(r"^\s+(\w+)::(.+)$", body, flags=re.M):
        m = re.match(r"NTuple\{(\d+),(\w+)\}", typ)
        total += int(m.group(1)) * sizes[m.group(2)] if m else sizes[typ.strip()]
    assert total == C.sizeof(_lib.Geometry) == 248


def test_bench_measures_one_gpu_only():
    # the bench's contract: more than one GPU is refused with a message,
    # before anything touches a device
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
