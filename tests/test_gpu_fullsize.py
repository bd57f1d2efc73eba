"""GPU parity at the benchmarked sizes: the same workloads bench.py times,
each compared with the CPU oracle element for element.

* the 1,000,000-atom cube (L = 271.44 A, 54^3 cells, ~26.2 M pairs);
* the 1,000,188-atom sheared FCC slab, pbc (T, T, F), 30 A of vacuum, most
  atoms with a fractional x outside [0, 1) (~42.4 M pairs);
* the 4,096-system batch (8..2,048 atoms, random triclinic cells and pbc,
  cutoffs 4..6 A, 10% cells smaller than the cutoff), through the first
  build, the autotuner and graph-replayed rebuilds, every system against the
  single-system oracle;
* a served stream up to 200,000 atoms through a default Server (systems of
  65,536 atoms and more resident), each result equal to the plain call.

The bar: i, j, S, first, perm, cell_id and cell_offsets equal element for
element, R equal bit for bit (the reference's own GPU tests compare sets,
test/test_gpu.jl:41-86).
"""
import os

import numpy as np
import pytest

import nlhip
import oracle
from nlhip import _lib, serve, workloads
from nlhip.batch import BatchBuilder, pack_systems

pytestmark = pytest.mark.gpu

NT = min(os.cpu_count() or 1, 16)


def _bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def _full_size(sysm, label):
    print(f"{label}: oracle ({sysm.nat} atoms, {NT} threads)", flush=True)
    ref = oracle.neighbour_list(sysm.X, sysm.cutoff, sysm.cell, sysm.pbc, want_R=True, nthreads=NT)
    print(f"{label}: {ref.npairs} pairs; GPU build without R", flush=True)
    h = _lib.Handle()
    nl = nlhip.neighbour_list(sysm.X, sysm.cutoff, sysm.cell, sysm.pbc, handle=h)
    assert nl.npairs() == ref.npairs
    for k in ("first", "i", "j", "S"):
        assert np.array_equal(getattr(nl, k), getattr(ref, k)), k
    del nl
    print(f"{label}: GPU build with R", flush=True)
    nl = nlhip.neighbour_list(sysm.X, sysm.cutoff, sysm.cell, sysm.pbc, want_R=True, handle=h)
    assert np.array_equal(nl.j, ref.j) and np.array_equal(nl.S, ref.S)
    assert np.array_equal(_bits(nl.R), _bits(ref.R))
    del nl
    print(f"{label}: cell list", flush=True)
    cl = nlhip.build_cell_list(sysm.X, sysm.cutoff, sysm.cell, sysm.pbc, handle=h)
    for k in ("perm", "cell_id", "cell_offsets"):
        assert np.array_equal(getattr(cl, k), getattr(ref, k)), k
    h.close()
    return ref


def test_cube_1M():
    s = workloads.random_cubic(1_000_000, density=0.05, seed=2, cutoff=5.0)
    ref = _full_size(s, "cube 1M")
    assert 25_000_000 < ref.npairs < 27_500_000


def test_fcc_slab_1M():
    s = workloads.triclinic_fcc_slab()
    assert s.nat == 1_000_188 and s.pbc == (True, True, False)
    # the slab's atoms stay in the orthogonal box: many lie outside the
    # sheared cell (fractional x < 0 or >= 1), so binning wraps them
    frac = s.X @ np.linalg.inv(s.cell)
    assert (frac[:, 0] < 0).sum() > 100_000
    ref = _full_size(s, "fcc slab")
    assert 42_000_000 < ref.npairs < 43_000_000
    # the vacuum: no pair crosses the open z boundary
    assert not ref.S[:, 2].any()


def _check_systems(res, systems, X, offs, label):
    for s, sysm in enumerate(systems):
        Xs = X[offs[s]:offs[s + 1]]
        ref = oracle.neighbour_list(Xs, sysm.cutoff, sysm.cell, sysm.pbc, want_R=False)
        pl = res.pairlist(s)
        assert pl.npairs() == ref.npairs, (label, s)
        assert np.array_equal(pl.first, ref.first), (label, s)
        assert np.array_equal(pl.i, ref.i) and np.array_equal(pl.j, ref.j), (label, s)
        assert np.array_equal(pl.S, ref.S), (label, s)


def test_batch_4096_systems():
    systems = workloads.batch_systems(4096, seed=7)
    X, offs = pack_systems(systems)
    assert len(systems) == 4096 and sum(1 for s in systems if s.nat <= 4) > 200
    bb = BatchBuilder()
    print("batch 4096: first build", flush=True)
    res = bb.run(systems)
    _check_systems(res, systems, X, offs, "run")
    print("batch 4096: autotune", flush=True)
    best = bb.autotune(X, reps=1)
    print(f"batch 4096: tuned {best}; graph rebuilds", flush=True)
    Xd = _lib.DeviceArray(X)
    for rep in range(3):
        res = bb.rebuild(X, X_device=Xd)
        assert bb.handle.graph_launches() >= rep + 1
    _check_systems(res, systems, X, offs, "rebuild")
    # new positions in the same device buffer: the replay reads them
    X2 = X + 0.05 * workloads.normal(161, X.size).reshape(X.shape)
    Xd.upload(X2)
    res = bb.rebuild(X2, X_device=Xd)
    _check_systems(res, systems, X2, offs, "rebuild, new positions")
    Xd.free()
    bb.handle.close()


def test_served_stream_to_200k_atoms():
    """A default Server (residents from 65,536 atoms, batches of up to 2^19
    atoms, 4 workers) against the plain call, request by request."""
    reqs = workloads.served_stream(120, rate=1e4, seed=11, nmax=200_000)
    big = [r for r in reqs if r.nat >= 65536]
    assert len(big) >= 5 and any(r.rebuild for r in big) and max(r.nat for r in reqs) > 150_000
    srv = serve.Server(keep_results=True)
    print("served stream: mixed", flush=True)
    mixed = srv.serve(reqs, realtime=False)
    assert len(srv.residents) > 0
    srv.close()
    print("served stream: alone", flush=True)
    alone = serve.serve_plain(reqs, realtime=False, keep_results=True)
    assert max(c.batch for c in mixed) > 1
    for m, a, r in zip(mixed, alone, reqs):
        assert m.npairs == a.npairs, r.rid
        assert np.array_equal(m.result.first, a.result.first), r.rid
        assert np.array_equal(m.result.i, a.result.i) and np.array_equal(m.result.j, a.result.j), r.rid
        assert np.array_equal(m.result.S, a.result.S), r.rid
    # the largest resident rebuild and the largest batched request, against the oracle
    for r in (max((q for q in reqs if q.rebuild and q.nat >= 65536), key=lambda q: q.nat),
              max((q for q in reqs if q.nat < 65536), key=lambda q: q.nat)):
        print(f"served stream: request {r.rid} ({r.nat} atoms) against the oracle", flush=True)
        ref = oracle.neighbour_list(r.X, r.cutoff, r.cell, r.pbc, want_R=False, nthreads=NT)
        m = mixed[r.rid]
        assert np.array_equal(m.result.j, ref.j) and np.array_equal(m.result.S, ref.S)
        assert np.array_equal(m.result.first, ref.first)
