"""GPU parity over randomly drawn geometries.

Each case draws a triclinic cell (sheared, elongated, right- or left-handed),
one of the eight pbc patterns, a density, a cutoff from a few neighbours per
atom up to 2.5 face heights (several periodic images per pair), atoms inside
and, for some, lattice vectors outside the cell, Int32 or Int64 indices, with
or without R.  The cell list (perm, cell_id, cell_offsets, sorted X) and the
pair list (i, j, S, first, R) are compared element for element with the
oracle, the C restatement of the reference's sort-based path
(cell_list.jl:632-748, gpu_kernels.jl:20-364).  The batch cases put 80 such
small systems through one batched build (the fused small-system kernel when
every system fits its limits) and check every system; the lazy cases query
every atom's neighbours from the cell list.
"""
import numpy as np
import pytest

import nlhip
import oracle
from nlhip import workloads

pytestmark = pytest.mark.gpu

ALL_PBC = [(True, True, True), (False, False, False), (True, False, False), (False, True, False),
           (False, False, True), (True, True, False), (True, False, True), (False, True, True)]


def _random_system(seed: int, nmax: int, kmin: float):
    u = workloads.uniform(7000 + seed, 24)
    N = max(1, int(round(np.exp(u[0] * np.log(nmax)))))
    rho = 0.02 + 0.08 * u[1]
    d = 0.4 + 1.6 * u[2:5]                                   # axis lengths, relative
    C = np.diag(d)
    C[0, 1], C[0, 2], C[1, 2] = (u[5:8] - 0.5) * 0.6 * d[:3][[0, 0, 1]]
    if u[8] < 0.2:
        C = C[[1, 0, 2]]                                     # left-handed
    C *= (N / rho / abs(np.linalg.det(C))) ** (1.0 / 3.0)
    pbc = ALL_PBC[int(u[9] * 8)]
    vol = abs(np.linalg.det(C))
    h = min(vol / np.linalg.norm(np.cross(C[(k + 1) % 3], C[(k + 2) % 3])) for k in range(3))
    k_nb = kmin + (100.0 - kmin) * u[10]
    rc = min((k_nb / (rho * 4.18879020478639)) ** (1.0 / 3.0), 2.5 * h)
    X = workloads.uniform(7100 + seed, 3 * N).reshape(N, 3) @ C
    if u[11] < 0.3:                                         # a tenth of the atoms moved by lattice vectors
        m = workloads.uniform(7200 + seed, N) < 0.1
        shift = np.floor(workloads.uniform(7300 + seed, 3 * N).reshape(N, 3) * 5.0) - 2.0
        X[m] += shift[m] @ C
    int_type = np.int64 if u[12] < 0.25 else np.int32
    return workloads.System(X, C, pbc, float(rc)), int_type, bool(u[13] < 0.5)


@pytest.mark.parametrize("seed", range(64))
def test_random_geometry(seed):
    s, int_type, want_R = _random_system(seed, 20000, 2.0)
    print(f"sweep seed={seed} N={s.X.shape[0]} pbc={s.pbc} rc={s.cutoff:.3f} ti={np.dtype(int_type).name} "
          f"R={want_R}", flush=True)
    ti = 32 if int_type == np.int32 else 64
    ref = oracle.neighbour_list(s.X, s.cutoff, s.cell, s.pbc, ti_bits=ti, want_R=want_R, nthreads=8)
    cl = nlhip.build_cell_list(s.X, s.cutoff, s.cell, s.pbc, int_type=int_type)
    assert np.array_equal(cl.perm, ref.perm)
    assert np.array_equal(cl.cell_id, ref.cell_id)
    assert np.array_equal(cl.cell_offsets, ref.cell_offsets)
    assert np.array_equal(cl.X, s.X[ref.perm - 1])
    nl = nlhip.materialize_pairlist(cl, want_R=want_R)
    assert nl.npairs() == ref.npairs
    assert nl.i.dtype == np.dtype(int_type)
    assert np.array_equal(nl.first, ref.first)
    assert np.array_equal(nl.i, ref.i)
    assert np.array_equal(nl.j, ref.j)
    assert np.array_equal(nl.S, ref.S)
    if want_R:
        assert np.array_equal(np.asarray(nl.R).view(np.int64), ref.R.view(np.int64))


@pytest.mark.parametrize("seed", range(8))
def test_random_geometry_batch(seed):
    systems = [_random_system(1000 * (seed + 1) + k, 2048, 6.0)[0] for k in range(80)]
    print(f"sweep batch seed={seed} atoms={sum(x.X.shape[0] for x in systems)}", flush=True)
    res = nlhip.batch_neighbour_list(systems, want_R=True)
    for k, sysm in enumerate(systems):
        ref = oracle.neighbour_list(sysm.X, sysm.cutoff, sysm.cell, sysm.pbc, want_R=True)
        pl = res.pairlist(k)
        assert pl.npairs() == ref.npairs, k
        assert np.array_equal(pl.first, ref.first), k
        assert np.array_equal(pl.i, ref.i) and np.array_equal(pl.j, ref.j) and np.array_equal(pl.S, ref.S), k
        assert np.array_equal(np.asarray(pl.R).view(np.int64), ref.R.view(np.int64)), k


@pytest.mark.parametrize("seed", range(16))
def test_random_geometry_lazy(seed):
    """The host lazy API (neighbours over a range of atoms, cell_list.jl:779-833)
    on the cell list of each draw: every atom's j, S and R equal the oracle's
    slice of the materialised list."""
    s, int_type, _ = _random_system(100 + seed, 20000, 2.0)
    nat = s.X.shape[0]
    print(f"sweep lazy seed={seed} N={nat} pbc={s.pbc}", flush=True)
    ti = 32 if int_type == np.int32 else 64
    ref = oracle.neighbour_list(s.X, s.cutoff, s.cell, s.pbc, ti_bits=ti, want_R=True, nthreads=8)
    cl = nlhip.build_cell_list(s.X, s.cutoff, s.cell, s.pbc, int_type=int_type)
    offs, js, Rs, Ss = nlhip.neighbours_range(cl, 1, nat)
    assert np.array_equal(offs + 1, ref.first)
    assert np.array_equal(js, ref.j) and np.array_equal(Ss, ref.S)
    assert np.array_equal(np.asarray(Rs).view(np.int64), ref.R.view(np.int64))
