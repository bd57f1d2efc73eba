"""GPU parity: libnlhip vs the CPU oracle on the same seeded inputs.

Bar (integer/index work): bit-exact.  Because the device emits pairs in the
reference traversal order, the raw arrays i, j, S, first (and R, computed in
the reference's evaluation order) must equal the oracle's element for
element -- a stronger check than the reference's own set comparison
(test/test_utils.jl:127-131).
"""
import glob
import os

import numpy as np
import pytest

import nlhip
import oracle
from nlhip import workloads

pytestmark = pytest.mark.gpu

ALL_PBC = [(True, True, True), (False, False, False), (True, False, False), (False, True, False),
           (False, False, True), (True, True, False), (True, False, True), (False, True, True)]


def _same(nl, ref, R=True):
    assert nl.npairs() == ref.npairs
    assert np.array_equal(nl.first, ref.first)
    assert np.array_equal(nl.i, ref.i)
    assert np.array_equal(nl.j, ref.j)
    assert np.array_equal(nl.S, ref.S)
    if R:
        assert np.array_equal(np.asarray(nl.R).view(np.int64), ref.R.view(np.int64))


def _check(X, rc, C, pbc, int_type=np.int32):
    """Both builds: with R, and without (4-byte slot rows written through
    range-checked buffer stores, the path the benchmarks time)."""
    ti = 32 if np.dtype(int_type) == np.int32 else 64
    nl = nlhip.neighbour_list(X, rc, C, pbc, int_type=int_type, want_R=True)
    ref = oracle.neighbour_list(X, rc, C, pbc, ti_bits=ti, want_R=True)
    assert nl.i.dtype == np.dtype(int_type)
    _same(nl, ref)
    _same(nlhip.neighbour_list(X, rc, C, pbc, int_type=int_type), ref, R=False)
    return nl, ref


def test_edge_cases():
    # test/test_utils.jl:480-494
    C = np.eye(3) * 10.0
    assert nlhip.neighbour_list(np.array([[5.0, 5, 5]]), 3.0, C, (True,) * 3).npairs() == 0
    nl = nlhip.neighbour_list(np.array([[5.0, 5, 5], [5, 5, 6]]), 3.0, C, (True,) * 3)
    assert nl.npairs() == 2
    assert nlhip.max_neighbours(nl) == 1
    assert nlhip.neighbour_list(np.array([[1.0, 1, 1], [8, 8, 8]]), 3.0, C, (False,) * 3).npairs() == 0
    empty = nlhip.neighbour_list(np.zeros((0, 3)), 3.0, C, (True,) * 3)
    assert empty.npairs() == 0 and list(empty.first) == [1]


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_random_configs(seed):
    N = 50 + (seed * 37) % 151
    s = workloads.random_cubic(N, seed=seed)
    L = s.cell[0, 0]
    print(f"random N={N} seed={seed}", flush=True)
    _check(s.X, L / 3, s.cell, (True, True, True))


@pytest.mark.parametrize("pbc", ALL_PBC)
def test_all_pbc(pbc):
    s = workloads.random_cubic(50, seed=21)
    L = s.cell[0, 0]
    print(f"pbc={pbc}", flush=True)
    _check(s.X, 0.33 * L, s.cell, pbc)


@pytest.mark.parametrize("pbc", ALL_PBC)
def test_triclinic(pbc):
    C = np.array([[10.0, 2.0, 1.0], [0.0, 9.0, 1.5], [0.0, 0.0, 8.0]])   # test_utils.jl:297-299
    s = workloads.random_in_cell(80, C, seed=5)
    print(f"triclinic pbc={pbc}", flush=True)
    _check(s.X, 3.0, C, pbc)


def test_elongated_and_large_cutoff():
    C = np.diag([5.0, 5.0, 20.0])                                       # test_sortbased.jl:35-40
    s = workloads.random_in_cell(80, C, seed=6)
    _check(s.X, 3.0, C, (True, True, True))
    s = workloads.random_cubic(30, seed=8)                              # test_sortbased.jl:49-55
    _check(s.X, 0.6 * s.cell[0, 0], s.cell, (True, True, True))


def test_high_density_multi_image():
    # test_gpu.jl:78-86: L=5, 200 atoms, cutoff 2
    X = 5.0 * workloads.uniform(9, 600).reshape(200, 3)
    _check(X, 2.0, np.eye(3) * 5.0, (True, True, True))
    # cells far smaller than the cutoff: nxyz > 1 on every axis
    X = workloads.uniform(10, 12).reshape(4, 3) @ np.diag([2.2, 3.0, 2.5])
    nl, ref = _check(X, 6.0, np.diag([2.2, 3.0, 2.5]), (True, True, True))
    assert nl.npairs() > 100


@pytest.mark.parametrize("int_type", [np.int32, np.int64])
def test_int_types(int_type):
    s = workloads.random_cubic(100, seed=12)
    _check(s.X, s.cell[0, 0] / 3, s.cell, (True, True, True), int_type)


def test_c0_1000_atoms():
    s = workloads.random_cubic(1000, density=0.05, seed=100, cutoff=5.0)
    nl, ref = _check(s.X, 5.0, s.cell, s.pbc)
    assert 24000 < nl.npairs() < 28500


def test_atoms_outside_box():
    # test_sortbased.jl:143-228 (issue #6): per-atom +-L shifts on periodic
    # axes, and the frac-just-below-zero rounding trap
    for pbc in ALL_PBC:
        s = workloads.random_cubic(60, seed=31)
        L = s.cell[0, 0]
        idx = np.arange(1, 61)
        sh = np.stack([(idx % 3) - 1, ((idx + 1) % 3) - 1, ((idx + 2) % 3) - 1], 1).astype(np.float64)
        Xp = s.X + L * sh * np.array(pbc, dtype=np.float64)[None, :]
        _check(Xp, 0.25 * L, s.cell, pbc)
    L = 8.0
    tiny = -5e-17 * L
    X = np.array([[0.5, L - 0.3, 0.5], [0.5, tiny, 0.5]])
    nl, _ = _check(X, 1.5, np.eye(3) * L, (True, True, True))
    J, R, S = nlhip.neighbours(nl, 1)
    assert 2 in list(J)
    k = list(J).index(2)
    assert abs(np.sqrt((R[k] ** 2).sum()) - 0.3) < 1e-9


def test_far_windings():
    # windings beyond the packed 10-bit range take the wide path
    s = workloads.random_cubic(64, seed=41)
    L = s.cell[0, 0]
    Xp = s.X.copy()
    Xp[::3, 0] += 1000.0 * L
    Xp[1::3, 1] -= 777.0 * L
    _check(Xp, 0.3 * L, s.cell, (True, True, True))


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf, 1.0e300])
def test_non_finite_positions_are_refused(bad):
    """A position whose cell index does not fit (non-finite, or beyond 1e9
    cells; the reference's unsafe_trunc is undefined there, cell_list.jl:66-74)
    makes the build fail with NLH_EINVAL, on the single-system path and on the
    fused small-system build of a batch; the next builds are unaffected."""
    from nlhip.batch import BatchBuilder

    s = workloads.random_cubic(3000, seed=61)
    X = s.X.copy()
    X[1234, 1] = bad
    print(f"bad position {bad}: single system", flush=True)
    with pytest.raises(nlhip.NLHInvalidError):
        nlhip.neighbour_list(X, 5.0, s.cell, s.pbc)
    _check(s.X, 5.0, s.cell, s.pbc)
    systems = workloads.batch_systems(80, seed=62, nmax=600)
    bad_systems = list(systems)
    Xb = systems[17].X.copy()
    Xb[0, 2] = bad
    bad_systems[17] = workloads.System(Xb, systems[17].cell, systems[17].pbc, systems[17].cutoff)
    print(f"bad position {bad}: batch", flush=True)
    bb = BatchBuilder()
    with pytest.raises(nlhip.NLHInvalidError):
        bb.run(bad_systems)
    res = bb.run(systems, want_R=True)
    for k in (0, 17, 79):
        sysm = systems[k]
        _same(res.pairlist(k), oracle.neighbour_list(sysm.X, sysm.cutoff, sysm.cell, sysm.pbc, want_R=True))


def test_sorted_cell_list_fields():
    s = workloads.random_cubic(500, seed=51)
    cl = nlhip.build_cell_list(s.X, 5.0, s.cell, s.pbc)
    ref = oracle.neighbour_list(s.X, 5.0, s.cell, s.pbc)
    assert np.array_equal(cl.perm, ref.perm)
    assert np.array_equal(cl.cell_id, ref.cell_id)
    assert np.array_equal(cl.cell_offsets, ref.cell_offsets)
    assert np.array_equal(cl.X, s.X[ref.perm - 1])
    # lazy per-atom queries equal the materialised list (cell_list.jl:779-833)
    nl = nlhip.materialize_pairlist(cl, want_R=True)
    for i in (1, 2, 77, 500):
        js, Rs, Ss = nlhip.neighbours(cl, i)
        a, b = ref.first[i - 1] - 1, ref.first[i] - 1
        assert np.array_equal(js, ref.j[a:b])
        assert np.array_equal(Ss, ref.S[a:b])
        assert np.array_equal(Rs.view(np.int64), ref.R[a:b].view(np.int64))
        assert nlhip.count_neighbours(cl, i) == b - a
        assert nlhip.num_neighbours(nl, i) == b - a


@pytest.mark.parametrize("cell_sort", [0, 1])
def test_cell_sort_forms(cell_sort):
    """Both forms of the cell sort (buckets counted in LDS, or one global
    atomic histogram) give the reference's cell list and pairs: a clustered
    system (most atoms in a few cells of each bucket), a sparse grid (many
    empty buckets) and a batch whose buckets straddle system boundaries."""
    from nlhip import _lib
    from nlhip.batch import BatchBuilder, pack_systems

    rng = np.random.default_rng(95)
    L = 40.0
    Xc = np.concatenate([rng.random((3000, 3)) * 6.0 + 10.0, rng.random((1000, 3)) * L])
    Xs = rng.random((700, 3)) * 400.0
    for X, C, rc, pbc in [(Xc, np.eye(3) * L, 4.0, (True, True, True)),
                          (Xs, np.eye(3) * 400.0, 3.0, (True, False, True))]:
        h = _lib.Handle()
        h.set_tuning(768, 256, 256, 0, cell_sort)
        cl = nlhip.build_cell_list(X, rc, C, pbc, handle=h)
        ref = oracle.neighbour_list(X, rc, C, pbc, want_R=True)
        assert np.array_equal(cl.perm, ref.perm)
        assert np.array_equal(cl.cell_id, ref.cell_id)
        assert np.array_equal(cl.cell_offsets, ref.cell_offsets)
        _same(nlhip.materialize_pairlist(cl, want_R=True), ref)
    systems = workloads.batch_systems(40, seed=96, nmax=500)
    X0, offs = pack_systems(systems)
    bb = BatchBuilder()
    bb.handle.set_tuning(768, 256, 256, 0, cell_sort)
    res = bb.run(systems, want_R=True)
    for s in (0, 1, 17, 39):
        sysm = systems[s]
        _same(res.pairlist(s), oracle.neighbour_list(sysm.X, sysm.cutoff, sysm.cell, sysm.pbc, want_R=True))


def test_batch_matches_single():
    systems = workloads.batch_systems(64, seed=3, nmax=400)
    res = nlhip.batch_neighbour_list(systems, want_R=True)
    for s, sysm in enumerate(systems):
        ref = oracle.neighbour_list(sysm.X, sysm.cutoff, sysm.cell, sysm.pbc, want_R=True)
        pl = res.pairlist(s)
        _same(pl, ref)


@pytest.mark.parametrize("pbc", ALL_PBC)
def test_bricks_medium(pbc):
    # 20k atoms: 14^3 cells, many full 4x4x4 bricks plus clipped edge bricks
    s = workloads.random_cubic(20000, density=0.05, seed=61)
    print(f"bricks pbc={pbc}", flush=True)
    _check(s.X, 5.0, s.cell, pbc)


@pytest.mark.parametrize("brick_atoms,count_block", [(200, 256), (460, 256), (700, 256), (920, 256), (1840, 512)])
def test_brick_launch_order(brick_atoms, count_block):
    """Bricks launch largest shape class first, each class in row-major order
    (count pass).  A batch of boxes of 7 x 9 x 11 and 13 x 6 x 5 cells, whose
    axes end in partial bricks of different sizes, so that the classes (full,
    partial along one, two or three axes) differ in size and order, through
    the batched build (no minimum brick count: each budget keeps its largest
    brick, 1 x 1 x 1 to 4 x 4 x 4); every home cell is counted exactly once."""
    from nlhip.batch import BatchBuilder
    rng = np.random.default_rng(83)
    systems = []
    for dims, pbc in [((7, 9, 11), (True, True, True)), ((13, 6, 5), (False, True, False)),
                      ((7, 9, 11), (True, False, True))]:
        C = np.diag([d * (5.0 + 0.05 * k) for k, d in enumerate(dims)])
        X = rng.random((int(0.05 * np.prod(np.diag(C))), 3)) @ C
        systems.append(workloads.System(X, C, pbc, 5.0))
    print(f"brick order brick_atoms={brick_atoms} count_block={count_block}", flush=True)
    bb = BatchBuilder()
    bb.handle.set_tuning(brick_atoms, 256, 256, count_block, 0, 1)
    res = bb.run(systems, want_R=True)
    for k, sysm in enumerate(systems):
        _same(res.pairlist(k), oracle.neighbour_list(sysm.X, sysm.cutoff, sysm.cell, sysm.pbc, want_R=True))


def test_slot_rows_overflow_and_regrow():
    # 80% of the atoms packed into one eighth of the box: the slot rows sized
    # from the mean density (about 8 neighbours, 32 slots) are too short for
    # the clustered atoms (about 50), so the count drops the entries past each
    # row's end and the build repeats on longer rows
    rng = np.random.default_rng(71)
    L, n = 60.0, 20000
    nc = int(0.8 * n)
    X = np.concatenate([rng.random((nc, 3)) * (L / 2), rng.random((n - nc, 3)) * L])
    C = np.eye(3) * L
    nl, ref = _check(X, 2.74, C, (True, True, True))
    assert nlhip.max_neighbours(nl) > 40
    _check(X, 2.74, C, (False, True, False))


def test_bricks_dense_triclinic():
    C = np.array([[40.0, 6.0, 3.0], [0.0, 38.0, 5.0], [0.0, 0.0, 36.0]])
    s = workloads.random_in_cell(4000, C, seed=62)
    _check(s.X, 4.0, C, (True, True, True))
    _check(s.X, 4.0, C, (True, False, True))


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_fixture(path):
    """Every committed fixture, bit for bit (pairs, R and the cell list)."""
    with np.load(path, allow_pickle=False) as z:
        f = {k: z[k] for k in z.files}
    X, C, rc = f["X"], f["cell"], float(f["cutoff"])
    pbc = tuple(bool(b) for b in f["pbc"])
    print(f"golden {os.path.basename(path)} nat={X.shape[0]}", flush=True)
    nl = nlhip.neighbour_list(X, rc, C, pbc, want_R=True)
    assert nl.npairs() == f["i"].shape[0]
    for k in ("i", "j", "S", "first"):
        assert np.array_equal(getattr(nl, k), f[k]), k
    assert np.array_equal(np.asarray(nl.R).view(np.int64), f["R"].view(np.int64))
    cl = nlhip.build_cell_list(X, rc, C, pbc)
    for k in ("perm", "cell_id", "cell_offsets"):
        assert np.array_equal(getattr(cl, k), f[k]), k


def test_materialize_twice():
    """Repeated materialisation of one cell list (with and without R) gives
    the same pairs (cells smaller than the cutoff send every brick down the
    count pass's global-memory path)."""
    X = 5.0 * workloads.uniform(19, 300).reshape(100, 3)
    for rc in (2.0, 1.2):
        cl = nlhip.build_cell_list(X, rc, np.eye(3) * 5.0, (True, True, True))
        a = nlhip.materialize_pairlist(cl)
        b = nlhip.materialize_pairlist(cl, want_R=True)
        c = nlhip.materialize_pairlist(cl)
        ref = oracle.neighbour_list(X, rc, np.eye(3) * 5.0, (True, True, True), want_R=True)
        for nl in (a, b, c):
            assert np.array_equal(nl.i, ref.i) and np.array_equal(nl.j, ref.j) and np.array_equal(nl.S, ref.S)
        assert np.array_equal(np.asarray(b.R).view(np.int64), ref.R.view(np.int64))


def _sparse_with_dimers(ngrid, ndimer, seed):
    """Atoms on a grid of spacing 6 (no neighbour within 5) plus a partner at
    distance 1 for `ndimer` of them, shuffled: most atoms own no pairs, so
    a tile of output pairs spans thousands of atoms."""
    rng = np.random.default_rng(seed)
    g = int(round(ngrid ** (1 / 3)))
    ax = 6.0 * np.arange(g) + 3.0
    grid = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    pick = rng.choice(grid.shape[0], ndimer, replace=False)
    X = np.concatenate([grid, grid[pick] + np.array([1.0, 0.0, 0.0])])
    return X[rng.permutation(X.shape[0])], np.eye(3) * (6.0 * g)


@pytest.mark.parametrize("fill_pairs", [128, 256, 512])
def test_fill_tiles_over_many_empty_atoms(fill_pairs):
    """Fill tiles whose pairs lie far apart in atom order (the per-lane
    search path) and tiles next to it (the 64-atom window path)."""
    h = nlhip._lib.Handle()
    h.set_tuning(768, 256, fill_pairs)
    X, C = _sparse_with_dimers(8000, 700, seed=91)
    print(f"sparse dimers fill_pairs={fill_pairs}", flush=True)
    for pbc in [(True, True, True), (False, True, False)]:
        nl = nlhip.neighbour_list(X, 5.0, C, pbc, want_R=True, handle=h)
        ref = oracle.neighbour_list(X, 5.0, C, pbc, want_R=True)
        assert ref.npairs == 1400
        _same(nl, ref)
    # dense clusters among empty atoms: windows that just cover, or just miss, a tile
    rng = np.random.default_rng(92)
    Xg, Cg = _sparse_with_dimers(4096, 0, seed=93)
    cl = 30.0 + 1.5 * rng.standard_normal((300, 3))
    Xc = np.concatenate([Xg, cl])[rng.permutation(Xg.shape[0] + 300)]
    nl = nlhip.neighbour_list(Xc, 5.0, Cg, (True, True, True), want_R=True, handle=h)
    ref = oracle.neighbour_list(Xc, 5.0, Cg, (True, True, True), want_R=True)
    _same(nl, ref)


def _cell_lists_both_sorts(X, rc, C, pbc, ref):
    from nlhip import _lib

    for cs in (0, 1):
        h = _lib.Handle()
        h.set_tuning(920, 256, 256, 0, cs)
        cl = nlhip.build_cell_list(X, rc, C, pbc, handle=h)
        for k in ("perm", "cell_id", "cell_offsets"):
            assert np.array_equal(getattr(cl, k), getattr(ref, k)), (k, cs)
        h.close()


def test_scan_paths():
    """Both scan forms: a single pass up to 64 tiles of 2048 items (128 for
    the bucketed sort's counts), reduce -> tile sums -> downsweep beyond.
    300k atoms: 147 tiles of pair counts (three kernels), 23 of cells for the
    atomic sort (single pass); a sparse 100k-atom box: over 3 million cells
    (the atomic sort's three-kernel cell scan; 781 buckets of 4096 cells for
    the bucketed one), 49 tiles of pair counts (single pass)."""
    s = workloads.random_cubic(300_000, density=0.05, seed=95)
    print("scan paths: 300k cube", flush=True)
    nl = nlhip.neighbour_list(s.X, 5.0, s.cell, s.pbc)
    ref = oracle.neighbour_list(s.X, 5.0, s.cell, s.pbc)
    _same(nl, ref, R=False)
    _cell_lists_both_sorts(s.X, 5.0, s.cell, s.pbc, ref)
    s = workloads.random_cubic(100_000, density=0.002, seed=96, cutoff=2.5)
    print("scan paths: sparse 100k, 3.2M cells", flush=True)
    _, ref = _check(s.X, 2.5, s.cell, (True, False, True))
    _cell_lists_both_sorts(s.X, 2.5, s.cell, (True, False, True), ref)


def test_huge_grid_takes_the_atomic_sort():
    """A grid beyond the bucketed sort's reach (8192 buckets of 4096 cells):
    500 atoms (200 of them in one cluster) in a 2000 A box at cutoff 5 A
    have 400^3 = 64M cells, and the default sort falls back to the atomic
    histogram."""
    rng = np.random.default_rng(97)
    L = 2000.0
    X = np.concatenate([rng.random((300, 3)) * L, rng.random((200, 3)) * 20.0 + 100.0])   # a cluster with pairs
    C = np.eye(3) * L
    print("huge grid: 64M cells", flush=True)
    cl = nlhip.build_cell_list(X, 5.0, C, (True, True, True))
    ref = oracle.neighbour_list(X, 5.0, C, (True, True, True))
    for k in ("perm", "cell_id", "cell_offsets"):
        assert np.array_equal(getattr(cl, k), getattr(ref, k)), k
    _same(nlhip.materialize_pairlist(cl), ref, R=False)


def test_repeated_builds_replay_the_pipeline():
    """Repeated single builds of one shape on a handle replay the captured
    pipeline (from the first repeat on); every result equals the oracle:
    new positions each call, a cell scaled within the same cell grid, another
    shape built in between, and a call that asks for the stages one by one."""
    h = nlhip._lib.Handle()
    s = workloads.random_cubic(3000, seed=97)            # L/rc = 7.83: 7 cells per axis
    other = workloads.random_cubic(1000, seed=98)
    for rep in range(6):
        X = s.X + 0.05 * workloads.normal(300 + rep, s.X.size).reshape(s.X.shape)
        C = s.cell * (1.0 + 0.001 * rep)
        print(f"repeat {rep} graph launches {h.graph_launches()}", flush=True)
        nl = nlhip.neighbour_list(X, 5.0, C, s.pbc, want_R=True, handle=h)
        _same(nl, oracle.neighbour_list(X, 5.0, C, s.pbc, want_R=True))
        if rep == 3:
            nl = nlhip.neighbour_list(other.X, 5.0, other.cell, other.pbc, want_R=True, handle=h)
            _same(nl, oracle.neighbour_list(other.X, 5.0, other.cell, other.pbc, want_R=True))
    # calls 2, 3, 4 (repeats) and 6 (the second repeat after the other shape)
    assert h.graph_launches() == 4
    n0 = h.graph_launches()
    X = s.X + 0.05 * workloads.normal(400, s.X.size).reshape(s.X.shape)
    nl = nlhip.neighbour_list(X, 5.0, s.cell, s.pbc, want_R=True, handle=h, no_graph=True)
    _same(nl, oracle.neighbour_list(X, 5.0, s.cell, s.pbc, want_R=True))
    assert h.graph_launches() == n0


@pytest.mark.parametrize("nat,replayed", [(200_000, False), (8_000, True)])
def test_builds_return_with_the_fill_in_flight(nat, replayed):
    """A build returns once its fill has published the build status; the
    fill may still be writing pairs.  The pair count is already final, and
    the device views wait for the fill: here they are read at once with plain
    copies that do not follow the handle's stream.  Repeats of the smaller
    system replay the captured pipeline, the larger one launches its stages
    directly.  The positions alternate between two sets, written into one
    device array (a captured pipeline reads them from a fixed address)."""
    import ctypes as C
    _lib = nlhip._lib
    L = _lib.lib()
    s = workloads.random_cubic(nat, seed=5)
    X2 = s.X + 0.3 * workloads.normal(77, s.X.size).reshape(s.X.shape)
    refs = [oracle.neighbour_list(X, s.cutoff, s.cell, s.pbc) for X in (s.X, X2)]
    Xs = [np.ascontiguousarray(X) for X in (s.X, X2)]
    d = _lib.DeviceArray(Xs[0])
    g = _lib.geometry(s.cell, s.cutoff, s.pbc, 4)
    h = _lib.Handle()
    n = C.c_int64()

    def copy(ptr, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            _lib.check(L.nlh_memcpy(out.ctypes.data, ptr, out.nbytes), "memcpy")
        return out

    for k in range(12):
        ref = refs[k % 2]
        d.upload(Xs[k % 2])   # a blocking copy; the previous build reads no positions after it returns
        n0 = h.graph_launches()
        _lib.check(L.nlh_neighbour_list(h.ptr, d.ptr, s.X.shape[0], C.byref(g), _lib.NLH_X_DEVICE, C.byref(n)),
                   "neighbour_list")
        assert n.value == ref.npairs
        if k in (5, 8, 11):
            assert h.graph_launches() == n0 + (1 if replayed else 0)
            print(f"build {k}: reading the views at once", flush=True)
            v = _lib.PairListView()
            _lib.check(L.nlh_pairlist_device(h.ptr, C.byref(v)), "pairlist_device")
            P = int(v.npairs)
            assert P == ref.npairs
            assert np.array_equal(copy(v.first, s.X.shape[0] + 1, np.int32), ref.first)
            assert np.array_equal(copy(v.i, P, np.int32), ref.i)
            assert np.array_equal(copy(v.j, P, np.int32), ref.j)
            assert np.array_equal(copy(v.S, (P, 3), np.int32), ref.S)
    if replayed:
        assert h.graph_launches() >= 9
    else:
        assert h.graph_launches() == 0
    h.close()
    d.free()


def test_lazy_sweep_over_every_atom():
    """for_each_neighbour / neighbours over all 100,000 atoms of a cell list
    (cell_list.jl:779-833) through the range query: a few device passes,
    under a second, each atom's neighbours bit for bit the oracle's slice of
    the materialised list (j, S, R and the per-atom counts)."""
    import time

    s = workloads.random_cubic(100_000, density=0.05, seed=1, cutoff=5.0)
    ref = oracle.neighbour_list(s.X, 5.0, s.cell, s.pbc, want_R=True, nthreads=os.cpu_count() or 1)
    cl = nlhip.build_cell_list(s.X, 5.0, s.cell, s.pbc)
    t0 = time.perf_counter()
    offs, js, Rs, Ss = nlhip.neighbours_range(cl, 1, 100_000)
    dt = time.perf_counter() - t0
    print(f"lazy sweep of 100k atoms: {dt:.3f} s", flush=True)
    assert dt < 1.0
    assert np.array_equal(offs + 1, ref.first)
    assert np.array_equal(js, ref.j) and np.array_equal(Ss, ref.S)
    assert np.array_equal(Rs.view(np.int64), ref.R.view(np.int64))
    # the per-atom Python API reads blocks of the same query
    t0 = time.perf_counter()
    n = 0
    for i in range(1, 100_001):
        n += nlhip.count_neighbours(cl, i)
    print(f"count_neighbours over 100k atoms: {time.perf_counter() - t0:.3f} s", flush=True)
    assert n == ref.npairs
    for i in (1, 65536, 65537, 99_999, 100_000):
        j, R, S = nlhip.neighbours(cl, i)
        a, b = ref.first[i - 1] - 1, ref.first[i] - 1
        assert np.array_equal(j, ref.j[a:b]) and np.array_equal(S, ref.S[a:b])
        assert np.array_equal(R.view(np.int64), ref.R[a:b].view(np.int64))
    with pytest.raises(nlhip.NLHInvalidError):
        nlhip.neighbours(cl, 100_001)


@pytest.mark.parametrize("nsys", [24, 96])
def test_lazy_range_in_a_batch(nsys):
    """The range query on one system of a batch (system-local indices); 96
    systems take the fused small-system build, whose cell list it reads."""
    import ctypes as C

    from nlhip import _lib
    from nlhip.batch import BatchBuilder

    systems = workloads.batch_systems(nsys, seed=151, nmax=700)
    bb = BatchBuilder(np.int64)
    bb.run(systems, copy=False)
    L = _lib.lib()
    for sidx in (0, 5, nsys - 1):
        sysm = systems[sidx]
        ref = oracle.neighbour_list(sysm.X, sysm.cutoff, sysm.cell, sysm.pbc, ti_bits=64, want_R=True)
        n = sysm.nat
        offs = np.zeros(n + 1, np.int64)
        op = offs.ctypes.data_as(C.POINTER(C.c_int64))
        _lib.check(L.nlh_neighbours_range(bb.handle.ptr, sidx, 1, n, 0, op, None, None, None))
        P = int(offs[-1])
        assert P == ref.npairs
        j = np.empty(max(P, 1), np.int64)
        S = np.empty((max(P, 1), 3), np.int64)
        R = np.empty((max(P, 1), 3), np.float64)
        _lib.check(L.nlh_neighbours_range(bb.handle.ptr, sidx, 1, n, P, op, j.ctypes.data, R.ctypes.data,
                                          S.ctypes.data))
        assert np.array_equal(offs + 1, ref.first)
        assert np.array_equal(j[:P], ref.j) and np.array_equal(S[:P], ref.S)
        assert np.array_equal(R[:P].view(np.int64), ref.R.view(np.int64))
        # a sub-range, and the bounds check
        k = max(n // 3, 1)
        o2 = np.zeros(k + 1, np.int64)
        _lib.check(L.nlh_neighbours_range(bb.handle.ptr, sidx, n - k + 1, k, 0,
                                          o2.ctypes.data_as(C.POINTER(C.c_int64)), None, None, None))
        assert np.array_equal(o2, ref.first[n - k:] - ref.first[n - k])
    assert L.nlh_neighbours_range(bb.handle.ptr, 0, 2, systems[0].nat, 0, op, None, None, None) == _lib.NLH_EINVAL
