"""CPU checks that pin the oracle (no GPU).

The oracle (a C restatement of the reference's sort-based build) is checked
against what fixes the answer independently of it:

* the known answers of the reference's own tests (test/test_utils.jl:480-494,
  test/test_unified_api.jl:67-83, 164-175, test/test_atoms_base.jl:13-34,
  71-120, test/test_sortbased.jl:118-228, test/test_gpu.jl:78-86);
* closed forms (FCC coordination shells; cell index = floor(frac n + 1));
* a brute force over every (i, j) and every periodic image in reach;
* the legacy linked-list PairList (cell_list.jl:185-505), a second
  algorithm, on atoms inside the cell;
* scipy's cKDTree on orthorhombic cells with cutoff < L/2;
* the committed golden fixtures (tests/golden/*.npz), bit for bit.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from nlhip import workloads

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))

ALL_PBC = [(True, True, True), (False, False, False), (True, False, False), (False, True, False),
           (False, False, True), (True, True, False), (True, False, True), (False, True, True)]
TRICLINIC = np.array([[10.0, 2.0, 1.0], [0.0, 9.0, 1.5], [0.0, 0.0, 8.0]])   # test_utils.jl:297-299


def _set(i, j, S):
    return set(zip(np.asarray(i).tolist(), np.asarray(j).tolist(), map(tuple, np.asarray(S).tolist())))


def _inside(X, cell):
    f = X @ np.linalg.inv(cell)
    return bool(np.all((f >= 0) & (f < 1)))


def _three_way(X, rc, C, pbc):
    """Sort-based == brute force (== legacy when atoms are inside the cell)."""
    ps = oracle.neighbour_list(X, rc, C, pbc, want_R=True)
    bf = oracle.bruteforce(X, rc, C, pbc)
    got = _set(ps.i, ps.j, ps.S)
    assert got == _set(bf.i, bf.j, bf.S)
    assert ps.npairs == bf.npairs
    if _inside(X, C):
        lg = oracle.legacy_pairlist(X, rc, C, pbc)
        assert got == _set(lg.i, lg.j, lg.S)
        assert np.array_equal(np.diff(lg.first), np.diff(ps.first))
    return ps


# ---------------------------------------------------------------- golden fixtures

def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_fixture(path):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    name = os.path.basename(path)[:-4]
    f = _load(path)
    X, C, pbc, rc = f["X"], f["cell"], tuple(bool(b) for b in f["pbc"]), float(f["cutoff"])
    ps = oracle.neighbour_list(X, rc, C, pbc, want_R=True)
    for k in ("i", "j", "first", "perm", "cell_id", "cell_offsets"):
        assert np.array_equal(getattr(ps, k), f[k]), k
    assert np.array_equal(ps.S, f["S"])
    assert np.array_equal(ps.R.view(np.int64), f["R"].view(np.int64))
    # the generator's inputs are reproducible and its known answers still hold
    Xg, Cg, pbcg, rcg, exp = mg.cases()[name]
    assert np.array_equal(np.asarray(Xg, dtype=np.float64), X) and np.array_equal(np.asarray(Cg, dtype=np.float64), C)
    mg.check_expectations(ps, exp)
    bf = oracle.bruteforce(X, rc, C, pbc)
    assert _set(ps.i, ps.j, ps.S) == _set(bf.i, bf.j, bf.S)


def test_golden_fixtures_present():
    assert len(GOLDEN) >= 20


# ---------------------------------------------------------------- known answers

def test_edge_cases():
    C = np.eye(3) * 10.0
    assert oracle.neighbour_list(np.array([[5.0, 5, 5]]), 3.0, C).npairs == 0
    ps = oracle.neighbour_list(np.array([[5.0, 5, 5], [5, 5, 6]]), 3.0, C)
    assert ps.npairs == 2 and list(ps.i) == [1, 2] and list(ps.j) == [2, 1]
    assert np.diff(ps.first).max() == 1                                  # max_neighbours
    assert oracle.neighbour_list(np.array([[1.0, 1, 1], [8, 8, 8]]), 3.0, C, (False,) * 3).npairs == 0
    e = oracle.neighbour_list(np.zeros((0, 3)), 3.0, C)
    assert e.npairs == 0 and list(e.first) == [1]
    assert list(e.cell_offsets) == [1] * (e.geometry.ncells_total + 1)   # cell_list.jl:731-734


@pytest.mark.parametrize("rc,z", [(1.0, 0), (2.6, 12), (3.5, 12), (3.7, 18), (5.0, 42), (5.2, 54)])
def test_fcc_coordination_closed_form(rc, z):
    # FCC shells at a/sqrt2, a, a sqrt(3/2), a sqrt2: 12, 6, 24, 12 neighbours
    X, C = workloads.fcc(3.61, (4, 2, 3))
    ps = _three_way(X, rc, C, (True, True, True))
    assert np.all(np.diff(ps.first) == z)
    if ps.npairs:
        d = np.sqrt((ps.R ** 2).sum(1))
        shells = 3.61 * np.array([np.sqrt(0.5), 1.0, np.sqrt(1.5), np.sqrt(2.0)])
        assert np.min(np.abs(d[:, None] - shells[None, :]), axis=1).max() < 1e-9


def test_isolated_h3():
    X = np.array([[0.0, 0, 0], [0, 0, 2], [0, 10, 0]])
    C = np.diag(X.max(0) - X.min(0) + 1.0)          # ext/NeighbourListsAtomsBaseExt.jl:17-31
    ps = _three_way(X, 5.0, C, (False,) * 3)
    assert list(ps.first) == [1, 2, 3, 3]
    assert list(ps.j) == [2, 1]
    assert np.array_equal(ps.R, np.array([[0.0, 0, 2], [0, 0, -2]]))


def test_issue6_translations():
    # test_sortbased.jl:145-169
    L, C = 8.0, np.eye(3) * 8.0

    def multiset(X):
        ps = oracle.neighbour_list(np.array(X), 1.5, C, (True,) * 3, want_R=True)
        return sorted((min(a, b), max(a, b), round(float(np.sqrt((r ** 2).sum())), 10))
                      for a, b, r in zip(ps.i.tolist(), ps.j.tolist(), ps.R))

    ref = multiset([[0.5, 0.5, 0.5], [0.5, 1.5, 0.5]])
    assert len(ref) == 2
    assert multiset([[0.5, 0.5, 0.5], [0.5, -0.5, 0.5]]) == ref
    assert multiset([[0.5, 0.5, 0.5], [0.5, 1.5 - L, 0.5]]) == ref


@pytest.mark.parametrize("pbc", ALL_PBC)
def test_issue6_shift_by_L_is_noop(pbc):
    # test_sortbased.jl:171-200: per-atom shifts by -L, 0, +L on periodic axes
    s = workloads.random_cubic(60, seed=31)
    L = s.cell[0, 0]
    rc = 0.25 * L

    def multiset(X):
        ps = oracle.neighbour_list(X, rc, s.cell, pbc, want_R=True)
        return sorted((min(a, b), max(a, b), round(float(np.sqrt((r ** 2).sum())), 10))
                      for a, b, r in zip(ps.i.tolist(), ps.j.tolist(), ps.R))

    idx = np.arange(1, 61)
    sh = np.stack([(idx % 3) - 1, ((idx + 1) % 3) - 1, ((idx + 2) % 3) - 1], 1).astype(np.float64)
    Xp = s.X + L * sh * np.array(pbc, dtype=np.float64)[None, :]
    assert multiset(Xp) == multiset(s.X)
    # and the shifted input agrees with the brute force
    bf = oracle.bruteforce(Xp, rc, s.cell, pbc)
    ps = oracle.neighbour_list(Xp, rc, s.cell, pbc)
    assert _set(ps.i, ps.j, ps.S) == _set(bf.i, bf.j, bf.S)


def test_issue6_rounding_trap():
    # test_sortbased.jl:202-228
    L = 8.0
    tiny = -5e-17 * L
    assert (tiny / L) - np.floor(tiny / L) == 1.0
    X = np.array([[0.5, L - 0.3, 0.5], [0.5, tiny, 0.5]])
    ps = oracle.neighbour_list(X, 1.5, np.eye(3) * L, (True,) * 3, want_R=True)
    lo, hi = ps.first[0] - 1, ps.first[1] - 1
    js = list(ps.j[lo:hi])
    assert 2 in js
    k = lo + js.index(2)
    assert abs(np.sqrt((ps.R[k] ** 2).sum()) - 0.3) < 1e-9
    assert np.allclose(X[1] - X[0] + np.eye(3) * L @ ps.S[k], ps.R[k], rtol=0, atol=1e-12)


# ---------------------------------------------------------------- sort-based vs legacy vs brute force

@pytest.mark.parametrize("seed", range(1, 9))
def test_random_configs(seed):
    N = 50 + (seed * 37) % 151                                # test_sortbased.jl:17-21: N in 50:200
    s = workloads.random_cubic(N, seed=seed)
    _three_way(s.X, s.cell[0, 0] / 3, s.cell, (True, True, True))


@pytest.mark.parametrize("pbc", ALL_PBC)
def test_all_pbc(pbc):
    s = workloads.random_cubic(50, seed=21)
    _three_way(s.X, 0.33 * s.cell[0, 0], s.cell, pbc)


@pytest.mark.parametrize("pbc", ALL_PBC)
def test_triclinic(pbc):
    s = workloads.random_in_cell(80, TRICLINIC, seed=5)
    _three_way(s.X, 3.0, TRICLINIC, pbc)


def test_elongated_and_large_cutoff():
    C = np.diag([5.0, 5.0, 20.0])
    _three_way(workloads.random_in_cell(80, C, seed=6).X, 3.0, C, (True, True, True))
    s = workloads.random_cubic(30, seed=8)
    ps = _three_way(s.X, 0.6 * s.cell[0, 0], s.cell, (True, True, True))
    assert np.abs(ps.S).max() >= 1


def test_high_density_multi_image():
    X = 5.0 * workloads.uniform(9, 600).reshape(200, 3)
    ps = _three_way(X, 2.0, np.eye(3) * 5.0, (True, True, True))
    # cells far smaller than the cutoff: several images of the same j
    C = np.diag([2.2, 3.0, 2.5])
    X = workloads.uniform(10, 12).reshape(4, 3) @ C
    ps = _three_way(X, 6.0, C, (True, True, True))
    assert ps.geometry.nxyz.min() >= 2 and ps.npairs > 100


def test_int_types_agree():
    s = workloads.random_cubic(100, seed=12)
    a = oracle.neighbour_list(s.X, s.cell[0, 0] / 3, s.cell, ti_bits=32)
    b = oracle.neighbour_list(s.X, s.cell[0, 0] / 3, s.cell, ti_bits=64)
    for k in ("i", "j", "S", "first", "perm", "cell_id", "cell_offsets"):
        assert np.array_equal(getattr(a, k), getattr(b, k))


@pytest.mark.parametrize("N", [500, 1000, 2000])
def test_counts_vs_legacy(N):
    # test_sortbased.jl:65-71
    s = workloads.random_cubic(N, seed=N)
    rc = s.cell[0, 0] / 4
    ps = oracle.neighbour_list(s.X, rc, s.cell)
    lg = oracle.legacy_pairlist(s.X, rc, s.cell)
    assert ps.npairs == lg.npairs
    assert np.array_equal(np.diff(ps.first), np.diff(lg.first))


def test_hard_cells_vs_bruteforce():
    # test_julip.jl:80-98: left-handed Si cell (negative determinant)
    X = np.array([[0.0, 0.0, 0.0], [1.92333044, 6.63816518e-17, -1.36], [1.92333044, 1.92333044, -2.72],
                  [3.84666089, 1.92333044, -4.08]])
    C = np.diag([3.84666089, 3.84666089, -5.44])
    ps = _three_way(X, 5.405, C, (True, True, True))
    # diamond-like Si at this cutoff: the same count around every atom
    assert len(set(np.diff(ps.first).tolist())) == 1
    # near-degenerate Ti cell with 1e-40 off-diagonals (test_julip.jl:126-129)
    C1 = np.array([[5.71757, -1.81834e-15, 9.74255e-41], [-2.85879, 4.95156, 4.93924e-25],
                   [4.56368e-40, 9.05692e-25, 9.05629]])
    X1 = workloads.uniform(77, 48).reshape(16, 3) @ C1
    for pbc in ALL_PBC:
        _three_way(X1, 7.25, C1, pbc)


def test_scipy_ckdtree():
    # orthorhombic, cutoff < L/2: one image per (i, j), so the unordered pair
    # set is what cKDTree.query_pairs returns
    from scipy.spatial import cKDTree

    for pbc, seed in (((True, True, True), 3), ((False, False, False), 4)):
        s = workloads.random_cubic(3000, density=0.05, seed=seed, cutoff=5.0)
        L = s.cell[0, 0]
        tree = cKDTree(s.X, boxsize=L if all(pbc) else None)
        ref = {(a + 1, b + 1) for a, b in tree.query_pairs(5.0 - 1e-12)} | \
              {(b + 1, a + 1) for a, b in tree.query_pairs(5.0 - 1e-12)}
        ps = oracle.neighbour_list(s.X, 5.0, s.cell, pbc, want_R=True)
        got = set(zip(ps.i.tolist(), ps.j.tolist()))
        # pairs within 1e-12 of the cutoff are the only ones the two may
        # classify differently
        d = np.sqrt((ps.R ** 2).sum(1))
        assert np.all(d < 5.0)
        edge = {p for p, dd in zip(zip(ps.i.tolist(), ps.j.tolist()), d) if dd > 5.0 - 1e-9}
        assert ref - got == set() and got - ref <= edge
        assert ps.npairs == len(got)


# ---------------------------------------------------------------- cell list invariants

@pytest.mark.parametrize("pbc", [(True, True, True), (False, True, False)])
def test_cell_list_invariants(pbc):
    s = workloads.random_in_cell(2000, np.diag([30.0, 25.0, 40.0]) + 2.0 * np.triu(np.ones((3, 3)), 1), seed=91)
    Xs = s.X.copy()
    Xs[::7] += 40.0            # atoms outside the cell (wrapped or clamped)
    ps = oracle.neighbour_list(Xs, 4.0, s.cell, pbc)
    n = ps.X.shape[0]
    assert np.array_equal(np.sort(ps.perm), np.arange(1, n + 1))                 # permutation
    assert np.all(np.diff(ps.cell_id) >= 0)                                      # sorted
    M = ps.geometry.ncells_total
    hist = np.bincount(ps.cell_id, minlength=M + 1)[1:]
    assert ps.cell_offsets[0] == 1 and np.array_equal(np.diff(ps.cell_offsets), hist)
    ids, _ = oracle.cell_ids(Xs, 4.0, s.cell, pbc)
    assert np.array_equal(ids[ps.perm - 1], ps.cell_id)
    # stable: within a cell the original indices ascend
    same = np.diff(ps.cell_id) == 0
    assert np.all(np.diff(ps.perm)[same] > 0)


def test_cell_index_closed_form():
    # c_k = floor(frac_k n_k + 1), wrapped into [1, n] with a winding on
    # periodic axes, clamped on open ones (cell_list.jl:33-74)
    C = np.diag([10.0, 10.0, 10.0])
    pts = np.array([[0.0, 0.0, 0.0], [9.999, 5.0, 2.5], [10.0, 0.0, 0.0], [-0.001, 0.0, 0.0],
                    [25.0, -15.0, 3.3], [0.0, -5e-16, 0.0]])
    ids, w = oracle.cell_ids(pts, 2.5, C, (True, True, True))    # n = 4 per axis
    n = 4
    c = np.floor(pts / 10.0 * n + 1.0).astype(np.int64)
    wrapped = (c - 1) % n + 1
    assert np.array_equal(w, (c - wrapped) // n)
    assert np.array_equal(ids, wrapped[:, 0] + (wrapped[:, 1] - 1) * n + (wrapped[:, 2] - 1) * n * n)
    ids_o, w_o = oracle.cell_ids(pts, 2.5, C, (False, False, False))
    cl = np.clip(c, 1, n)
    assert np.all(w_o == 0)
    assert np.array_equal(ids_o, cl[:, 0] + (cl[:, 1] - 1) * n + (cl[:, 2] - 1) * n * n)


def test_geometry_closed_form():
    C = np.array([[10.0, 2.0, 1.0], [0.0, 9.0, 1.5], [0.0, 0.0, 8.0]])
    g = oracle.analyze(C, 3.0, (True, True, True))
    assert np.allclose(g.inv_cell, np.linalg.inv(C), rtol=1e-15, atol=1e-15)
    V = abs(np.linalg.det(C))
    h = np.array([V / np.linalg.norm(np.cross(C[1], C[2])), V / np.linalg.norm(np.cross(C[2], C[0])),
                  V / np.linalg.norm(np.cross(C[0], C[1]))])
    assert np.allclose(g.lens, h, rtol=1e-14)
    assert list(g.ncells) == [max(int(np.floor(x / 3.0)), 1) for x in h]
    assert list(g.nxyz) == [int(np.ceil(3.0 * n / x)) for n, x in zip(g.ncells, h)]
    assert g.cutoff_sq == 9.0
    # a cutoff beyond the cell: one cell per axis, reach of several cells
    g = oracle.analyze(np.eye(3) * 2.0, 5.0)
    assert list(g.ncells) == [1, 1, 1] and list(g.nxyz) == [3, 3, 3]


def test_index_overflow():
    # cell_list.jl:655-659: prod(ncells) must fit the index type
    with pytest.raises(OverflowError):
        oracle.analyze(np.eye(3) * 1e4, 0.01, ti_bits=32)
    g = oracle.analyze(np.eye(3) * 1e4, 0.01, ti_bits=64)
    assert g.ncells_total == 10 ** 18


def test_planted_mistake_is_caught():
    """A wrong cutoff comparison (<= instead of <) must fail the brute-force check.

    The mistake is planted in a copy of the oracle's answer on the CPU: an
    extra pair at exactly |R| = rc."""
    X = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    ps = oracle.neighbour_list(X, 3.0, np.eye(3) * 10.0, (False,) * 3)
    assert ps.npairs == 0                           # strict: r^2 < rc^2
    bf = oracle.bruteforce(X, 3.0, np.eye(3) * 10.0, (False,) * 3)
    planted = _set([1, 2], [2, 1], [[0, 0, 0], [0, 0, 0]])
    assert planted != _set(bf.i, bf.j, bf.S)
