"""Batched builds: graph-captured rebuilds and launch configurations.

The reference has no batching API, so every system of a batch is checked
against the single-system oracle (bit for bit: i, j, S, first, R).  A
graph replay is checked on every replay against the stage-by-stage
(eager) rebuild of the same positions, with eager work and synchronisations
between the replays.
"""
import numpy as np
import pytest

import nlhip
import oracle
from nlhip import _lib, workloads
from nlhip.batch import BatchBuilder, pack_systems

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert np.array_equal(a.i, b.i) and np.array_equal(a.j, b.j) and np.array_equal(a.S, b.S)
    assert np.array_equal(a.first, b.first)
    if a.R is not None and b.R is not None:
        assert np.array_equal(np.asarray(a.R).view(np.int64), np.asarray(b.R).view(np.int64))


def _jitter(X, seed, sigma):
    return X + sigma * workloads.normal(seed, X.size).reshape(X.shape)


def _check_oracle(res, systems, Xcat, offs, which, R=True):
    for s in which:
        sysm = systems[s]
        Xs = Xcat[offs[s]:offs[s + 1]]
        ref = oracle.neighbour_list(Xs, sysm.cutoff, sysm.cell, sysm.pbc, want_R=R)
        pl = res.pairlist(s)
        assert pl.npairs() == ref.npairs
        assert np.array_equal(pl.i, ref.i) and np.array_equal(pl.j, ref.j) and np.array_equal(pl.S, ref.S)
        assert np.array_equal(pl.first, ref.first)
        if R:
            assert np.array_equal(np.asarray(pl.R).view(np.int64), ref.R.view(np.int64))


def test_graph_replays_match_eager():
    systems = workloads.batch_systems(96, seed=5, nmax=600)
    X0, offs = pack_systems(systems)
    bb = BatchBuilder()
    bb.run(systems, want_R=True)
    ref_bb = BatchBuilder()
    ref_bb.run(systems, want_R=True)
    for rep in range(4):
        X = _jitter(X0, 100 + rep, 0.1)
        print(f"graph replay {rep}", flush=True)
        g = bb.rebuild(X, want_R=True, graph=True)
        e = ref_bb.rebuild(X, want_R=True, graph=False)          # eager work between replays
        assert g.npairs == e.npairs
        _same(g, e)
        assert np.array_equal(g.pair_offsets, e.pair_offsets)
        _check_oracle(g, systems, X, offs, [0, 7, 31, 95])


def test_graph_replay_with_device_positions():
    # built without R, rebuilt with R: the replay must size the R slots and
    # outputs before capturing (regression: they were left unallocated)
    systems = workloads.batch_systems(32, seed=6, nmax=500)
    X0, offs = pack_systems(systems)
    Xd = _lib.DeviceArray(X0)
    bb = BatchBuilder()
    bb.run(systems, X_device=Xd, copy=False)
    for rep in range(3):
        X = _jitter(X0, 200 + rep, 0.05)
        Xd.upload(X)
        g = bb.rebuild(X, X_device=Xd, want_R=True)
        _check_oracle(g, systems, X, offs, range(32))
    Xd.free()


def test_graph_replay_outgrows_capacity():
    # compress every system: more pairs than the captured buffers hold, so
    # the replay's fill skips itself and the stage path regrows the outputs
    systems = workloads.batch_systems(24, seed=8, nmax=300)
    X0, offs = pack_systems(systems)
    bb = BatchBuilder()
    bb.run(systems)
    bb.rebuild(X0)                                          # captured here
    dense = []
    for s in systems:
        c = s.X.mean(0)
        dense.append(workloads.System(c + 0.4 * (s.X - c), s.cell, s.pbc, s.cutoff))
    Xd_, _ = pack_systems(dense)
    g = bb.rebuild(Xd_, want_R=True)
    _check_oracle(g, dense, Xd_, offs, range(24))
    g2 = bb.rebuild(Xd_, want_R=True)                       # re-captured on the grown buffers
    _same(g, g2)
    plain = BatchBuilder()                                  # the same without R
    plain.run(systems)
    plain.rebuild(X0)
    _check_oracle(plain.rebuild(Xd_), dense, Xd_, offs, range(24), R=False)


@pytest.mark.parametrize("brick_atoms,fill_block,fill_pairs,count_block,cell_sort",
                         [(64, 64, 128, 0, 1), (512, 128, 512, 256, 0), (1024, 512, 256, 512, 1),
                          (2048, 1024, 128, 256, 0), (1536, 256, 512, 512, 1), (768, 256, 256, 256, 0)])
def test_tuning_does_not_change_results(brick_atoms, fill_block, fill_pairs, count_block, cell_sort):
    s = workloads.random_cubic(6000, density=0.05, seed=81)
    h = _lib.Handle()
    h.set_tuning(brick_atoms, fill_block, fill_pairs, count_block, cell_sort)
    t = h.get_tuning()
    assert (t.brick_atoms, t.fill_block, t.fill_pairs, t.count_block, t.cell_sort) == (
        brick_atoms, fill_block, fill_pairs, count_block, cell_sort)
    nl = nlhip.neighbour_list(s.X, 5.0, s.cell, (True, False, True), want_R=True, handle=h)
    ref = oracle.neighbour_list(s.X, 5.0, s.cell, (True, False, True), want_R=True)
    _same(nl, ref)


def test_tuning_rejects_bad_values():
    h = _lib.Handle()
    with pytest.raises(_lib.NLHInvalidError):
        h.set_tuning(16, 256)
    with pytest.raises(_lib.NLHInvalidError):
        h.set_tuning(1024, 100)
    with pytest.raises(_lib.NLHInvalidError):
        h.set_tuning(1024, 256, 200)
    with pytest.raises(_lib.NLHInvalidError):
        h.set_tuning(1024, 256, 256, 128)
    with pytest.raises(_lib.NLHInvalidError):
        h.set_tuning(1024, 256, 256, 256, 2)
    with pytest.raises(_lib.NLHInvalidError):
        h.set_tuning(1024, 256, 256, 256, 0, 2)


def test_autotune_keeps_results():
    systems = workloads.batch_systems(128, seed=9, nmax=800)
    X0, offs = pack_systems(systems)
    bb = BatchBuilder()
    before = bb.run(systems, want_R=True)
    best = bb.autotune(X0, reps=1)
    assert best.brick_atoms in (460, 920, 1840) and best.fill_block in (128, 256, 512)
    assert best.count_block in (256, 512)
    assert best.fill_pairs in (256, 512)
    t = bb.handle.get_tuning()
    assert (t.brick_atoms, t.fill_block, t.fill_pairs, t.count_block) == (best.brick_atoms, best.fill_block,
                                                                          best.fill_pairs, best.count_block)
    after = bb.rebuild(X0, want_R=True)
    _same(before, after)
    _check_oracle(after, systems, X0, offs, [0, 1, 64, 127])
    plain = bb.rebuild(X0)                      # the tuned batch without R
    _check_oracle(plain, systems, X0, offs, [0, 1, 64, 127], R=False)


def test_replay_after_entries_widen():
    """A replay whose shifts outgrow the packed slot entries (windings beyond
    +-511 cells) is redone on the stage path with 16-byte entries; the next
    replays recapture the pipeline for the wide entries instead of reusing
    the narrow capture (the regrown slot buffer may come back at the same
    address).  Batched rebuilds and repeated single builds both."""
    systems = [workloads.random_cubic(3000, seed=131), workloads.random_cubic(1500, seed=132)]
    X0, offs = pack_systems(systems)
    bb = BatchBuilder()
    bb.run(systems, want_R=True)
    bb.rebuild(X0, want_R=True)                            # captured (narrow entries)
    far = X0.copy()
    L0 = systems[0].cell[0, 0]
    far[0:3000:7, 0] += 1000.0 * L0                        # system 0: x windings of 1000
    far[1:3000:11, 1] -= 777.0 * L0
    for rep in range(3):
        print(f"widened replay {rep}", flush=True)
        g = bb.rebuild(far, want_R=True)
        _check_oracle(g, systems, far, offs, [0, 1])
    # repeated single builds of one shape replay below 12,288 atoms
    s = systems[0]
    h = _lib.Handle()
    for rep in range(2):
        nl = nlhip.neighbour_list(s.X, s.cutoff, s.cell, s.pbc, want_R=True, handle=h)
        _same(nl, oracle.neighbour_list(s.X, s.cutoff, s.cell, s.pbc, want_R=True))
    Xf = far[:3000]
    for rep in range(3):
        print(f"widened single build {rep}", flush=True)
        nl = nlhip.neighbour_list(Xf, s.cutoff, s.cell, s.pbc, want_R=True, handle=h)
        _same(nl, oracle.neighbour_list(Xf, s.cutoff, s.cell, s.pbc, want_R=True))
    assert h.graph_launches() >= 2
    h.close()


# ---------------------------------------------------------------------------
# fused build of small systems
# ---------------------------------------------------------------------------

def _fused_off(bb):
    t = bb.handle.get_tuning()
    bb.handle.set_tuning(t.brick_atoms, t.fill_block, t.fill_pairs, t.count_block, t.cell_sort, 1)


def _cell_lists(bb, nat, ncell_entries, ti):
    perm = np.empty(nat, ti)
    cid = np.empty(nat, ti)
    coff = np.empty(ncell_entries, ti)
    Xs = np.empty((nat, 3))
    _lib.check(_lib.lib().nlh_copy_celllist(bb.handle.ptr, Xs.ctypes.data, perm.ctypes.data, cid.ctypes.data,
                                            coff.ctypes.data))
    return Xs, perm, cid, coff


@pytest.mark.parametrize("int_type,want_R", [(nlhip.Int32, False), (nlhip.Int32, True), (nlhip.Int64, True)])
def test_fused_small_systems_match_the_general_pipeline(int_type, want_R):
    """A batch of small systems built by the fused kernel (one workgroup per
    unit of systems: binning, sort, cell list and count in LDS) equals the
    general pipeline's build of the same batch in every output (pair list,
    per-system offsets, sorted positions, perm, cell_id, cell_offsets), and
    the oracle on sampled systems."""
    systems = workloads.batch_systems(600, seed=181, nmax=2048)
    X, offs = pack_systems(systems)
    fused = BatchBuilder(int_type)
    plain = BatchBuilder(int_type)
    _fused_off(plain)
    print(f"fused batch {int_type} R={want_R}", flush=True)
    a = fused.run(systems, want_R=want_R)
    b = plain.run(systems, want_R=want_R)
    assert a.npairs == b.npairs
    _same(a, b)
    assert np.array_equal(a.pair_offsets, b.pair_offsets)
    ti = fused.ti
    dt = np.int32 if ti == 4 else np.int64
    M = sum(int(_lib.geometry(s.cell, s.cutoff, s.pbc, ti).ncells_total) for s in systems)
    ca = _cell_lists(fused, X.shape[0], M + len(systems), dt)
    cb = _cell_lists(plain, X.shape[0], M + len(systems), dt)
    for u, v in zip(ca, cb):
        assert np.array_equal(u, v)
    _check_oracle(a, systems, X, offs, [0, 1, 299, 599], R=want_R)


def test_fused_small_windings_and_fallback():
    """Atoms outside their cell (windings of one image: 8-byte slot rows) stay
    on the fused path; a winding beyond +-511 cells makes the build fall back
    to the general pipeline, with the same results; repeated rebuilds."""
    systems = workloads.batch_systems(200, seed=182, nmax=1500)
    moved = []
    for k, s in enumerate(systems):
        Xs = s.X.copy()
        if k % 3 == 0 and s.nat > 1:
            Xs[::2] += s.cell[0] * (1 if s.pbc[0] else 0) - s.cell[1] * (1 if s.pbc[1] else 0)
        moved.append(workloads.System(Xs, s.cell, s.pbc, s.cutoff))
    X1, offs = pack_systems(moved)
    bb = BatchBuilder()
    print("fused build with windings", flush=True)
    res = bb.run(moved, want_R=True)
    _check_oracle(res, moved, X1, offs, range(0, 200, 7))
    for rep in range(2):
        print(f"fused rebuild {rep}", flush=True)
        g = bb.rebuild(X1, want_R=True)
        _same(g, res)
    far = X1.copy()
    per = [k for k, s in enumerate(systems) if s.pbc[0] and s.nat > 1][:2]
    assert len(per) == 2
    a, b = per
    far[offs[a]:offs[a + 1]] += 1000.0 * systems[a].cell[0]   # every atom 1000 cells along a_1
    far[offs[b]] += 700.0 * systems[b].cell[0]                # one atom 700 cells along a_1
    print("fused build with far windings", flush=True)
    g = bb.rebuild(far, want_R=True)
    _check_oracle(g, moved, far, offs, [a, b, 7, 100])
    g2 = bb.rebuild(far, want_R=True)
    _same(g, g2)
    plain = BatchBuilder()
    _fused_off(plain)
    plain.run(moved, want_R=True)
    _same(plain.rebuild(far, want_R=True), g)


def test_fused_small_tiny_cells_and_empty_systems():
    """Systems of 1-4 atoms in cells smaller than the cutoff (reach 2-3
    cells: up to 343 images, several segment batches per home cell), and
    systems with no atoms, through the fused path."""
    rng = np.random.default_rng(183)
    systems = []
    for k in range(150):
        if k % 10 == 0:
            systems.append(workloads.System(np.zeros((0, 3)), np.eye(3) * 9.0, (True, True, True), 4.0))
            continue
        side = 2.2 + 2.0 * rng.random()
        C = np.diag([side, side * (1 + 0.1 * rng.random()), side])
        n = 1 + k % 4
        systems.append(workloads.System(rng.random((n, 3)) @ C, C, (bool(k % 2), True, bool(k % 3)),
                                        4.0 + 2.0 * rng.random()))
    X, offs = pack_systems(systems)
    bb = BatchBuilder()
    print("fused tiny cells", flush=True)
    res = bb.run(systems, want_R=True)
    _check_oracle(res, systems, X, offs, range(150))
    plain = BatchBuilder()
    _fused_off(plain)
    _same(plain.run(systems, want_R=True), res)


def test_fused_small_rows_and_outputs_regrow():
    """Clumped systems: every atom sees all others, far above the density
    estimate the slot rows and pair outputs are first sized for, so the fused
    build regrows both and counts again; equal to the general pipeline and
    the oracle.  A rebuild with spread-out positions then fits the grown
    buffers."""
    rng = np.random.default_rng(184)
    systems = []
    for k in range(80):
        n = 40 + (k * 37) % 260
        side = 18.0 + 4.0 * rng.random()
        C = np.diag([side, side, side])
        X = np.array([side / 2] * 3) + 2.5 * (rng.random((n, 3)) - 0.5)
        systems.append(workloads.System(X, C, (bool(k % 2), True, bool(k % 3)), 4.5 + rng.random()))
    X, offs = pack_systems(systems)
    bb = BatchBuilder()
    print("fused clumped build", flush=True)
    res = bb.run(systems, want_R=True)
    _check_oracle(res, systems, X, offs, [0, 1, 17, 40, 79])
    plain = BatchBuilder()
    _fused_off(plain)
    _same(plain.run(systems, want_R=True), res)
    spread = [workloads.System(rng.random((s.nat, 3)) @ s.cell, s.cell, s.pbc, s.cutoff) for s in systems]
    X2, _ = pack_systems(spread)
    print("fused rebuild spread", flush=True)
    g = bb.rebuild(X2, want_R=True)
    _check_oracle(g, spread, X2, offs, [0, 3, 79])


def test_fused_small_unit_limits():
    """Unit edges of the fused build: 300 one-atom systems (a unit holds at
    most 256 systems), a system of exactly 2048 atoms, and a rod of 256
    cells along one axis (the most a system may have); then the same batch
    with a 257-cell rod, which the general pipeline builds.  Every case
    equals the oracle and the general pipeline."""
    rng = np.random.default_rng(185)
    systems = []
    for k in range(300):
        side = 6.0 + 2.0 * rng.random()
        systems.append(workloads.System(rng.random((1, 3)) * side, np.eye(3) * side, (True, bool(k % 2), True), 5.0))
    d = (2048 / 0.06) ** (1.0 / 3.0)
    systems.append(workloads.System(rng.random((2048, 3)) * d, np.eye(3) * d, (True, True, False), 5.0))

    def rod(ncell):
        C = np.diag([5.0 * ncell + 1.0, 6.0, 6.0])
        return workloads.System(rng.random((400, 3)) @ C, C, (True, False, True), 5.0)

    for tail in (rod(256), rod(257)):
        batch = systems + [tail]
        g = _lib.geometry(tail.cell, tail.cutoff, tail.pbc, 4)
        print(f"fused unit limits, rod of {g.ncells[0]} cells", flush=True)
        X, offs = pack_systems(batch)
        bb = BatchBuilder()
        res = bb.run(batch, want_R=True)
        _check_oracle(res, batch, X, offs, [0, 255, 256, 299, 300, 301])
        plain = BatchBuilder()
        _fused_off(plain)
        _same(plain.run(batch, want_R=True), res)
