"""nlh_neighbour_list_noalloc, the build that allocates no device memory,
against nlh_neighbour_list (the entry point it varies), bit for bit: i, j,
S, first and R, on calls in a row of decreasing size, after slot rows grow,
after the entries widen, with Int64 indices, with too small outputs, with
no atoms and beyond the static scratch.
"""
import ctypes as C

import numpy as np
import pytest

import nlhip
from nlhip import _lib, workloads

pytestmark = pytest.mark.gpu


def _noalloc(X, cell, rc, pbc, want_R=True, ti=4, cap=None):
    """One call through the C ABI; returns (status, npairs, arrays)."""
    L = _lib.lib()
    dt = np.int32 if ti == 4 else np.int64
    nat = X.shape[0]
    g = _lib.geometry(cell, rc, pbc, ti)
    Xd = _lib.DeviceArray(np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3))
    c1 = max(int(cap), 1)
    i = _lib.DeviceArray(np.zeros(c1, dt))
    j = _lib.DeviceArray(np.zeros(c1, dt))
    S = _lib.DeviceArray(np.zeros((c1, 3), dt))
    first = _lib.DeviceArray(np.zeros(nat + 1, dt))
    R = _lib.DeviceArray(np.zeros((c1, 3))) if want_R else None
    n = C.c_int64(-1)
    flags = _lib.NLH_X_DEVICE | (_lib.NLH_WANT_R if want_R else 0)
    st = L.nlh_neighbour_list_noalloc(None, Xd.ptr, nat, C.byref(g), flags, int(cap), i.ptr, j.ptr, S.ptr,
                                      first.ptr, R.ptr if R is not None else None, C.byref(n))
    P = int(n.value)
    out = None
    if st == _lib.NLH_OK:
        out = {"i": i.download()[:P], "j": j.download()[:P], "S": S.download()[:P], "first": first.download(),
               "R": R.download()[:P] if R is not None else None}
    for a in (Xd, i, j, S, first, R):
        if a is not None:
            a.free()
    return st, P, out


def _same_as_handle_path(X, cell, rc, pbc, want_R=True, ti=4, slack=0):
    it = np.int32 if ti == 4 else np.int64
    ref = nlhip.neighbour_list(X, rc, cell, pbc, want_R=want_R, int_type=it)
    P = ref.npairs()
    st, n, out = _noalloc(X, cell, rc, pbc, want_R, ti, cap=P + slack)
    assert st == _lib.NLH_OK, _lib.lib().nlh_last_error()
    assert n == P
    for k in ("i", "j", "S", "first"):
        assert np.array_equal(out[k], getattr(ref, k)), k
    if want_R:
        assert np.array_equal(out["R"].view(np.int64), np.asarray(ref.R).view(np.int64))
    return P


def test_two_calls_in_a_row_second_smaller():
    s = workloads.random_cubic(40_000, density=0.05, seed=171)
    print("no-alloc: 40k cube, with R", flush=True)
    P1 = _same_as_handle_path(s.X, s.cell, 5.0, s.pbc, slack=1000)
    C2 = np.array([[40.0, 6.0, 3.0], [0.0, 38.0, 5.0], [0.0, 0.0, 36.0]])
    s2 = workloads.random_in_cell(6000, C2, seed=172)
    print("no-alloc: 6k triclinic, pbc (T, F, T), with R", flush=True)
    P2 = _same_as_handle_path(s2.X, C2, 4.0, (True, False, True))
    assert 0 < P2 < P1
    print("no-alloc: 100k cube, no R", flush=True)
    s3 = workloads.random_cubic(100_000, density=0.05, seed=1)
    _same_as_handle_path(s3.X, s3.cell, 5.0, s3.pbc, want_R=False)
    print("no-alloc: 3k cube, Int64, no R", flush=True)
    s4 = workloads.random_cubic(3000, density=0.05, seed=173)
    _same_as_handle_path(s4.X, s4.cell, 5.0, (False, True, True), want_R=False, ti=8)


def test_rows_grow_and_entries_widen():
    # 80% of the atoms in one eighth of the box: rows sized from the mean
    # density are too short, the count runs again on longer rows
    rng = np.random.default_rng(174)
    L, n = 60.0, 20000
    X = np.concatenate([rng.random((16000, 3)) * (L / 2), rng.random((n - 16000, 3)) * L])
    print("no-alloc: clustered atoms (slot rows grow)", flush=True)
    _same_as_handle_path(X, np.eye(3) * L, 2.74, (True, True, True))
    # windings beyond +-511 cells: 16-byte slot entries
    s = workloads.random_cubic(3000, seed=175)
    Xf = s.X.copy()
    Xf[::5, 0] += 1000.0 * s.cell[0, 0]
    print("no-alloc: far windings (wide entries)", flush=True)
    _same_as_handle_path(Xf, s.cell, s.cutoff, s.pbc)


def test_errors_and_edges():
    L = _lib.lib()
    s = workloads.random_cubic(5000, seed=176)
    ref = nlhip.neighbour_list(s.X, 5.0, s.cell, s.pbc)
    # outputs too small: NLH_ENOMEM, the pair count reported, then a retry
    st, n, _ = _noalloc(s.X, s.cell, 5.0, s.pbc, want_R=False, cap=ref.npairs() // 2)
    assert st == _lib.NLH_ENOMEM and n == ref.npairs()
    assert b"pair outputs" in L.nlh_last_error()
    st, n, out = _noalloc(s.X, s.cell, 5.0, s.pbc, want_R=False, cap=n)
    assert st == _lib.NLH_OK and np.array_equal(out["j"], ref.j) and np.array_equal(out["first"], ref.first)
    # no atoms: first = [1]
    st, n, out = _noalloc(np.zeros((0, 3)), s.cell, 5.0, s.pbc, want_R=False, cap=0)
    assert st == _lib.NLH_OK and n == 0 and list(out["first"]) == [1]
    # a non-finite position: NLH_EINVAL once the binning has seen it
    Xn = s.X.copy()
    Xn[4321, 0] = np.nan
    st, n, _ = _noalloc(Xn, s.cell, 5.0, s.pbc, want_R=False, cap=ref.npairs())
    assert st == _lib.NLH_EINVAL and b"finite" in L.nlh_last_error()
    # beyond the static scratch: a clear error before any launch
    big = workloads.random_cubic(1_000_000, seed=177)
    st, n, _ = _noalloc(big.X, big.cell, 5.0, big.pbc, want_R=False, cap=16)
    assert st == _lib.NLH_ENOMEM and b"scratch" in L.nlh_last_error()
    # and the scratch still serves the next call
    print("no-alloc: after the refusals", flush=True)
    _same_as_handle_path(s.X, s.cell, 5.0, s.pbc)


@pytest.mark.parametrize("seed", range(16))
def test_random_geometries(seed):
    """The random geometries of the parity sweep through the entry point
    that allocates nothing, one call after another on the shared scratch,
    against the oracle."""
    import oracle
    from test_gpu_sweep import _random_system

    s, int_type, want_R = _random_system(seed, 20000, 2.0)
    ti = 4 if int_type == np.int32 else 8
    print(f"no-alloc sweep seed={seed} N={s.X.shape[0]} pbc={s.pbc} ti={ti} R={want_R}", flush=True)
    ref = oracle.neighbour_list(s.X, s.cutoff, s.cell, s.pbc, ti_bits=8 * ti, want_R=want_R, nthreads=8)
    st, n, out = _noalloc(s.X, s.cell, s.cutoff, s.pbc, want_R=want_R, ti=ti, cap=ref.npairs)
    assert st == _lib.NLH_OK, _lib.lib().nlh_last_error()
    assert n == ref.npairs
    for k in ("i", "j", "S", "first"):
        assert np.array_equal(out[k], getattr(ref, k)), k
    if want_R:
        assert np.array_equal(out["R"].view(np.int64), ref.R.view(np.int64))
