"""Serving loop: continuous batching, HBM-budget admission, resident rebuilds.

The identity test serves a request stream mixed (several workers, batched
small systems, resident large systems rebuilt through graph replays) and
each request alone through the plain single-system call, and requires the
two pair lists to be equal bit for bit (i, j, S, first); a few small
requests are also checked against the oracle.
"""
import numpy as np
import pytest

import oracle
from nlhip import serve, workloads

pytestmark = pytest.mark.gpu


def _stream(n, seed, nmax):
    return workloads.served_stream(n, rate=1e4, seed=seed, nmin=16, nmax=nmax)


def _same(a, b):
    assert np.array_equal(a.i, b.i) and np.array_equal(a.j, b.j)
    assert np.array_equal(a.S, b.S) and np.array_equal(a.first, b.first)


@pytest.mark.parametrize("max_residents", [32, 2])
def test_alone_vs_mixed(max_residents):
    # max_residents=2: new large systems take over evicted residents' buffers
    reqs = _stream(48, seed=21, nmax=30000)
    assert any(r.rebuild for r in reqs) and any(r.nat >= 2048 for r in reqs)
    srv = serve.Server(workers=3, small_atoms=2048, batch_atoms=20000, keep_results=True,
                       max_residents=max_residents)
    mixed = srv.serve(reqs, realtime=False)
    if max_residents == 2:
        # residents in use by the other workers are not recycled
        assert srv.evictions > 0 and len(srv.residents) <= 2 + 3 - 1
    srv.close()
    alone = serve.serve_plain(reqs, realtime=False, keep_results=True)
    assert [c.rid for c in mixed] == [r.rid for r in reqs]
    assert max(c.batch for c in mixed) > 1                   # small systems were batched
    for m, a, r in zip(mixed, alone, reqs):
        print(f"request {r.rid} nat={r.nat} rebuild={r.rebuild}", flush=True)
        assert m.npairs == a.npairs
        _same(m.result, a.result)
    for r, m in list(zip(reqs, mixed))[:12]:
        if r.nat <= 3000:
            ref = oracle.neighbour_list(r.X, r.cutoff, r.cell, r.pbc)
            assert np.array_equal(m.result.j, ref.j) and np.array_equal(m.result.S, ref.S)


def test_single_request_units_match_alone():
    # batch_atoms=1: every small request is a unit of its own (the
    # single-system call on the worker's handle), large ones are residents
    reqs = _stream(24, seed=25, nmax=8000)
    srv = serve.Server(workers=2, small_atoms=4096, batch_atoms=1, keep_results=True)
    mixed = srv.serve(reqs, realtime=False)
    srv.close()
    assert max(c.batch for c in mixed) == 1
    alone = serve.serve_plain(reqs, realtime=False, keep_results=True)
    for m, a in zip(mixed, alone):
        assert m.npairs == a.npairs
        _same(m.result, a.result)


def test_open_loop_latency_report():
    reqs = workloads.served_stream(60, rate=2000.0, seed=22, nmax=20000)
    srv = serve.Server(workers=2, small_atoms=4096)
    done = srv.serve(reqs, realtime=True)
    srv.close()
    s = serve.summarize(done)
    assert s["requests"] == 60 and s["p99_ms"] >= s["p50_ms"] > 0
    assert all(c.end >= c.start >= c.arrival - 1e-3 for c in done)


def test_admission_under_a_small_budget():
    # budget for a few systems at a time: everything still completes, the
    # admitted estimate never exceeds the budget, residents get evicted
    reqs = _stream(30, seed=23, nmax=40000)
    budget = max(serve.estimate_bytes(r.nat, r.cell, r.cutoff) for r in reqs) * 2
    srv = serve.Server(hbm_budget=budget, workers=3, small_atoms=1024, keep_results=True)
    done = srv.serve(reqs, realtime=False)
    assert len(done) == 30 and srv.peak_bytes <= budget
    srv.close()
    alone = serve.serve_plain(reqs, realtime=False, keep_results=True)
    for m, a in zip(done, alone):
        _same(m.result, a.result)


def test_request_larger_than_budget_is_refused():
    reqs = _stream(3, seed=24, nmax=2000)
    srv = serve.Server(hbm_budget=1 << 20)
    with pytest.raises(MemoryError):
        srv.serve(reqs, realtime=False)


def test_resident_requests_with_new_geometry():
    """Requests for a resident system that change its cell (strain), its
    cutoff or its atom count are built with their own geometry, not the
    resident's first one: every result equals the plain call and the
    oracle."""
    from nlhip.serve import Request

    base = workloads.random_cubic(6000, density=0.05, seed=141)
    rng = np.random.default_rng(142)
    X0, C0 = base.X, base.cell
    reqs = [Request(0, 0, X0, C0, (True, True, True), 5.0, False, 0.0)]
    X1 = X0 + 0.1 * rng.standard_normal(X0.shape)
    reqs.append(Request(1, 0, X1, C0, (True, True, True), 5.0, True, 0.0))           # plain rebuild
    strain = np.array([[1.03, 0.02, 0.0], [0.0, 0.98, 0.01], [0.0, 0.0, 1.01]])
    reqs.append(Request(2, 0, X1 @ strain, C0 @ strain, (True, True, True), 5.0, True, 0.0))   # new cell
    reqs.append(Request(3, 0, X1 @ strain, C0 @ strain, (True, False, True), 4.2, True, 0.0))  # pbc, cutoff
    reqs.append(Request(4, 0, X1[:5500], C0, (True, True, True), 5.0, True, 0.0))   # fewer atoms
    reqs.append(Request(5, 0, X1[:5500] + 0.05, C0, (True, True, True), 5.0, True, 0.0))
    srv = serve.Server(workers=2, small_atoms=2048, keep_results=True)
    mixed = srv.serve(reqs, realtime=False)
    srv.close()
    alone = serve.serve_plain(reqs, realtime=False, keep_results=True)
    for m, a, r in zip(mixed, alone, reqs):
        print(f"request {r.rid} nat={r.nat} rc={r.cutoff}", flush=True)
        ref = oracle.neighbour_list(r.X, r.cutoff, r.cell, r.pbc, want_R=False)
        assert m.npairs == a.npairs == ref.npairs
        _same(m.result, a.result)
        assert np.array_equal(m.result.j, ref.j) and np.array_equal(m.result.S, ref.S)
        assert np.array_equal(m.result.first, ref.first)


def test_served_latency_covers_the_fill():
    """A completion's latency runs to the complete pair list; the call-return
    time is reported beside it and never later."""
    reqs = _stream(40, seed=27, nmax=120000)
    srv = serve.Server(workers=2, small_atoms=8192)
    done = srv.serve(reqs, realtime=False)
    srv.close()
    assert all(c.end >= c.returned >= c.start for c in done)
    s = serve.summarize(done)
    assert s["p50_ms"] >= s["p50_return_ms"] and s["p99_ms"] >= s["p99_return_ms"]


def test_random_geometry_stream():
    """Requests drawn by the parity sweep's generator (random triclinic cells,
    pbc patterns, cutoffs up to 2.5 face heights, atoms outside the cell),
    a third of them rebuilds of an earlier system with new positions, served
    mixed (batched small systems, residents for the large ones); every
    completion is checked against the oracle."""
    from test_gpu_sweep import _random_system

    u = workloads.uniform(4242, 48)
    reqs, base = [], []
    for k in range(48):
        if base and u[k] < 0.33:
            sid = int(workloads.uniform(4343 + k, 1)[0] * len(base))
            b = base[sid]
            X = b.X + 0.05 * workloads.normal(4444 + k, b.X.size).reshape(b.X.shape)
            reqs.append(serve.Request(k, sid, X, b.cell, b.pbc, b.cutoff, True))
            continue
        s, _, _ = _random_system(300 + k, 20000, 2.0)
        reqs.append(serve.Request(k, len(base), s.X, s.cell, s.pbc, s.cutoff, False))
        base.append(s)
    assert any(r.rebuild for r in reqs) and any(r.nat >= 2048 for r in reqs)
    srv = serve.Server(workers=3, small_atoms=2048, batch_atoms=20000, keep_results=True)
    done = srv.serve(reqs, realtime=False)
    srv.close()
    assert [c.rid for c in done] == [r.rid for r in reqs]
    for c, r in zip(done, reqs):
        print(f"request {r.rid} nat={r.nat} pbc={r.pbc} rebuild={r.rebuild}", flush=True)
        ref = oracle.neighbour_list(r.X, r.cutoff, r.cell, r.pbc, want_R=False, nthreads=8)
        assert c.npairs == ref.npairs
        _same(c.result, ref)
