"""Serving loop: neighbour-list requests with continuous batching on one GPU.

The reference serves nothing: each `neighbour_list` call builds one system
(src/cell_list.jl:897-916).  This loop takes a stream of requests -- new
systems of any size, or new positions for a system served before -- and
keeps the device busy:

* admission: a request enters the active set only while the estimated
  device memory of everything in flight plus the resident systems fits the
  HBM budget; resident systems are evicted least-recently-used first when
  a new request would not fit;
* continuous batching: whenever a worker is free, the scheduler hands it the
  next unit of work -- one large system, or as many admitted small systems
  as fit `batch_atoms` in a single batched build -- so requests that arrive
  while others run join the next batch instead of waiting for a drain;
* overlap: each worker thread owns its own handle (its own HIP stream and
  buffers); the C calls release the GIL, so several builds are on the
  device at once;
* residency: a large system keeps its positions and its cell list on the
  device in a handle of its own; a rebuild request with the same atom
  count, cell, pbc and cutoff uploads the new positions into the resident
  buffer (through the handle's pinned staging) and rebuilds on the
  resident's buffers (nlh_batch_rebuild); a request for the system with a
  new cell (NPT, strain), pbc, cutoff or atom count is a fresh build on the
  resident's handle with its own geometry.  At most `max_residents`
  systems stay resident: a new one then takes over the handle of the least
  recently used idle resident (device buffers only grow, so a steady
  stream allocates nothing; device frees synchronise the whole device, so
  they happen only under HBM-budget pressure).  A resident whose system
  has a request admitted or running is never evicted or recycled.
* latency: a completion records when its build call returned (`returned`:
  the pair count is known, the last kernel may still be writing pairs) and
  when the pair list is complete on the device (`end`, after waiting for
  the worker's stream); `latency` is measured to `end`.

Every result equals the single-system build of the same input bit for bit
(the batch and replay paths emit in the reference traversal order), which
the alone-vs-mixed test checks.
"""
from __future__ import annotations

import ctypes as C
import gc
import math
import queue
import sys
import threading
import time
from collections import OrderedDict
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import Handle, check, geometry, lib
from .api import Int32, PairList
from .batch import BatchBuilder, pack_systems
from .workloads import System


@dataclass
class Request:
    rid: int
    system_id: int
    X: np.ndarray
    cell: np.ndarray
    pbc: tuple
    cutoff: float
    rebuild: bool = False     # new positions for a system served before
    arrival: float = 0.0      # seconds from the start of the stream

    @property
    def nat(self) -> int:
        return int(self.X.shape[0])


@dataclass
class Completion:
    rid: int
    system_id: int
    npairs: int
    arrival: float
    start: float
    end: float                # the pair list is complete on the device
    batch: int                # systems in the unit of work that served it
    result: PairList | None = None
    returned: float = 0.0     # the build call returned (pair count known)

    @property
    def latency(self) -> float:
        """Arrival to a complete pair list on the device."""
        return self.end - self.arrival

    @property
    def return_latency(self) -> float:
        """Arrival to the build call's return (the fill may still run)."""
        return self.returned - self.arrival


def _geometry_key(cell, pbc, cutoff) -> tuple:
    """What a resident's cell list depends on besides the positions."""
    return (np.asarray(cell, dtype=np.float64).tobytes(), tuple(bool(b) for b in pbc), float(cutoff))


def _cross(u, v):
    return (u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0])


def estimate_bytes(nat: int, cell, cutoff: float, ti_bytes: int = 4) -> int:
    """Device memory of one build: positions, cell list, neighbour slots and
    pair outputs (pair count from the mean density; over-estimates by ~25%).
    Scalar arithmetic on the 3x3 cell: numpy's per-call cost on 3-vectors
    would put ~0.1 ms on the scheduler for every admitted request."""
    a = np.asarray(cell, dtype=np.float64).tolist()
    faces = [_cross(a[(k + 1) % 3], a[(k + 2) % 3]) for k in range(3)]   # |face k| x height k = volume
    vol = abs(a[0][0] * faces[0][0] + a[0][1] * faces[0][1] + a[0][2] * faces[0][2])
    rho = nat / vol if vol > 0 else 0.0
    nb = rho * 4.18879020478639 * cutoff ** 3
    M = 1.0
    for f in faces:
        h = vol / max(math.sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]), 1e-300)
        M *= max(math.floor(h / cutoff), 1.0)
    cap = int((nb * 1.5 + 16 + 7) // 8 * 8)
    pairs = nat * (nb * 1.25 + 4)
    per_atom = 24 + 6 * 4 + 32 + 12 + 2 * ti_bytes + 8 + ti_bytes + 8 * cap
    b = nat * per_atom + M * (12 + ti_bytes) + pairs * 5 * ti_bytes
    return int(b * 1.15) + (1 << 20)


class _Resident:
    """A large system kept on the device in a handle of its own: positions
    (the handle's position buffer, written through its pinned staging),
    cell list, pair list and captured pipeline.  `recycled` is the (handle,
    bytes held) of an evicted resident to build on."""

    def __init__(self, req: Request, int_type, recycled=None):
        self.system_id = req.system_id
        self.ti = _lib.ti_bytes_of(int_type)
        self.nat = req.nat
        self.geom = geometry(req.cell, req.cutoff, req.pbc, self.ti)
        self.key = _geometry_key(req.cell, req.pbc, req.cutoff)
        self.bytes = estimate_bytes(req.nat, req.cell, req.cutoff, self.ti)
        self.handle = None
        if recycled is not None:
            self.handle, held = recycled
            self.bytes = max(self.bytes, held)      # its buffers never shrink
        if self.handle is None:
            self.handle = Handle()
        self.lock = threading.Lock()
        self.built = False
        self.rebuilds = 0

    # rebuilds before the pipeline is captured as a graph: capturing and
    # instantiating costs milliseconds, more than a few eager rebuilds; and a
    # resident (at least small_atoms atoms) builds faster with its stages
    # launched directly than replayed (the host queues the next build while
    # the fill runs), so by default it is never captured
    GRAPH_AFTER = 1 << 62

    def build(self, req: Request) -> int:
        """First build, or a rebuild on new positions (upload into the
        resident buffer; stages launched directly for the first GRAPH_AFTER
        rebuilds, replays of the captured pipeline after that).  A request
        whose atom count, cell, pbc or cutoff differ from the resident's is
        a fresh build with its own geometry on the resident's handle."""
        n = C.c_int64()
        X = np.ascontiguousarray(req.X, dtype=np.float64)
        key = _geometry_key(req.cell, req.pbc, req.cutoff)
        if self.built and key == self.key and req.nat == self.nat:
            flags = _lib.NLH_NO_GRAPH if self.rebuilds < self.GRAPH_AFTER else 0
            self.rebuilds += 1
            check(lib().nlh_batch_rebuild(self.handle.ptr, X.ctypes.data, flags, C.byref(n)), "rebuild")
        else:
            if key != self.key or req.nat != self.nat:
                self.geom = geometry(req.cell, req.cutoff, req.pbc, self.ti)
                self.key, self.nat, self.rebuilds = key, req.nat, 0
            check(lib().nlh_neighbour_list(self.handle.ptr, X.ctypes.data, self.nat, C.byref(self.geom), 0,
                                           C.byref(n)), "neighbour_list")
            self.built = True
        return int(n.value)

    def synchronize(self):
        """Wait until the last build's pair list is complete."""
        check(lib().nlh_synchronize(self.handle.ptr), "synchronize")

    def detach(self):
        """Hand the handle over to a new resident."""
        out = (self.handle, self.bytes)
        self.handle = None
        return out

    def result(self, X_host: np.ndarray) -> PairList:
        from .api import _collect_pairs
        v = _lib.PairListView()
        check(lib().nlh_pairlist_device(self.handle.ptr, C.byref(v)))
        return _collect_pairs(self.handle, self.geom, X_host, int(v.npairs), False, None)

    def free(self):
        if self.handle is not None:
            self.handle.close()


class Server:
    """Continuous-batching neighbour-list server on one GPU.

    hbm_budget    bytes of device memory the server may hold (in flight plus
                  resident systems)
    workers       worker threads, each with its own handle (HIP stream and
                  grow-only buffers, reused from one unit of work to the next)
    small_atoms   systems below this size are batched together; larger ones
                  are resident (own handle, positions and cell list kept on
                  the device, rebuilt on their own buffers)
    batch_atoms   atoms per batched build
    max_residents resident systems kept; beyond it a new one recycles the
                  least recently used idle resident's device buffers (a
                  resident with a request admitted or running is not idle,
                  so up to max_residents + workers - 1 can be held)
    keep_results  copy each result back to the host (PairList) on completion
    """

    def __init__(self, hbm_budget: int = 64 << 30, workers: int = 4, small_atoms: int = 65536,
                 batch_atoms: int = 1 << 19, int_type=Int32, keep_results: bool = False, max_residents: int = 32):
        self.budget = int(hbm_budget)
        self.nworkers = int(workers)
        self.small_atoms = int(small_atoms)
        self.batch_atoms = int(batch_atoms)
        self.int_type = int_type
        self.keep = keep_results
        self.max_residents = max(int(max_residents), 1)
        self.residents: OrderedDict[int, _Resident] = OrderedDict()
        self.resident_bytes = 0
        self.inflight_bytes = 0
        self.peak_bytes = 0          # largest in-flight + resident estimate admitted
        self.evictions = 0
        self._lock = threading.Lock()
        self._work: queue.Queue = queue.Queue()
        self._done: queue.Queue = queue.Queue()
        self._free_workers = threading.Semaphore(0)
        self._threads: list[threading.Thread] = []
        self._errors: list[BaseException] = []
        # one batch builder per worker slot, kept across serve() calls: its
        # device buffers only grow, so a warm server allocates nothing
        self._builders: list[BatchBuilder | None] = [None] * self.nworkers
        self._est: dict[int, int] = {}   # id(request) -> estimate_bytes, for one serve() call
        # systems with a request admitted or running (under _lock): their
        # residents are neither evicted nor recycled
        self._busy: set[int] = set()
        self.t0 = 0.0

    # ------------------------------------------------------------------ workers

    def _worker(self, slot: int):
        local = {"slot": slot}
        self._free_workers.release()
        while True:
            item = self._work.get()
            if item is None:
                return
            try:
                done = self._run_item(item, local)
            except BaseException as e:   # reported by serve()
                self._errors.append(e)
                done = [("error", r, 0) for r in item[1]]
            # free the slot first: the scheduler, woken by the completions,
            # then finds this worker available
            self._free_workers.release()
            for d in done:
                self._done.put(d)

    def _run_item(self, item, local) -> list:
        """Run one unit of work; returns its completions."""
        kind, reqs, nbytes = item
        start = time.perf_counter() - self.t0
        if kind == "large":
            r = reqs[0]
            res = self._resident_for(r, nbytes)
            nbytes = 0                       # the reservation became the resident's
            with res.lock:
                P = res.build(r)
                ret = time.perf_counter() - self.t0
                res.synchronize()            # the pair list is complete
                out = res.result(r.X) if self.keep else None
            end = time.perf_counter() - self.t0
            return [("ok", Completion(r.rid, r.system_id, P, r.arrival, start, end, 1, out, ret), nbytes)]
        bb = self._builders[local["slot"]]
        if bb is None:
            bb = self._builders[local["slot"]] = BatchBuilder(self.int_type)
        if len(reqs) == 1:
            # a batch of one: the single-system call on the slot's handle
            # (no packing, system table or per-system offsets to fetch)
            r = reqs[0]
            g = geometry(r.cell, r.cutoff, r.pbc, _lib.ti_bytes_of(self.int_type))
            X = np.ascontiguousarray(r.X, dtype=np.float64)
            n = C.c_int64()
            check(lib().nlh_neighbour_list(bb.handle.ptr, X.ctypes.data, X.shape[0], C.byref(g), 0, C.byref(n)),
                  "neighbour_list")
            ret = time.perf_counter() - self.t0
            bb.handle.synchronize()          # the pair list is complete
            out = None
            if self.keep:
                from .api import _collect_pairs
                out = _collect_pairs(bb.handle, g, X, int(n.value), False, None)
            end = time.perf_counter() - self.t0
            return [("ok", Completion(r.rid, r.system_id, int(n.value), r.arrival, start, end, 1, out, ret),
                     nbytes)]
        systems = [System(r.X, r.cell, r.pbc, r.cutoff) for r in reqs]
        bb.run(systems, copy=False)
        ret = time.perf_counter() - self.t0
        bb.handle.synchronize()              # every pair list of the batch is complete
        po = np.zeros(len(reqs) + 1, dtype=np.int64)
        check(lib().nlh_batch_pair_offsets(bb.handle.ptr, po.ctypes.data_as(C.POINTER(C.c_int64))))
        res = None
        if self.keep:
            X, _ = pack_systems(systems)
            res = bb.collect(X, False)
        end = time.perf_counter() - self.t0
        done = []
        for s, r in enumerate(reqs):
            out = res.pairlist(s) if res is not None else None
            done.append(("ok", Completion(r.rid, r.system_id, int(po[s + 1] - po[s]), r.arrival, start, end,
                                          len(reqs), out, ret), nbytes if s == 0 else 0))
        return done

    def _resident_for(self, r: Request, reserved: int) -> _Resident:
        """The resident entry of r's system; the bytes reserved for it at
        admission become the resident's (a new resident's estimate, or what
        an existing one's buffers may grow by for a larger request)."""
        with self._lock:
            res = self.residents.get(r.system_id)
            if res is not None:
                self.residents.move_to_end(r.system_id)
                self.inflight_bytes -= reserved
                self.resident_bytes += reserved
                res.bytes += reserved
                return res
            # at the resident limit, the least recently used idle resident
            # hands its device buffers to the new one
            recycled = None
            if len(self.residents) >= self.max_residents:
                for sid in list(self.residents):
                    if sid in self._busy:
                        continue             # admitted or running: not idle
                    old = self.residents[sid]
                    if old.lock.acquire(blocking=False):
                        try:
                            del self.residents[sid]
                            self.resident_bytes -= old.bytes
                            self.evictions += 1
                            recycled = old.detach()
                        finally:
                            old.lock.release()
                        break
        try:
            res = _Resident(r, self.int_type, recycled)
        except BaseException:
            if recycled is not None:
                recycled[0].close()
            raise
        with self._lock:
            self.residents[r.system_id] = res
            self.inflight_bytes -= reserved
            self.resident_bytes += res.bytes
        return res

    # ---------------------------------------------------------------- scheduler

    def _fits(self, nbytes: int) -> bool:
        with self._lock:
            if self.inflight_bytes + self.resident_bytes + nbytes <= self.budget:
                return True
            # evict idle residents, least recently used first
            for sid in list(self.residents):
                if self.inflight_bytes + self.resident_bytes + nbytes <= self.budget:
                    break
                if sid in self._busy:
                    continue                 # admitted or running: not idle
                res = self.residents[sid]
                if res.lock.acquire(blocking=False):
                    try:
                        del self.residents[sid]
                        self.resident_bytes -= res.bytes
                        self.evictions += 1
                        res.free()
                    finally:
                        res.lock.release()
            return self.inflight_bytes + self.resident_bytes + nbytes <= self.budget

    def serve(self, requests: list[Request], *, realtime: bool = True) -> list[Completion]:
        """Serve a request stream.  realtime=True releases each request at its
        arrival time (open loop); False makes every request available at once
        (saturation)."""
        order = sorted(requests, key=lambda r: r.arrival)
        for r in order:
            if estimate_bytes(r.nat, r.cell, r.cutoff) > self.budget:
                raise _lib.NLHMemoryError(_lib.NLH_ENOMEM, f"request {r.rid} ({r.nat} atoms) alone exceeds the "
                                                           f"HBM budget of {self.budget} bytes")
        self._est = {}
        # a fresh count of free workers for this call's threads (the previous
        # call's threads have exited)
        self._free_workers = threading.Semaphore(0)
        self._threads = [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(self.nworkers)]
        for t in self._threads:
            t.start()
        out: list[Completion] = []
        waiting: list[Request] = []
        k = 0
        n = len(order)
        self.t0 = time.perf_counter()
        with self._lock:
            self._busy = set()
        inflight_systems = self._busy

        def complete(d):
            st, c, nb = d
            if st == "error":
                raise self._errors[0]
            with self._lock:
                self.inflight_bytes -= nb       # completions release memory
                inflight_systems.discard(c.system_id)
            out.append(c)

        # no cyclic garbage collection while serving: a collection pass over a
        # large process stalls every thread for milliseconds (reference
        # counting still frees the per-request objects)
        gc_on = gc.isenabled()
        gc.disable()
        # the scheduler and the workers hand the interpreter to each other
        # within 0.2 ms rather than the default 5 ms: a thread woken by a
        # completion or an arrival otherwise waits out a busy thread's slice
        switch = sys.getswitchinterval()
        sys.setswitchinterval(min(switch, 2e-4))
        try:
            while len(out) < n:
                now = time.perf_counter() - self.t0
                while k < n and (not realtime or order[k].arrival <= now):
                    waiting.append(order[k])
                    k += 1
                while True:
                    try:
                        complete(self._done.get_nowait())
                    except queue.Empty:
                        break
                # dispatch while a worker is free
                dispatched = False
                while waiting and self._free_workers.acquire(blocking=False):
                    item = self._next_item(waiting, inflight_systems)
                    if item is None:
                        self._free_workers.release()
                        break
                    self._work.put(item)
                    dispatched = True
                if not dispatched and len(out) < n:
                    # sleep until the next completion or the next arrival
                    # (bounded, so the loop also re-checks admission)
                    t = 2e-3
                    if k < n and realtime:
                        t = min(t, max(order[k].arrival - (time.perf_counter() - self.t0), 0.0))
                    try:
                        complete(self._done.get(timeout=t) if t > 0 else self._done.get_nowait())
                    except queue.Empty:
                        pass
        finally:
            for _ in self._threads:
                self._work.put(None)
            for t in self._threads:
                t.join(timeout=60)
            sys.setswitchinterval(switch)
            if gc_on:
                gc.enable()
        return sorted(out, key=lambda c: c.rid)

    def _estimate(self, r: Request) -> int:
        """estimate_bytes of a request, computed once (a request may be
        looked at by several dispatch passes before it is admitted)."""
        e = self._est.get(id(r))   # the stream's requests stay alive until serve() returns
        if e is None:
            e = self._est[id(r)] = estimate_bytes(r.nat, r.cell, r.cutoff)
        return e

    def _next_item(self, waiting: list[Request], inflight_systems: set):
        """The next unit of work from the admitted head of the queue: one
        large system, or a batch of small ones (FIFO order)."""
        for idx, r in enumerate(waiting):
            if r.system_id in inflight_systems:
                continue                     # a system is built by one worker at a time
            if r.nat >= self.small_atoms:
                with self._lock:
                    # marked busy before the admission check, so that no
                    # eviction in it takes this system's own resident
                    inflight_systems.add(r.system_id)
                    res = self.residents.get(r.system_id)
                # an existing resident grows by what the new request needs
                # beyond its buffers (a larger atom count or a denser cell)
                nb = self._estimate(r) if res is None else max(self._estimate(r) - res.bytes, 0)
                if not self._fits(nb):
                    with self._lock:
                        inflight_systems.discard(r.system_id)
                    return None
                with self._lock:
                    self.inflight_bytes += nb
                    self.peak_bytes = max(self.peak_bytes, self.inflight_bytes + self.resident_bytes)
                del waiting[idx]
                return ("large", [r], nb)
            batch, atoms, nb = [], 0, 0
            rest = []
            for q in waiting[idx:]:
                # (the first request of a batch goes in even above batch_atoms)
                if (q.nat < self.small_atoms and q.system_id not in inflight_systems and
                        (not batch or atoms + q.nat <= self.batch_atoms) and
                        q.system_id not in {b.system_id for b in batch}):
                    e = self._estimate(q)
                    if not self._fits(nb + e):
                        rest.append(q)
                        continue
                    batch.append(q)
                    atoms += q.nat
                    nb += e
                else:
                    rest.append(q)
            if not batch:
                return None
            waiting[idx:] = rest
            with self._lock:
                self.inflight_bytes += nb
                self.peak_bytes = max(self.peak_bytes, self.inflight_bytes + self.resident_bytes)
                for q in batch:
                    inflight_systems.add(q.system_id)
            return ("batch", batch, nb)
        return None

    def close(self):
        with self._lock:
            for res in self.residents.values():
                res.free()
            self.residents.clear()
            self.resident_bytes = 0
            for bb in self._builders:
                if bb is not None:
                    bb.handle.close()
            self._builders = [None] * self.nworkers


def serve_plain(requests: list[Request], *, realtime: bool = True, int_type=Int32,
                keep_results: bool = False, handle: Handle | None = None) -> list[Completion]:
    """The plain per-request path: one single-system build call per request,
    in arrival order, each waiting for its result (nlh_neighbour_list on
    host positions), on `handle` (kept by the caller across streams) or a
    handle of its own."""
    h = handle or Handle()
    ti = _lib.ti_bytes_of(int_type)
    order = sorted(requests, key=lambda r: r.arrival)
    out = []
    gc_on = gc.isenabled()
    gc.disable()   # as in Server.serve
    t0 = time.perf_counter()
    n = C.c_int64()
    try:
        _plain_loop(h, ti, order, out, t0, n, realtime, keep_results)
    finally:
        if gc_on:
            gc.enable()
    if handle is None:
        h.close()
    return sorted(out, key=lambda c: c.rid)


def _plain_loop(h, ti, order, out, t0, n, realtime, keep_results):
    for r in order:
        if realtime:
            while (wait := r.arrival - (time.perf_counter() - t0)) > 0:
                time.sleep(min(wait, 2e-4))
        start = time.perf_counter() - t0
        g = geometry(r.cell, r.cutoff, r.pbc, ti)
        X = np.ascontiguousarray(r.X, dtype=np.float64)
        check(lib().nlh_neighbour_list(h.ptr, X.ctypes.data, X.shape[0], C.byref(g), 0, C.byref(n)), "plain")
        ret = time.perf_counter() - t0
        h.synchronize()   # the pair list is complete
        res = None
        if keep_results:
            from .api import _collect_pairs
            res = _collect_pairs(h, g, X, int(n.value), False, None)
        end = time.perf_counter() - t0
        out.append(Completion(r.rid, r.system_id, int(n.value), r.arrival, start, end, 1, res, ret))


def summarize(done: list[Completion], wall: float | None = None) -> dict:
    """p50_ms / p99_ms: arrival to a complete pair list on the device;
    p50_return_ms / p99_return_ms: arrival to the build call's return (pair
    count known, the last kernel possibly still writing pairs)."""
    lat = np.array([c.latency for c in done]) * 1e3
    ret = np.array([c.return_latency for c in done]) * 1e3
    t_end = max(c.end for c in done)
    t_first = min(c.arrival for c in done)
    span = wall if wall is not None else (t_end - t_first)
    pairs = sum(c.npairs for c in done)
    return {"requests": len(done), "span_s": round(span, 4), "systems_per_s": round(len(done) / span, 1),
            "pairs_per_s": round(pairs / span, 1), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "p50_return_ms": round(float(np.percentile(ret, 50)), 3),
            "p99_return_ms": round(float(np.percentile(ret, 99)), 3),
            "mean_batch": round(float(np.mean([c.batch for c in done])), 2)}
