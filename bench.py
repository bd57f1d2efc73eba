"""Benchmark of the sort-based neighbour-list build on one MI355X.

Headline (the reference's README benchmark, README.md:119-133 /
scripts/benchmark.jl:57-151): pairs per second of a full build --
cell list + pair materialisation -- of 100,000 atoms uniform-random in a cube
at 0.05 atoms/A^3 (L = 125.99 A), cutoff 5 A, full pbc, Float64, with the
positions already resident in HBM.

Also reported on the same JSON line:
  roofline      fill kernel (PairList writer) at 1,000,000 atoms, HBM-bound:
                algorithmic bytes 20 P + 32 nat + 4 (M+1) per launch over its
                mean duration from HIP events on the build stream
  count_fp64    the distance-test kernel's FP64 rate (11 flop per pair test)
  count_valu    the same kernel's VALU issue rate: wave64 VALU instructions per
                launch (rocprofv3 SQ_INSTS_VALU, scaled to the whole chip,
                profiles/count_valu.json) over its duration measured here
  builds        1M-atom cube and the 1,000,188-atom sheared FCC slab
  batched       4,096 small systems in one batched build vs one call per system
  cpu_baseline  the CPU oracle (C restatement of the reference's sort-based
                path) on the host cores, bounded sample
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "neighbourlists.jl_amd"))
sys.path.insert(0, REPO)

import numpy as np  # noqa: E402

A4500_PAIRS_PER_S = 2.618e6 / 6.9e-3   # README.md:129: 100k atoms, 2.6M pairs, 6.9 ms on an RTX A4500
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
# one wave64 VALU instruction per SIMD every 4 cycles (16 lanes): 256 CUs x 4 SIMDs x 2.4 GHz / 4
PEAK_VALU_GIPS = 256 * 4 * 2.4 / 4


def _sync():
    """Device-wide wait (the builds already wait on their own stream)."""
    from nlhip import _lib

    _lib.lib().nlh_device_synchronize()


def _single(s, steps: int, warmup: int, profile: bool = False):
    """Time `steps` full builds of one system (X resident on the device)."""
    from nlhip import _lib

    X = _lib.DeviceArray(np.ascontiguousarray(s.X, dtype=np.float64))   # resident in HBM
    g = _lib.geometry(s.cell, s.cutoff, s.pbc, 4)
    h = _lib.Handle()
    n = C.c_int64()
    L = _lib.lib()

    def step():
        _lib.check(L.nlh_neighbour_list(h.ptr, X.ptr, X.shape[0], C.byref(g), _lib.NLH_X_DEVICE, C.byref(n)))

    for _ in range(warmup):
        step()
    _sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    _sync()
    wall = (time.perf_counter() - t0) / steps
    # isolated builds: each one waits until its pair list is complete before
    # the next starts (no overlap of a build's fill with the next call)
    iso = []
    for _ in range(max(5, min(steps, 20))):
        t1 = time.perf_counter()
        step()
        h.synchronize()
        iso.append(time.perf_counter() - t1)
    phases = None
    if profile:   # separate, event-instrumented runs for the per-kernel durations
        h.set_profiling(True)
        acc = {}
        reps = max(5, min(steps, 20))
        for _ in range(reps):
            step()
            for k, v in h.phase_times().items():
                acc.setdefault(k, []).append(v)
        h.set_profiling(False)
        phases = {k: statistics.mean(v) for k, v in acc.items()}
    return {"npairs": int(n.value), "nat": int(X.shape[0]), "ms": wall * 1e3, "pairs_per_s": n.value / wall,
            "ncells": int(g.ncells_total), "phases_ms": phases,
            "isolated_ms": statistics.median(iso) * 1e3, "isolated_min_ms": min(iso) * 1e3}


def _host_inclusive(s, steps: int) -> dict:
    """The same build with positions handed over from host memory and i, j,
    S, first copied back to host arrays (PCIe both ways, inside the timing)."""
    from nlhip import _lib

    X = np.ascontiguousarray(s.X, dtype=np.float64)
    g = _lib.geometry(s.cell, s.cutoff, s.pbc, 4)
    h = _lib.Handle()
    n = C.c_int64()
    L = _lib.lib()
    _lib.check(L.nlh_neighbour_list(h.ptr, X.ctypes.data, X.shape[0], C.byref(g), 0, C.byref(n)))
    P = int(n.value)
    i = np.empty(P, np.int32)
    j = np.empty(P, np.int32)
    S = np.empty((P, 3), np.int32)
    first = np.empty(X.shape[0] + 1, np.int32)
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        _lib.check(L.nlh_neighbour_list(h.ptr, X.ctypes.data, X.shape[0], C.byref(g), 0, C.byref(n)))
        _lib.check(L.nlh_copy_pairlist(h.ptr, i.ctypes.data, j.ctypes.data, S.ctypes.data, first.ctypes.data, None))
        ts.append(time.perf_counter() - t0)
    t = statistics.median(ts)
    return {"ms": round(t * 1e3, 4), "pairs_per_s": round(P / t, 1),
            "note": "positions uploaded from pageable host memory and the PairList copied back, median of steps"}


def _roofline(r) -> dict:
    P, nat, M = r["npairs"], r["nat"], r["ncells"]
    b = 20 * P + 32 * nat + 4 * (M + 1)
    t = r["phases_ms"]["fill"] * 1e-3
    achieved = b / t / 1e9
    traffic = None
    tf = os.path.join(REPO, "profiles", "fill_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("fill_bytes_per_launch")
    return {"kernel": "fill_kernel", "workload": f"cube {nat} atoms", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
            "algorithmic_bytes": b, "fill_us": round(t * 1e6, 2)}


def _count_rate(r, cand_per_atom: float) -> dict:
    tests = r["nat"] * cand_per_atom
    t = r["phases_ms"]["count"] * 1e-3
    tflops = tests * 11 / t / 1e12
    return {"kernel": "count_kernel", "bound": "fp64-vector", "achieved": round(tflops, 3), "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tflops / PEAK_FP64_TFLOPS, 4), "count_us": round(t * 1e6, 2)}


def _count_valu(r) -> dict | None:
    """VALU issue rate of the count kernel at 1M atoms: the instruction count
    of one launch (a property of the workload) from the committed counter
    profile, over the launch's duration measured in this run."""
    pf = os.path.join(REPO, "profiles", "count_valu.json")
    if not os.path.exists(pf):
        return None
    with open(pf) as f:
        n = json.load(f)["valu_instructions_per_launch"]
    t = r["phases_ms"]["count"] * 1e-3
    gips = n / t / 1e9
    return {"kernel": "count_kernel", "workload": f"cube {r['nat']} atoms", "bound": "valu-issue",
            "achieved": round(gips, 1), "peak": PEAK_VALU_GIPS, "unit": "G wave64-instr/s",
            "frac": round(gips / PEAK_VALU_GIPS, 4), "instructions": n, "count_us": round(t * 1e6, 2)}


def _batched(nsys: int, steps: int):
    from nlhip import _lib, workloads
    from nlhip.batch import BatchBuilder, pack_systems

    systems = workloads.batch_systems(nsys, seed=7)
    X, offs = pack_systems(systems)
    Xd = _lib.DeviceArray(X)
    bb = BatchBuilder()
    bb.run(systems, X_device=Xd, copy=False)   # warm-up (buffers, capacities)
    tuned = bb.autotune(X, X_device=Xd, reps=1)   # launch configuration for this batch (untimed)
    bb.rebuild(X, X_device=Xd, copy=False)     # capture the pipeline with it
    _sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        P = bb.rebuild(X, X_device=Xd, copy=False)   # same systems, positions in HBM
    _sync()
    tb = (time.perf_counter() - t0) / steps
    # plain path: the single-system call per system, X resident, each call waits
    geoms = [_lib.geometry(s.cell, s.cutoff, s.pbc, 4) for s in systems]
    h = _lib.Handle()
    n = C.c_int64()
    L = _lib.lib()
    base = Xd.ptr

    def plain():
        tot = 0
        for k, s in enumerate(systems):
            _lib.check(L.nlh_neighbour_list(h.ptr, base + 24 * int(offs[k]), int(offs[k + 1] - offs[k]),
                                            C.byref(geoms[k]), _lib.NLH_X_DEVICE, C.byref(n)))
            tot += n.value
        return tot

    plain()
    _sync()
    t0 = time.perf_counter()
    Pp = plain()
    _sync()
    tp = time.perf_counter() - t0
    assert Pp == P, (Pp, P)
    return {"nsys": nsys, "atoms": int(offs[-1]), "pairs": int(P), "batched_ms": round(tb * 1e3, 3),
            "tuning": {"brick_atoms": tuned.brick_atoms, "fill_block": tuned.fill_block,
                       "fill_pairs": tuned.fill_pairs, "count_block": tuned.count_block,
                       "fused_small": tuned.fused_small},
            "batched_systems_per_s": round(nsys / tb, 1), "plain_ms": round(tp * 1e3, 3),
            "plain_systems_per_s": round(nsys / tp, 1), "speedup": round(tp / tb, 2)}


def _batched_rebuild(nsys: int, steps: int, warmup: int) -> dict:
    """The batched rebuild alone, default launch configuration (counter runs:
    no plain path, no autotuner trials)."""
    from nlhip import _lib, workloads
    from nlhip.batch import BatchBuilder, pack_systems

    systems = workloads.batch_systems(nsys, seed=7)
    X, offs = pack_systems(systems)
    Xd = _lib.DeviceArray(X)
    bb = BatchBuilder()
    bb.run(systems, X_device=Xd, copy=False)
    for _ in range(warmup):
        bb.rebuild(X, X_device=Xd, copy=False)
    _sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        P = bb.rebuild(X, X_device=Xd, copy=False)
    _sync()
    tb = (time.perf_counter() - t0) / max(steps, 1)
    return {"nsys": nsys, "atoms": int(offs[-1]), "pairs": int(P), "batched_ms": round(tb * 1e3, 3)}


def _served(nreq: int, seed: int = 11) -> dict:
    """Request stream (Poisson arrivals; 16..200,000 atoms; 30% rebuilds of a
    resident system; 64 GB budget) through the continuous-batching server and
    through the plain per-request call, both warmed by two full passes over
    the stream.  Saturation (all requests available at once) for the capacities,
    then open loop at 75% of the plain path's capacity for the latencies."""
    from nlhip import serve, workloads

    from nlhip import _lib

    reqs = workloads.served_stream(nreq, rate=1.0, seed=seed)
    for r in reqs:
        r.arrival = 0.0
    # both paths at steady state: two full passes over the stream first
    # (device buffers grown, resident systems held), as a long-running
    # service is
    ph = _lib.Handle()
    for _ in range(2):
        serve.serve_plain(reqs, realtime=False, handle=ph)
    plain_sat = serve.serve_plain(reqs, realtime=False, handle=ph)
    sp = serve.summarize(plain_sat)
    srv = serve.Server()
    for _ in range(2):
        srv.serve(reqs, realtime=False)
    served_sat = srv.serve(reqs, realtime=False)
    rate = 0.75 * sp["systems_per_s"]
    gaps = -np.log(1.0 - workloads.uniform(seed + 5, nreq)) / rate
    t = np.cumsum(gaps) - gaps[0]
    for r, a in zip(reqs, t):
        r.arrival = float(a)
    served_open = srv.serve(reqs, realtime=True)
    plain_open = serve.serve_plain(reqs, realtime=True, handle=ph)
    srv.close()
    ph.close()
    so, po, ss = serve.summarize(served_open), serve.summarize(plain_open), serve.summarize(served_sat)
    return {"requests": nreq, "mean_atoms": round(float(np.mean([r.nat for r in reqs])), 1),
            "rebuild_frac": round(float(np.mean([r.rebuild for r in reqs])), 3),
            "arrival_rate_per_s": round(rate, 1),
            "open_loop": {"served": so, "plain": po},
            "saturation": {"served": ss, "plain": sp,
                           "speedup_systems_per_s": round(ss["systems_per_s"] / sp["systems_per_s"], 2)}}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_share() -> tuple[int, int, int]:
    """(nproc: CPUs this process may run on, os.cpu_count(), threads used).
    The threads are nproc capped at 16, the host-CPU share of one GPU on the
    benchmark machines (their OMP_NUM_THREADS says the same)."""
    try:
        nproc = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        nproc = os.cpu_count() or 1
    cap = 16
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        cap = min(cap, int(env))
    return nproc, os.cpu_count() or 1, max(1, min(nproc, cap))


def _time_cpu(s, nt: int, samples: int) -> tuple[int, list]:
    import oracle

    oracle.time_neighbour_list(s.X, s.cutoff, s.cell, s.pbc, nt)   # warm-up
    ts, P = [], 0
    for _ in range(samples):
        t0 = time.perf_counter()
        P = oracle.time_neighbour_list(s.X, s.cutoff, s.cell, s.pbc, nt)
        ts.append(time.perf_counter() - t0)
    return P, ts


def _cpu_baseline(s) -> dict:
    """The C restatement of the reference's CPU sort-based path (the
    reference's own `julia -t N` cannot run here: no Julia), timed as the
    reference's benchmark does (scripts/benchmark.jl:124-151: minimum of at
    least 5 samples after a warm-up), at N threads and at one thread."""
    nproc, ncpu, nt = _cpu_share()
    P, ts = _time_cpu(s, nt, 20)   # about 15 s of CPU time at 16 threads
    _, t1 = _time_cpu(s, 1, 10)    # about 5 s
    return {"value": round(P / min(ts), 1), "unit": "pairs/s", "cores": nt, "kind": "port",
            "median_value": round(P / statistics.median(ts), 1),
            "value_1_thread": round(P / min(t1), 1), "median_value_1_thread": round(P / statistics.median(t1), 1),
            "nproc": nproc, "cpu_count": ncpu, "cpu_model": _cpu_model(),
            "sample": f"the 100k-atom cube, 1 warm-up then 20 full builds at {nt} threads and 10 at one thread "
                      f"(min reported, median beside it); C restatement of the reference CPU sort-based path "
                      f"(serial binning, sort and offsets; count and fill over the threads); threads = nproc "
                      f"({nproc}) capped at 16, one GPU's host-CPU share"}


def _copy_calibration(nbytes: int, reps: int) -> dict:
    """Device-to-device copies of a known size: the reference point that turns
    the HBM byte counters of a profiler run into bytes."""
    from nlhip import _lib

    a = _lib.DeviceArray(np.zeros(nbytes // 8))
    b = _lib.DeviceArray(np.zeros(nbytes // 8))
    for _ in range(reps):
        _lib.check(_lib.lib().nlh_memcpy(b.ptr, a.ptr, nbytes), "copy")
    _sync()
    a.free()
    b.free()
    return {"copy_bytes": nbytes, "reps": reps}


def _only(name: str, steps: int, warmup: int) -> dict:
    """One workload of the full run, alone (for profiler runs)."""
    from nlhip import workloads

    if name == "cube_100k":
        return _single(workloads.random_cubic(100_000, density=0.05, seed=1, cutoff=5.0), steps, warmup)
    if name == "cube_1M":
        return _single(workloads.random_cubic(1_000_000, density=0.05, seed=2, cutoff=5.0), steps, warmup)
    if name == "fcc_slab_1M":
        return _single(workloads.triclinic_fcc_slab(), steps, warmup)
    if name == "batched":
        return _batched(4096, steps)
    if name == "batched_rebuild":
        return _batched_rebuild(4096, steps, warmup)
    if name == "served":
        return _served(steps)
    if name == "copy":
        return _copy_calibration(512 << 20, steps)
    raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--only", choices=["cube_100k", "cube_1M", "fcc_slab_1M", "batched", "batched_rebuild", "served",
                                       "copy"],
                    help="run one workload alone (profiler runs); prints its own JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (--gpus 1): the library builds each neighbour list on one device", file=sys.stderr)
        sys.exit(2)

    import nlhip
    from nlhip import workloads

    nlhip.init(0)
    if args.only:
        r = _only(args.only, args.steps, args.warmup)
        print(json.dumps({"workload": args.only, **{k: v for k, v in r.items() if k != "phases_ms"}}), flush=True)
        return
    c1 = workloads.random_cubic(100_000, density=0.05, seed=1, cutoff=5.0)
    head = _single(c1, args.steps, args.warmup, profile=True)
    out = {
        "metric": "pairs/sec at 100k atoms",
        "value": round(head["pairs_per_s"], 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(head["pairs_per_s"] / A4500_PAIRS_PER_S, 2),
        "baseline_source": "reference README.md:129 (RTX A4500, 100k atoms, 2.6M pairs, 6.9 ms => 3.79e8 pairs/s)",
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cube 100,000 atoms, uniform-random (seed 1), density 0.05/A^3, L=125.99 A, "
                               "cutoff 5.0 A, pbc=(T,T,T), Float64, Int32 indices",
                   "npairs": head["npairs"], "ncells": head["ncells"]},
        "phases_ms_100k": {k: round(v, 4) for k, v in head["phases_ms"].items()},
        "isolated_100k": {"ms": round(head["isolated_ms"], 4), "min_ms": round(head["isolated_min_ms"], 4),
                          "pairs_per_s": round(head["npairs"] / (head["isolated_ms"] * 1e-3), 1),
                          "note": "one build at a time, each waited for until its pair list is complete "
                                  "(median); `value` is the pipelined rate of back-to-back builds, each "
                                  "returning when its pair count is final"},
    }
    if not args.quick:
        out["host_inclusive_100k"] = _host_inclusive(c1, max(5, min(args.steps, 20)))
        c1m = workloads.random_cubic(1_000_000, density=0.05, seed=2, cutoff=5.0)
        big = _single(c1m, 10, 3, profile=True)
        out["roofline"] = _roofline(big)
        out["count_fp64"] = _count_rate(big, 27 * big["nat"] / big["ncells"])
        out["count_valu"] = _count_valu(big)
        slab = workloads.triclinic_fcc_slab()
        sl = _single(slab, 10, 3, profile=True)
        out["builds"] = {
            "cube_1M": {"atoms": big["nat"], "pairs": big["npairs"], "ms": round(big["ms"], 3),
                        "pairs_per_s": round(big["pairs_per_s"], 1),
                        "phases_ms": {k: round(v, 4) for k, v in big["phases_ms"].items()}},
            "fcc_slab_1M": {"atoms": sl["nat"], "pairs": sl["npairs"], "ms": round(sl["ms"], 3),
                            "pairs_per_s": round(sl["pairs_per_s"], 1),
                            "phases_ms": {k: round(v, 4) for k, v in sl["phases_ms"].items()}},
        }
        out["batched"] = _batched(4096, 5)
        out["served"] = _served(600)
        out["cpu_baseline"] = _cpu_baseline(c1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
