"""Seeded synthetic inputs for the benchmark and test workloads.

A portable SplitMix64 stream (u = (x >> 11) * 2^-53) drives every
generator so inputs are reproducible anywhere.  Position conventions follow
the reference's test helpers: cubic boxes X = L * u (test/test_utils.jl:28-34,
scripts/benchmark.jl:57-64) and triclinic cells X = C' u
(test/test_utils.jl:306-310), with ROWS of the cell matrix the lattice
vectors.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n consecutive SplitMix64 outputs of the stream `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n uniforms in [0, 1) with 53-bit resolution."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, n: int) -> np.ndarray:
    """n standard normals by Box-Muller on the uniform stream."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[:m]          # (0, 1]
    u2 = u[m:]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


@dataclass
class System:
    X: np.ndarray          # (nat, 3) float64
    cell: np.ndarray       # 3x3, rows = lattice vectors
    pbc: tuple
    cutoff: float

    @property
    def nat(self) -> int:
        return int(self.X.shape[0])


def random_cubic(N: int, density: float = 0.05, seed: int = 1, cutoff: float = 5.0,
                 pbc=(True, True, True)) -> System:
    """rand_config / benchmark generate_system: uniform positions in a cube of
    side L = (N/density)^(1/3)."""
    L = (N / density) ** (1.0 / 3.0)
    X = (L * uniform(seed, 3 * N)).reshape(N, 3)
    return System(X, np.eye(3) * L, tuple(bool(b) for b in pbc), float(cutoff))


def random_in_cell(N: int, cell, seed: int = 1, cutoff: float = 3.0, pbc=(True, True, True)) -> System:
    """X = C' u for u uniform in [0,1)^3 (test_utils.jl:306-310)."""
    C = np.asarray(cell, dtype=np.float64)
    u = uniform(seed, 3 * N).reshape(N, 3)
    return System(u @ C, C, tuple(bool(b) for b in pbc), float(cutoff))


def fcc(a: float, reps=(1, 1, 1)) -> tuple[np.ndarray, np.ndarray]:
    """Cubic FCC crystal: 4-atom conventional cell replicated reps times."""
    basis = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]], dtype=np.float64)
    nx, ny, nz = reps
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    X = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    return X, np.diag([nx * a, ny * a, nz * a]).astype(np.float64)


def triclinic_fcc_slab(n: int = 63, a: float = 3.61, sigma: float = 0.05, vacuum: float = 30.0,
                       seed: int = 3, cutoff: float = 5.0) -> System:
    """FCC a=3.61, n^3 conventional cells (4 n^3 atoms) jittered by N(0, sigma),
    sheared cell rows A1=(n a, 0, 0), A2=(round(0.3 n) a, n a, 0) (a lattice
    translation, so the periodic crystal stays seamless), A3=(0.2 n a, 0,
    n a + vacuum); pbc = (true, true, false).  Atoms stay in the orthogonal
    box, so many have fractional coordinates outside [0, 1)."""
    X, _ = fcc(a, (n, n, n))
    X = X + sigma * normal(seed, X.size).reshape(X.shape)
    cell = np.array([[n * a, 0.0, 0.0],
                     [round(0.3 * n) * a, n * a, 0.0],
                     [0.2 * n * a, 0.0, n * a + vacuum]])
    return System(X, cell, (True, True, False), float(cutoff))


def batch_systems(nsys: int = 4096, seed: int = 7, nmin: int = 8, nmax: int = 2048,
                  tiny_frac: float = 0.10) -> list[System]:
    """Many small systems: N log-uniform in [nmin, nmax]; triclinic cell with
    diagonal (N/rho)^(1/3), rho ~ U[0.04, 0.09], off-diagonals U[-0.15, 0.15]
    times the diagonal; per-axis pbc ~ Bernoulli(1/2); cutoff ~ U[4, 6];
    X = C' u.  A fraction tiny_frac are 1-4 atom cells smaller than the
    cutoff (several periodic images per pair)."""
    u = uniform(seed, 16 * nsys).reshape(nsys, 16)
    out = []
    for s in range(nsys):
        r = u[s]
        tiny = r[0] < tiny_frac
        if tiny:
            N = 1 + int(r[1] * 4)
        else:
            N = int(round(np.exp(np.log(nmin) + r[1] * (np.log(nmax) - np.log(nmin)))))
        rho = 0.04 + 0.05 * r[2]
        d = (N / rho) ** (1.0 / 3.0)
        C = np.diag([d, d, d])
        off = (r[3:9] * 0.30 - 0.15) * d
        C[0, 1], C[0, 2], C[1, 0], C[1, 2], C[2, 0], C[2, 1] = off
        pbc = tuple(bool(b) for b in (r[9:12] < 0.5))
        if tiny:
            pbc = (True, True, True) if r[12] < 0.5 else pbc
        rc = 4.0 + 2.0 * r[13]
        X = uniform(seed * 1000003 + s, 3 * N).reshape(N, 3) @ C
        out.append(System(X, C, pbc, float(rc)))
    return out


def served_stream(nreq: int, rate: float, seed: int = 11, nmin: int = 16, nmax: int = 200_000,
                  rebuild_frac: float = 0.30, sigma: float = 0.1, cutoff: float = 5.0):
    """A stream of neighbour-list requests: Poisson arrivals at `rate` per
    second; sizes log-uniform in [nmin, nmax]; density U[0.05, 0.09]; cubic
    or triclinic cells (off-diagonals U[-0.15, 0.15] times the diagonal);
    per-axis pbc ~ Bernoulli(1/2); cutoff 5.  A fraction `rebuild_frac` of the
    requests are new positions of a system requested before (every atom
    displaced by N(0, sigma)).  Returns serve.Request objects."""
    from .serve import Request

    u = uniform(seed, 12 * nreq).reshape(nreq, 12)
    gaps = -np.log(1.0 - uniform(seed + 1, nreq)) / rate
    t = np.cumsum(gaps) - gaps[0]
    out, systems = [], []
    for k in range(nreq):
        r = u[k]
        if systems and r[0] < rebuild_frac:
            sid = int(r[1] * len(systems))
            base = systems[sid]
            X = base.X + sigma * normal(seed * 7919 + k, base.X.size).reshape(base.X.shape)
            out.append(Request(k, sid, X, base.cell, base.pbc, base.cutoff, True, float(t[k])))
            continue
        N = int(round(np.exp(np.log(nmin) + r[2] * (np.log(nmax) - np.log(nmin)))))
        rho = 0.05 + 0.04 * r[3]
        d = (N / rho) ** (1.0 / 3.0)
        C = np.diag([d, d, d])
        if r[4] < 0.5:
            off = (r[5:11] * 0.30 - 0.15) * d
            C[0, 1], C[0, 2], C[1, 0], C[1, 2], C[2, 0], C[2, 1] = off
        pbc = tuple(bool(b) for b in (uniform(seed * 31 + k, 3) < 0.5))
        X = uniform(seed * 1000003 + k, 3 * N).reshape(N, 3) @ C
        sysm = System(X, C, pbc, float(cutoff))
        sid = len(systems)
        systems.append(sysm)
        out.append(Request(k, sid, X, C, pbc, float(cutoff), False, float(t[k])))
    return out
