"""Synthetic inputs for the LJ / SPH / DEM workloads.

All randomness comes from SplitMix64 (53-bit uniforms, Box-Muller normals)
written here, so every machine generates identical bits for a seed.  The
reference drivers use std::mt19937 + std::normal_distribution
(testMD1.cu:20-53), whose normal algorithm is implementation-defined, so the
workloads are restated, not reproduced.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n consecutive SplitMix64 outputs of the stream started at `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform53(seed: int, n: int, offset: int = 0) -> np.ndarray:
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normals(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n standard normals by Box-Muller on pairs of 53-bit uniforms."""
    m = (n + 1) // 2
    u = uniform53(seed, 2 * m, offset)
    r = np.sqrt(-2.0 * np.log1p(-u[0::2]))
    th = 2.0 * math.pi * u[1::2]
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(th)
    out[1::2] = r * np.sin(th)
    return out[:n]


def fcc_sites(ncell: int, a: float) -> np.ndarray:
    """FCC lattice of ncell^3 cubic cells x 4 basis sites, cell-major order."""
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    g = np.arange(ncell, dtype=np.float64)
    cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    return ((cells[:, None, :] + basis[None, :, :]) * a).reshape(-1, 3)


def mb_velocities(n: int, T: float, m: np.ndarray, seed: int) -> np.ndarray:
    """Maxwell-Boltzmann velocities, zero net momentum, exactly T with 3N-3 dof."""
    v = normals(seed, 3 * n, offset=0).reshape(n, 3) * np.sqrt(T / m)[:, None]
    p = (m[:, None] * v).sum(axis=0) / m.sum()
    v -= p[None, :]
    tcur = float((m[:, None] * v * v).sum()) / (3 * n - 3)
    return v * math.sqrt(T / tcur)


@dataclass
class LJSystem:
    r: np.ndarray          # (n,3) float64
    v: np.ndarray          # (n,3)
    m: np.ndarray          # (n,)
    type: np.ndarray       # (n,) uint32
    cell: np.ndarray       # (6,) xmin,xmax,ymin,ymax,zmin,zmax
    rho: float
    T: float
    seed: int
    sigma_eps: np.ndarray = field(default_factory=lambda: np.array([1.0, 1.0]))
    rc: float = 2.5

    @property
    def n(self):
        return self.r.shape[0]

    @property
    def L(self):
        return float(self.cell[1] - self.cell[0])


def lj_fcc(ncell: int, rho: float = 0.8442, T: float = 0.72, seed: int = 1,
           natoms: int | None = None) -> LJSystem:
    """LJ fluid start: FCC lattice (first `natoms` sites of ncell^3 x 4),
    box L = (N/rho)^(1/3), MB velocities.  Reduced units sigma=eps=m=kB=1."""
    nsite = 4 * ncell ** 3
    n = nsite if natoms is None else int(natoms)
    assert 0 < n <= nsite
    L = (n / rho) ** (1.0 / 3.0)
    a = L / ncell
    r = fcc_sites(ncell, a)[:n]
    m = np.ones(n)
    v = mb_velocities(n, T, m, seed)
    cell = np.array([0.0, L, 0.0, L, 0.0, L])
    return LJSystem(r=r, v=v, m=m, type=np.zeros(n, dtype=np.uint32), cell=cell, rho=rho, T=T,
                    seed=seed)


def lj_config1() -> LJSystem:
    """N=4000 (FCC 10^3 x 4), rho 0.8442, T 0.72, seed 1."""
    return lj_fcc(10, 0.8442, 0.72, seed=1)


def lj_config2() -> LJSystem:
    """N=1,048,576 (FCC 64^3 x 4), rho 0.8442, T 0.72, seed 2."""
    return lj_fcc(64, 0.8442, 0.72, seed=2)


@dataclass
class Batch:
    systems: list

    @property
    def counts(self):
        return np.array([s.n for s in self.systems], dtype=np.uint32)

    def concat(self):
        r = np.concatenate([s.r for s in self.systems])
        v = np.concatenate([s.v for s in self.systems])
        m = np.concatenate([s.m for s in self.systems])
        t = np.concatenate([s.type for s in self.systems])
        cells = np.concatenate([s.cell for s in self.systems])
        return r, v, m, t, cells


def lj_replicas(nrep: int = 256, nmin: int = 500, nmax: int = 20000, seed: int = 5) -> Batch:
    """Batched LJ replicas: N_k = round(exp(U(ln nmin, ln nmax))), rho_k ~
    U(0.7, 0.9), T_k ~ U(0.7, 2.0); replica k uses seed `seed + k`."""
    u = uniform53(seed, 3 * nrep, offset=1 << 40).reshape(nrep, 3)
    out = []
    for k in range(nrep):
        n = int(round(math.exp(math.log(nmin) + u[k, 0] * (math.log(nmax) - math.log(nmin)))))
        rho = 0.7 + 0.2 * u[k, 1]
        T = 0.7 + 1.3 * u[k, 2]
        nc = int(math.ceil((n / 4.0) ** (1.0 / 3.0) - 1e-12))
        out.append(lj_fcc(nc, rho, T, seed=seed + k, natoms=n))
    return Batch(out)
