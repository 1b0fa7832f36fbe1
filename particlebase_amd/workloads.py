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


# ------------------------------------------------------------------ SPH
@dataclass
class SPHSystem:
    r: np.ndarray        # (n, 3)
    m: np.ndarray
    fixed: np.ndarray    # uint8, 1 = wall particle
    mu0: np.ndarray
    c: np.ndarray        # sound speed
    h: float
    cell: np.ndarray     # (6,) box
    dx: float
    rho0: float = 1.0    # rest density (config physics: Tait reference density)

    @property
    def n(self):
        return self.r.shape[0]


def sph_dam_break(dx: float = 0.4 / 63, box=(1.6, 0.4, 0.8), column=(0.4, 0.4, 0.8),
                  layers: int = 3, hfac: float = 1.7, g: float = 9.81, seed: int = 3,
                  jitter: float = 0.0) -> SPHSystem:
    """3-D dam break in the reference's SPH units: a fluid column on a cubic
    lattice of spacing dx in the corner of a box lined with `layers` layers
    of fixed wall particles.  Density is relative (rho0 = 1, the constant
    cudaParticleSPH_NS.cu:91 uses), so every particle has mass dx^3; the
    kernel radius is h = hfac dx (Lucy kernel, support h); the sound speed
    is c0 = 10 sqrt(2 g H) (ten times the largest free-fall speed).  With
    dx = 0.4/63 the column holds 63 x 63 x 126 = 500,094 fluid particles.
    `jitter` (in units of dx) displaces fluid particles by uniforms."""
    nx, ny, nz = (int(round(column[a] / dx)) for a in range(3))
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    fluid = (np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1) + 0.5) * dx
    if jitter > 0:
        fluid += (uniform53(seed, fluid.size).reshape(-1, 3) - 0.5) * (jitter * dx)
    # walls: floor and four side walls, `layers` deep, outside the box
    wx, wy, wz = (int(round(box[a] / dx)) for a in range(3))
    gx = np.arange(-layers, wx + layers)
    gy = np.arange(-layers, wy + layers)
    gz = np.arange(-layers, wz)
    WX, WY, WZ = np.meshgrid(gx, gy, gz, indexing="ij")
    Wp = np.stack([WX.ravel(), WY.ravel(), WZ.ravel()], 1)
    inside = ((Wp[:, 0] >= 0) & (Wp[:, 0] < wx) & (Wp[:, 1] >= 0) & (Wp[:, 1] < wy) &
              (Wp[:, 2] >= 0))
    wall = (Wp[~inside] + 0.5) * dx
    r = np.concatenate([wall, fluid])
    n, nw = r.shape[0], wall.shape[0]
    fixed = np.zeros(n, dtype=np.uint8)
    fixed[:nw] = 1
    m = np.full(n, dx ** 3)
    H = column[2]
    c0 = 10.0 * math.sqrt(2.0 * g * H)
    mu0 = np.full(n, 1.0e-6)
    mu0[:nw] = 1.0e-3
    lo = -layers * dx
    cell = np.array([lo, box[0] - lo, lo, box[1] - lo, lo, box[2] + layers * dx])
    return SPHSystem(r=r, m=m, fixed=fixed, mu0=mu0, c=np.full(n, c0), h=hfac * dx, cell=cell,
                     dx=dx)


def m4_lattice_sum(hfac: float) -> float:
    """sum over the cubic lattice Z^3 of W_M4(|n| dx) dx^3 for h = hfac dx
    (the M4 cubic spline of support 2h, 3-D normalisation 1/(pi h^3))."""
    k = int(math.ceil(2 * hfac))
    g = np.arange(-k, k + 1, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    q = np.sqrt(X * X + Y * Y + Z * Z).ravel() / hfac
    t1 = np.maximum(1.0 - q, 0.0)
    t2 = np.maximum(2.0 - q, 0.0)
    return float((0.25 * t2 ** 3 - t1 ** 3).sum() / (math.pi * hfac ** 3))


def sph_dam_break_config(dx: float = 0.4 / 63, box=(1.6, 0.4, 0.8), column=(0.4, 0.4, 0.8),
                         layers: int = 3, hfac: float = 1.3, g: float = 9.81,
                         rho0: float = 1000.0, seed: int = 3, jitter: float = 0.0) -> SPHSystem:
    """The dam break with the config physics (cubic spline of support 2h,
    Tait, artificial viscosity) in SI units: the same lattice, walls and
    box as sph_dam_break, h = 1.3 dx, rest density 1000 kg/m^3 and
    c0 = 10 sqrt(2 g H).  The particle mass is rho0 dx^3 divided by the
    kernel's lattice sum, so a particle with full support on the initial
    lattice has density rho0 exactly (p = 0 at rest).  Three wall layers
    (3 dx) cover the support 2.6 dx."""
    w = sph_dam_break(dx, box, column, layers, hfac, g, seed, jitter)
    w.m = np.full(w.n, rho0 * dx ** 3 / m4_lattice_sum(hfac))
    w.rho0 = rho0
    w.mu0 = np.zeros(w.n)
    return w


# ------------------------------------------------------------------ DEM
@dataclass
class DEMSystem:
    r: np.ndarray
    m: np.ndarray
    rad: np.ndarray
    fixed: np.ndarray
    cell: np.ndarray
    v: np.ndarray | None = None

    @property
    def n(self):
        return self.r.shape[0]


def dem_settle(n: int = 200_000, rmin: float = 0.9e-3, rmax: float = 1.1e-3,
               box=(0.1, 0.1, 0.4), density: float = 7874.0, seed: int = 4,
               gap: float = 0.05, fill: str = "random") -> DEMSystem:
    """Granular settle: n spheres with radii uniform in [rmin, rmax] at
    non-overlapping random positions in the box.  Centres are drawn as the
    first n sites of a jittered cubic lattice of pitch 2 rmax (1 + gap),
    taken in random order, with a jitter small enough that no two spheres
    overlap; 200,000 spheres of ~1 mm in 0.1 x 0.1 x 0.4 m fill ~21 % of the
    volume.  fill="bottom" takes the lowest n sites instead (a packed bed;
    with gap < 0 neighbouring spheres overlap slightly).  Masses are those of
    solid spheres of the given density."""
    pitch = 2.0 * rmax * (1.0 + gap)
    nx, ny, nz = (int(box[a] // pitch) for a in range(3))
    if nx * ny * nz < n:
        raise ValueError("box too small for n spheres at this pitch")
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    sites = (np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1) + 0.5) * pitch
    if fill == "bottom":
        sites = sites[np.argsort(sites[:, 2], kind="stable")[:n]]
    else:
        order = np.argsort(uniform53(seed, sites.shape[0]))
        sites = sites[np.sort(order[:n])]
    jit = (uniform53(seed + 1, 3 * n).reshape(n, 3) - 0.5) * (pitch - 2.0 * rmax)
    r = sites + jit
    rad = rmin + (rmax - rmin) * uniform53(seed + 2, n)
    m = density * 4.0 / 3.0 * math.pi * rad ** 3
    cell = np.array([0.0, box[0], 0.0, box[1], 0.0, box[2]])
    return DEMSystem(r=r, m=m, rad=rad, fixed=np.zeros(n, dtype=np.uint8), cell=cell)


def dem_settle_config(n: int = 200_000, rmin: float = 0.9e-3, rmax: float = 1.1e-3,
                      box=(0.1, 0.1, 0.4), density: float = 7874.0, seed: int = 4,
                      vscale: float = 0.0, batch: int = 65536) -> DEMSystem:
    """Granular settle as the workload states it: n spheres with radii
    uniform in [rmin, rmax] at random non-overlapping positions inside the
    box (random sequential insertion: candidates with uniform centres in
    [rad, L - rad] per axis are taken in stream order, each kept when it
    overlaps no sphere kept before it).  200,000 spheres of ~1 mm in
    0.1 x 0.1 x 0.4 m fill ~21 % of the volume, below the insertion's
    jamming limit (~38 %).  vscale > 0 gives every sphere normal velocity
    components of that standard deviation (a granular gas); the workload
    itself starts at rest.  Masses are those of solid spheres of the given
    density; none is fixed (the walls are the box's reflecting faces)."""
    from scipy.spatial import cKDTree

    L = np.asarray(box, dtype=np.float64)
    pos = np.empty((0, 3))
    rad = np.empty(0)
    off = 0
    while rad.size < n:
        k = batch
        u = uniform53(seed, 4 * k, offset=4 * off).reshape(k, 4)
        off += k
        rc = rmin + (rmax - rmin) * u[:, 3]
        c = rc[:, None] + u[:, :3] * (L[None, :] - 2.0 * rc[:, None])
        ok = np.ones(k, dtype=bool)
        if rad.size:
            tree = cKDTree(pos)
            d, idx = tree.query(c, k=16, distance_upper_bound=rmax + rmax)
            hit = np.isfinite(d)
            idx = np.where(hit, idx, 0)
            ok &= ~np.any(hit & (d < rc[:, None] + rad[idx]), axis=1)
        # conflicts inside the batch: the later candidate of an overlapping
        # pair goes when the earlier one is still kept
        pairs = cKDTree(c).query_pairs(2.0 * rmax, output_type="ndarray")
        if pairs.size:
            pairs = np.sort(pairs, axis=1)
            dd = np.sqrt(((c[pairs[:, 0]] - c[pairs[:, 1]]) ** 2).sum(1))
            pairs = pairs[dd < rc[pairs[:, 0]] + rc[pairs[:, 1]]]
            pairs = pairs[np.lexsort((pairs[:, 0], pairs[:, 1]))]
            for a, b in pairs:
                if ok[a] and ok[b]:
                    ok[b] = False
        keep = np.nonzero(ok)[0][: n - rad.size]
        pos = np.concatenate([pos, c[keep]])
        rad = np.concatenate([rad, rc[keep]])
    m = density * 4.0 / 3.0 * math.pi * rad ** 3
    cell = np.array([0.0, box[0], 0.0, box[1], 0.0, box[2]])
    sysd = DEMSystem(r=pos, m=m, rad=rad, fixed=np.zeros(n, dtype=np.uint8), cell=cell)
    if vscale > 0:
        sysd.v = normals(seed + 3, 3 * n).reshape(n, 3) * vscale
    return sysd
