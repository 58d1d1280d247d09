"""Seeded synthetic scenes for the CK-MPM workloads.

Positions are generated cell by cell; every random draw comes from a counter-based
splitmix64 stream (``Splitmix``), so a scene is a pure function of its seed and is
identical on every machine and numpy version.

Workloads (sizes as stated by the benchmark configurations):

* ``colliding_spheres``  - two fixed-corotated spheres, r = 0.12, dx = 1/64, 16^3 blocks,
  8 particles per cell (2x2x2 stratified sub-cells with uniform jitter), v = (+-1, 0, 0).
  Shape follows tests/colliding_sphere/test_colliding_sphere.cuh:140-199 of the
  reference (lattice (i + 0.25 + 0.5 d) dx, one model per sphere).
* ``jelly_drop``         - Neo-Hookean cube of 50^3 cells x 8 ppc = 1,000,000 particles,
  dx = 1/256, 64^3 blocks, gravity -9.8.
* ``mixed_dam_sand``     - 5,000,000 weakly compressible fluid particles in a column whose
  x-extent is 0.25 of the domain plus a 3,000,000-particle Drucker-Prager sand cone,
  dx = 1/512, 128^3 blocks, per-cell counts log-uniform in [4, 16].
* ``elastic_bar``        - fixed-corotated bar 0.6 x 0.1 x 0.1, dx = 1/128, 8 ppc.
* ``batch_scenes``       - 64 independent scenes: grids 64^3 / 128^3 / 256^3 cells, particle
  counts log-uniform in [20k, 500k], material uniform over {FC, NH, DP, fluid}, a box or a
  sphere at 8 ppc, one velocity per scene ~ N(0, 0.5 m/s) per component.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

FIXED_COROTATED, NEO_HOOKEAN, DRUCKER_PRAGER, FLUID, VON_MISES, CAM_CLAY = 0, 1, 2, 3, 4, 5

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


class Splitmix:
    """Counter-based splitmix64: draw i of stream s is a hash of (s, i)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            i = np.arange(self.counter, self.counter + n, dtype=np.uint64)
            self.counter += n
            z = self.seed * np.uint64(0xD1342543DE82EF95) + (i + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        """n doubles uniform in [0, 1)."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def lame(E: float, nu: float):
    """mpm_material.cuh:276-283 (ComputeLameParameters)."""
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


@dataclasses.dataclass
class Material:
    kind: int
    lam: float = 0.0
    mu: float = 0.0
    bulk: float = 0.0
    gamma: float = 0.0
    viscosity: float = 0.0
    J0: float = 1.0
    friction_deg: float = 30.0
    yield_stress: float = 0.0       # von Mises
    alpha0: float = 0.0             # Cam-clay: initial logJp, beta, hardening factor, M, hardening
    beta: float = 0.0
    hardening_factor: float = 0.0
    m: float = 0.0
    hardening_on: bool = True

    def params(self) -> List[float]:
        return [self.lam, self.mu, self.bulk, self.gamma, self.viscosity, self.J0, self.friction_deg,
                self.yield_stress, self.alpha0, self.beta, self.hardening_factor, self.m, float(self.hardening_on)]


# grid boundary (ckmpm_set_boundary): walls and region kinds
WALLS_STICKY, WALLS_SLIP, WALLS_NONE = 0, 1, 2
REGION_FIXED, REGION_VELOCITY, REGION_ROTATION = 0, 1, 2


@dataclasses.dataclass
class Region:
    """Cells lo <= c < hi; v = 0 (fixed), value (velocity) or value x (c dx - center)."""
    lo: Sequence[int]
    hi: Sequence[int]
    kind: int
    value: Sequence[float] = (0.0, 0.0, 0.0)
    center: Sequence[float] = (0.0, 0.0, 0.0)


@dataclasses.dataclass
class Boundary:
    """The reference's ProcessGridCellVelocity scene hook as data: regions in order, then
    the walls (within 8 cells of the domain faces)."""
    walls: int = WALLS_STICKY
    regions: List[Region] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class ModelSpec:
    material: Material
    mass: float
    volume: float
    x: np.ndarray            # (n, 3) float32
    v: np.ndarray            # (1, 3) or (n, 3) float32
    lam_p: Optional[np.ndarray] = None   # per-particle Lame parameters (float32) or None
    mu_p: Optional[np.ndarray] = None
    rigid_count: int = 0                 # the first rigid_count particles are rigid
    rigid_velocity: tuple = (0.0, 0.0, 0.0)


@dataclasses.dataclass
class Scene:
    name: str
    dim: tuple
    offset: tuple
    dx: float
    gravity: float
    dt: float
    substeps: int
    models: List[ModelSpec]
    mass_clamp: float = 0.0
    capacity: int = 0        # block-list capacity (0: derived from the particle count)
    boundary: Optional[Boundary] = None  # None: the default sticky walls

    @property
    def n_particles(self) -> int:
        return int(sum(m.x.shape[0] for m in self.models))

    def block_capacity(self) -> int:
        if self.capacity:
            return self.capacity
        nblocks = int(np.prod(self.dim))
        # each particle block brings at most 27 exterior blocks; a body of n particles
        # at >= 4 ppc occupies at most n/4 cells = n/256 blocks
        est = 8 * (self.n_particles // 256 + 64) + 4096
        return int(min(nblocks, est))


def _stratified_cells(rng: Splitmix, cells: np.ndarray, dx: float, ppc_side: int = 2) -> np.ndarray:
    """ppc_side^3 particles per cell on stratified sub-cells with uniform jitter."""
    sub = np.stack(np.meshgrid(*[np.arange(ppc_side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    corner = cells[:, None, :] + sub[None, :, :] / ppc_side
    n = corner.shape[0] * corner.shape[1]
    jit = rng.uniform(3 * n).reshape(cells.shape[0], sub.shape[0], 3) / ppc_side
    return ((corner + jit) * dx).reshape(-1, 3)


def colliding_spheres(seed: int = 1, radius: float = 0.12, speed: float = 1.0) -> Scene:
    dx = 1.0 / 64
    E, nu, rho = 1e4, 0.3, 1000.0
    lam, mu = lame(E, nu)
    vol = dx ** 3 / 8
    rng = Splitmix(seed)
    models = []
    for cx, vx in ((0.3, speed), (0.7, -speed)):
        c = np.array([cx, 0.5, 0.5])
        lo = np.floor((c - radius) / dx).astype(int) - 1
        hi = np.ceil((c + radius) / dx).astype(int) + 1
        g = np.stack(np.meshgrid(*[np.arange(lo[a], hi[a]) for a in range(3)], indexing="ij"), -1).reshape(-1, 3)
        sub = np.stack(np.meshgrid(*[np.arange(2)] * 3, indexing="ij"), -1).reshape(-1, 3)
        # sub-cell centres (i + 0.25 + 0.5 d) dx, kept if inside the sphere
        centres = (g[:, None, :] + 0.25 + 0.5 * sub[None, :, :]).reshape(-1, 3) * dx
        keep = np.linalg.norm(centres - c, axis=1) <= radius
        centres = centres[keep]
        jit = (rng.uniform(3 * centres.shape[0]).reshape(-1, 3) - 0.5) * 0.5 * dx
        x = (centres + jit).astype(np.float32)
        v = np.array([[vx, 0.0, 0.0]], dtype=np.float32)
        models.append(ModelSpec(Material(FIXED_COROTATED, lam, mu), rho * vol, vol, x, v))
    return Scene("colliding_spheres", (16, 16, 16), (0, 0, 0), dx, 0.0, 1e-4, 100, models)


def jelly_drop(seed: int = 2, side_cells: int = 50) -> Scene:
    dx = 1.0 / 256
    E, nu, rho = 5e4, 0.4, 1000.0
    lam, mu = lame(E, nu)
    vol = dx ** 3 / 8
    rng = Splitmix(seed)
    lo = np.array([128 - side_cells // 2, 26, 128 - side_cells // 2])
    g = np.stack(np.meshgrid(*[np.arange(lo[a], lo[a] + side_cells) for a in range(3)], indexing="ij"), -1).reshape(-1, 3)
    x = _stratified_cells(rng, g.astype(np.float64), dx).astype(np.float32)
    v = np.zeros((1, 3), np.float32)
    models = [ModelSpec(Material(NEO_HOOKEAN, lam, mu), rho * vol, vol, x, v)]
    return Scene("jelly_drop", (64, 64, 64), (0, 0, 0), dx, -9.8, 5e-5, 500, models)


def _loguniform_fill(rng: Splitmix, cells: np.ndarray, dx: float, target: int) -> np.ndarray:
    """Per-cell count round(exp(U[ln 4, ln 16])), uniform positions, truncated at target."""
    counts = np.rint(np.exp(math.log(4.0) + rng.uniform(cells.shape[0]) * (math.log(16.0) - math.log(4.0)))).astype(np.int64)
    csum = np.cumsum(counts)
    if csum[-1] < target:
        raise ValueError("region too small for the requested particle count")
    last = int(np.searchsorted(csum, target))
    counts = counts[: last + 1].copy()
    counts[-1] -= int(csum[last] - target)
    cell_of = np.repeat(np.arange(last + 1), counts)
    pos = cells[cell_of].astype(np.float64) + rng.uniform(3 * target).reshape(-1, 3)
    return (pos * dx).astype(np.float32)


def mixed_dam_sand(seed: int = 3, n_fluid: int = 5_000_000, n_sand: int = 3_000_000, scale: int = 1) -> Scene:
    """scale > 1 shrinks the domain and counts by scale^3 (reduced copies for tests)."""
    dx = 1.0 / (512 // scale)
    ncell = 512 // scale
    rho = 1000.0
    vol = dx ** 3 / 8
    rng = Splitmix(seed)
    # fluid column: x in [0.0625, 0.3125] (0.25 of the domain), z in [0.45, 0.55]
    x0, x1 = int(0.0625 * ncell), int(0.3125 * ncell)
    z0, z1 = int(0.45 * ncell), int(0.55 * ncell)
    y0 = int(0.035 * ncell)
    per_col = (x1 - x0) * (z1 - z0)
    height = int(math.ceil(n_fluid / (per_col * 8.0))) + 2
    g = np.stack(np.meshgrid(np.arange(y0, y0 + height), np.arange(x0, x1), np.arange(z0, z1), indexing="ij"), -1).reshape(-1, 3)
    g = g[:, [1, 0, 2]]   # fill bottom layer first
    xf = _loguniform_fill(rng, g, dx, n_fluid)
    # sand cone centred at (0.68, *, 0.5), base radius r, height h on the floor
    r = 0.16 * ncell
    h = 0.13 * ncell
    cx, cz = 0.68 * ncell, 0.5 * ncell
    xs = np.arange(int(cx - r) - 1, int(cx + r) + 2)
    zs = np.arange(int(cz - r) - 1, int(cz + r) + 2)
    ys = np.arange(y0, y0 + int(h) + 2)
    gy, gx, gz = np.meshgrid(ys, xs, zs, indexing="ij")
    rad = np.sqrt((gx + 0.5 - cx) ** 2 + (gz + 0.5 - cz) ** 2)
    inside = rad <= r * (1.0 - (gy + 0.5 - y0) / h)
    gs = np.stack([gx[inside], gy[inside], gz[inside]], -1)
    xsand = _loguniform_fill(rng, gs, dx, n_sand)
    lam_s, mu_s = lame(3.5e5, 0.3)
    models = [
        ModelSpec(Material(FLUID, bulk=1e5, gamma=7.0, viscosity=0.1, J0=1.0), rho * vol, vol, xf, np.zeros((1, 3), np.float32)),
        ModelSpec(Material(DRUCKER_PRAGER, lam_s, mu_s, friction_deg=30.0), rho * vol, vol, xsand, np.zeros((1, 3), np.float32)),
    ]
    nb = 128 // scale
    return Scene("mixed_dam_sand", (nb, nb, nb), (0, 0, 0), dx, -9.8, 2e-5, 200, models)


def elastic_bar(seed: int = 5) -> Scene:
    dx = 1.0 / 128
    E, nu, rho = 1e4, 0.3, 1000.0
    lam, mu = lame(E, nu)
    vol = dx ** 3 / 8
    rng = Splitmix(seed)
    lo = np.array([int(0.2 / dx), int(0.45 / dx), int(0.45 / dx)])
    ext = np.array([int(round(0.6 / dx)), int(round(0.1 / dx)), int(round(0.1 / dx))])
    g = np.stack(np.meshgrid(*[np.arange(lo[a], lo[a] + ext[a]) for a in range(3)], indexing="ij"), -1).reshape(-1, 3)
    x = _stratified_cells(rng, g.astype(np.float64), dx).astype(np.float32)
    v = np.zeros((1, 3), np.float32)
    return Scene("elastic_bar", (32, 32, 32), (0, 0, 0), dx, 0.0, 1e-4, 300,
                 [ModelSpec(Material(FIXED_COROTATED, lam, mu), rho * vol, vol, x, v)])


def cantilever(seed: int = 6, scale: int = 1, log_e_sigma: float = 0.3) -> Scene:
    """The differentiation workload: a fixed-corotated bar 0.6 x 0.1 x 0.1 at 8 ppc on a
    128^3 grid (dx = 1/128), its left end inside the sticky 8-cell boundary band (so it
    is clamped there and bends under gravity), per-particle Young's modulus with
    log E ~ N(log 1e4, log_e_sigma), nu = 0.3, per-particle initial velocity (zero).
    scale > 1 divides the grid and the bar by scale per axis."""
    g = 128 // scale
    dx = 1.0 / g
    rho, nu = 1000.0, 0.3
    vol = dx ** 3 / 8
    rng = Splitmix(seed)
    # cells 0 and 1 bin below block 0, so the clamped end starts at cell 2 at least
    lo = np.array([max(2, int(round(0.03 * g))), int(round(0.45 * g)), int(round(0.45 * g))])
    ext = np.array([max(8, int(round(0.6 * g))), max(2, int(round(0.1 * g))), max(2, int(round(0.1 * g)))])
    cells = np.stack(np.meshgrid(*[np.arange(lo[a], lo[a] + ext[a]) for a in range(3)], indexing="ij"), -1).reshape(-1, 3)
    x = _stratified_cells(rng, cells.astype(np.float64), dx).astype(np.float32)
    logE = math.log(1e4) + log_e_sigma * rng.normal(x.shape[0])
    E = np.exp(logE)
    lam = (E * nu / ((1 + nu) * (1 - 2 * nu))).astype(np.float32)
    mu = (E / (2 * (1 + nu))).astype(np.float32)
    lam0, mu0 = lame(1e4, nu)
    v = np.zeros((x.shape[0], 3), np.float32)
    m = ModelSpec(Material(FIXED_COROTATED, lam0, mu0), rho * vol, vol, x, v, lam_p=lam, mu_p=mu)
    return Scene("cantilever", (g // 4, g // 4, g // 4), (0, 0, 0), dx, -9.8, 1e-4, 300, [m])


def _body_cells(shape: str, n_target: int, g: int) -> np.ndarray:
    """Cells of a centred box or sphere holding about n_target / 8 cells, kept inside
    the middle 70% of a g^3 grid (clear of the 8-cell sticky band)."""
    cells = n_target / 8.0
    limit = int(0.7 * g)
    c = g / 2.0
    if shape == "box":
        side = min(max(2, int(round(cells ** (1.0 / 3.0)))), limit)
        lo = int(c - side / 2)
        r = np.arange(lo, lo + side)
        return np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    rad = min(max(1.5, (3.0 * cells / (4.0 * math.pi)) ** (1.0 / 3.0)), limit / 2.0)
    r = np.arange(int(c - rad) - 1, int(c + rad) + 2)
    gx = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    keep = np.linalg.norm(gx + 0.5 - c, axis=1) <= rad
    return gx[keep]


def batch_scenes(count: int = 64, seed: int = 4, dt: float = 5e-5) -> List[Scene]:
    rho = 1000.0
    out = []
    for i in range(count):
        rng = Splitmix(seed * 1_000_003 + i)
        u = rng.uniform(5)
        g = (64, 128, 256)[min(2, int(u[0] * 3))]
        n_target = int(round(math.exp(math.log(20_000) + u[1] * (math.log(500_000) - math.log(20_000)))))
        kind = min(3, int(u[2] * 4))
        shape = "box" if u[3] < 0.5 else "sphere"
        dx = 1.0 / g
        cells = _body_cells(shape, n_target, g)
        x = _stratified_cells(rng, cells.astype(np.float64), dx).astype(np.float32)
        v = (0.5 * rng.normal(3)).astype(np.float32).reshape(1, 3)
        vol = dx ** 3 / 8
        if kind == FLUID:
            mat = Material(FLUID, bulk=1e5, gamma=7.0, viscosity=0.1, J0=1.0)
        elif kind == DRUCKER_PRAGER:
            lam, mu = lame(3.5e5, 0.3)
            mat = Material(DRUCKER_PRAGER, lam, mu, friction_deg=30.0)
        else:
            lam, mu = lame(5e4, 0.3)
            mat = Material(kind, lam, mu)
        nb = g // 4
        out.append(Scene(f"batch_{i:02d}", (nb, nb, nb), (0, 0, 0), dx, -9.8, dt, 200,
                         [ModelSpec(mat, rho * vol, vol, x, v)]))
    return out


SCENES = {
    "colliding_spheres": colliding_spheres,
    "jelly_drop": jelly_drop,
    "mixed_dam_sand": mixed_dam_sand,
    "elastic_bar": elastic_bar,
    "cantilever": cantilever,
}
