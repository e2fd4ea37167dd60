"""Seeded synthetic inputs for the pressure projection.

Every input is reproducible bit-for-bit on any host: a SplitMix64 stream
(Steele, Lea & Flood 2014) drives all random draws, float32 uniforms in [-1, 1)
are ``((u >> 40) * 2**-24) * 2 - 1`` (exact in float32), and geometry is
evaluated at cell centres ``(i + 0.5) * h``.

Problem layout (shared with ``include/mhdp.h``):
  * cells ``x + nx*(y + ny*z)``; ``fluid`` is one uint8 per cell (1 = liquid or
    air, 0 = solid wall); ``alpha`` the liquid VOF fraction per cell (float32)
    or ``None`` for single-phase flow;
  * face velocities are three SoA planes packed x|y|z; plane ``a`` has dims
    ``n + e_a``.

The workloads follow the scope of the reference's projection
(src/fluid/mod.rs:1015 ``FluidState::project``): configs 1-5 are the
benchmark problems, the ``*_fixture`` helpers restate the procedural
fixtures of the reference's own pressure tests.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
BASE_SEED = 0x6D686400
VEL_TAG = 0xA5A5A5A5A5A5A5A5
SOLID_TAG = 0x5A5A5A5A5A5A5A5A
RHO_LIQUID = 7000.0
RHO_AIR = 1.2


@dataclasses.dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    periodic: tuple = (False, False, False)
    h: float = 1.0
    two_phase: bool = False
    rho_liquid: float = 1.0
    rho_air: float = 1.0

    @property
    def dims(self):
        return (self.nx, self.ny, self.nz)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    @property
    def periodic_bits(self):
        return sum(1 << a for a in range(3) if self.periodic[a])

    def plane_dims(self, axis):
        d = [self.nx, self.ny, self.nz]
        d[axis] += 1
        return tuple(d)

    def plane_sizes(self):
        return [int(np.prod(self.plane_dims(a))) for a in range(3)]

    @property
    def faces(self):
        return sum(self.plane_sizes())


@dataclasses.dataclass
class Problem:
    name: str
    grid: Grid
    fluid: np.ndarray            # uint8 [cells]
    alpha: np.ndarray | None     # float32 [cells]
    ustar: np.ndarray            # float32 [faces] packed planes
    rel_tol: float = 1e-6
    seed: int = 0


# ----------------------------------------------------------------- SplitMix64

def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


class SplitMix64:
    """Sequential SplitMix64 for scalar draws (radii, centres)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next(self) -> int:
        with np.errstate(over="ignore"):
            self.state = self.state + GOLDEN
            return int(_mix(np.array([self.state], dtype=np.uint64))[0])

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * 2.0 ** -53)

    def normal(self) -> float:
        u1 = max(self.uniform(), 2.0 ** -53)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def counter_stream(base: int, start: int, count: int) -> np.ndarray:
    """u_k = mix(base + (k+1)*golden) for k in [start, start+count)."""
    k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(np.uint64(base & 0xFFFFFFFFFFFFFFFF) + k * GOLDEN)


def uniform_pm1(base: int, count: int, chunk: int = 1 << 24) -> np.ndarray:
    """float32 uniforms in [-1, 1): ((u>>40) * 2^-24) * 2 - 1, exact in float32."""
    out = np.empty(count, dtype=np.float32)
    for s in range(0, count, chunk):
        n = min(chunk, count - s)
        u = counter_stream(base, s, n) >> np.uint64(40)
        out[s:s + n] = (u.astype(np.float64) * 2.0 ** -23 - 1.0).astype(np.float32)
    return out


def unit_uniform(base: int, count: int) -> np.ndarray:
    u = counter_stream(base, 0, count) >> np.uint64(40)
    return u.astype(np.float64) * 2.0 ** -24


def seeded_faces(grid: Grid, seed: int) -> np.ndarray:
    return uniform_pm1(seed ^ VEL_TAG, grid.faces)


# ----------------------------------------------------------------- geometry helpers

def cell_centres(grid: Grid):
    h = grid.h
    x = (np.arange(grid.nx, dtype=np.float64) + 0.5) * h
    y = (np.arange(grid.ny, dtype=np.float64) + 0.5) * h
    z = (np.arange(grid.nz, dtype=np.float64) + 0.5) * h
    return x, y, z


def tanh_fraction(dist_minus_r: np.ndarray, h: float, width: float = 1.5) -> np.ndarray:
    return 0.5 * (1.0 - np.tanh(dist_minus_r / (width * h)))


def add_droplet(alpha3: np.ndarray, grid: Grid, centre, radius: float, margin: float = 9.0):
    """alpha = max(alpha, tanh profile of one sphere), evaluated in its bounding box."""
    h = grid.h
    lo = [max(0, int(math.floor((centre[a] - radius) / h - margin))) for a in range(3)]
    hi = [min(grid.dims[a], int(math.ceil((centre[a] + radius) / h + margin)) + 1) for a in range(3)]
    if any(hi[a] <= lo[a] for a in range(3)):
        return
    z = (np.arange(lo[2], hi[2]) + 0.5) * h - centre[2]
    y = (np.arange(lo[1], hi[1]) + 0.5) * h - centre[1]
    x = (np.arange(lo[0], hi[0]) + 0.5) * h - centre[0]
    d = np.sqrt(z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2)
    frac = tanh_fraction(d - radius, h).astype(np.float32)
    view = alpha3[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    np.maximum(view, frac, out=view)


# ----------------------------------------------------------------- benchmark configs

def config1(n: int = 64, seed: int | None = None) -> Problem:
    """Single-phase 64^3, beta = 1, 4-cell solid shell, uniform [-1,1] faces, tol 1e-6."""
    seed = BASE_SEED + 1 if seed is None else seed
    g = Grid(n, n, n, two_phase=False, rho_liquid=1.0, rho_air=1.0)
    i = np.arange(n)
    inner = (i >= 4) & (i < n - 4)
    fluid = (inner[:, None, None] & inner[None, :, None] & inner[None, None, :]).astype(np.uint8).ravel()
    return Problem("config1-single-phase-shell", g, fluid, None, seeded_faces(g, seed), 1e-6, seed)


def config2(n: int = 128, seed: int | None = None) -> Problem:
    """Two-phase 128^3: seeded-radius sphere (tanh over 1.5 h), solid z-cylinder r=0.1L."""
    seed = BASE_SEED + 2 if seed is None else seed
    rng = SplitMix64(seed)
    g = Grid(n, n, n, two_phase=True, rho_liquid=RHO_LIQUID, rho_air=RHO_AIR)
    L = n * g.h
    r = rng.uniform(0.2, 0.35) * L
    alpha3 = np.zeros((n, n, n), dtype=np.float32)
    add_droplet(alpha3, g, (0.5 * L, 0.5 * L, 0.5 * L), r, margin=float(n))
    x, y, _ = cell_centres(g)
    cyl = ((x[None, :] - 0.5 * L) ** 2 + (y[:, None] - 0.5 * L) ** 2) < (0.1 * L) ** 2  # [y, x]
    solid = np.broadcast_to(cyl[None, :, :], (n, n, n))
    fluid = (~solid).astype(np.uint8).ravel()
    return Problem("config2-two-phase-sphere-cylinder", g, fluid, alpha3.ravel(), seeded_faces(g, seed), 1e-6, seed)


def nozzle_geometry(n: int):
    """Revolved hotend profile (scripts/make_fixtures.py:38-39) about the z axis at
    (0.5, 0.5)L: chamber OD 0.6L / ID 0.3L, tapering to a 0.1L outlet; liquid fills
    the chamber to 0.7875L.  The outlet (bottom, 0.1L) and the chamber top (0.85L) open into
    the surrounding air, so liquid and air form one connected region.
    Returns (solid, liquid) boolean [z, y, x]."""
    L = float(n)
    c = (np.arange(n) + 0.5)
    rad = np.sqrt((c[None, :] - 0.5 * L) ** 2 + (c[:, None] - 0.5 * L) ** 2)  # [y, x]
    # profile units: radius 1 unit = 0.1 L; height z_unit in [-2, 10] -> [0.1L, 0.85L]
    zu = (c - 0.1 * L) / (0.0625 * L) - 2.0  # [z]
    r_u = rad / (0.1 * L)
    zz = zu[:, None, None]
    rr = r_u[None, :, :]
    # inner (liquid side) and outer wall radii along the profile
    cone = np.clip((zz + 2.0) / 2.0, 0.0, 1.0)
    inner = np.where(zz >= 0.0, 1.5, 0.5 + (1.5 - 0.5) * cone)
    outer = np.where(zz >= 0.0, 3.0, 0.8 + (3.0 - 0.8) * cone)
    in_height = (zz >= -2.0) & (zz <= 10.0)
    wall = in_height & (rr >= inner) & (rr < outer)
    liquid = (zz >= -2.0) & (zz <= 9.0) & (rr < inner)
    idx = np.arange(n)
    shell1 = (idx < 2) | (idx >= n - 2)
    shell = shell1[:, None, None] | shell1[None, :, None] | shell1[None, None, :]
    solid = wall | shell
    return solid, liquid & ~solid


def config3(n: int = 256, seed: int | None = None) -> Problem:
    """Nozzle 256^3, liquid chamber in air at 7000/1.2, 2-cell outer shell, tol 1e-8."""
    seed = BASE_SEED + 3 if seed is None else seed
    g = Grid(n, n, n, two_phase=True, rho_liquid=RHO_LIQUID, rho_air=RHO_AIR)
    solid, liquid = nozzle_geometry(n)
    fluid = (~solid).astype(np.uint8).ravel()
    alpha = liquid.astype(np.float32).ravel()
    return Problem("config3-nozzle", g, fluid, alpha, seeded_faces(g, seed), 1e-8, seed)


BATCH_SIDES = (32, 48, 64, 96)


def droplet_box(n: int, seed: int, droplets_lo: int = 1, droplets_hi: int = 4,
                r_lo: float = 0.1, r_hi: float = 0.3, solid_fraction: float = 0.05) -> Problem:
    rng = SplitMix64(seed)
    g = Grid(n, n, n, two_phase=True, rho_liquid=RHO_LIQUID, rho_air=RHO_AIR)
    L = n * g.h
    count = droplets_lo + int(rng.uniform() * (droplets_hi - droplets_lo + 1))
    alpha3 = np.zeros((n, n, n), dtype=np.float32)
    for _ in range(count):
        r = rng.uniform(r_lo, r_hi) * L
        centre = tuple(rng.uniform(0.2, 0.8) * L for _ in range(3))
        add_droplet(alpha3, g, centre, r)
    solid = unit_uniform(seed ^ SOLID_TAG, g.cells) < solid_fraction
    fluid = (~solid).astype(np.uint8)
    return Problem(f"droplets-{n}", g, fluid, alpha3.ravel(), seeded_faces(g, seed), 1e-8, seed)


def config4(count: int = 64, seed: int | None = None) -> list[Problem]:
    """64 independent droplet problems, sides drawn from {32,48,64,96}."""
    seed = BASE_SEED + 4 if seed is None else seed
    rng = SplitMix64(seed)
    out = []
    for k in range(count):
        side = BATCH_SIDES[int(rng.uniform() * len(BATCH_SIDES))]
        out.append(droplet_box(side, seed * 1000003 + k))
    return out


def config5(n: int = 512, seed: int | None = None, droplets: int = 200, plate: int = 8) -> Problem:
    """512^3: 200 lognormal droplets (radius clip(exp(N(ln 10, 0.45)), 4, 24) cells)
    in air above an 8-cell solid plate, 7000/1.2, tol 1e-8."""
    seed = BASE_SEED + 5 if seed is None else seed
    rng = SplitMix64(seed)
    g = Grid(n, n, n, two_phase=True, rho_liquid=RHO_LIQUID, rho_air=RHO_AIR)
    scale = n / 512.0
    alpha3 = np.zeros((n, n, n), dtype=np.float32)
    for _ in range(droplets):
        r = min(max(math.exp(math.log(10.0) + 0.45 * rng.normal()), 4.0), 24.0) * scale
        cx = rng.uniform(r, n - r)
        cy = rng.uniform(r, n - r)
        cz = rng.uniform(plate * scale + r, n - r)
        add_droplet(alpha3, g, (cx, cy, cz), r)
    zc = np.arange(n)
    solid = np.broadcast_to((zc < plate * scale)[:, None, None], (n, n, n))
    fluid = (~solid).astype(np.uint8).ravel()
    return Problem("config5-droplets-plate", g, fluid, alpha3.ravel(), seeded_faces(g, seed), 1e-8, seed)


CONFIGS = {1: config1, 2: config2, 3: config3, 5: config5}


# ----------------------------------------------------------------- reference test fixtures

def mg_fixture(dims, periodic, walls: bool, reverse: bool):
    """src/fluid/multigrid/tests.rs:8-63: full-box liquid, optional internal wall at
    x = d0/2 plus a notch (x=1, y=1); alpha = (x < d0/2) xor reverse; 7000:1 with
    air density 1."""
    nx, ny, nz = dims
    g = Grid(nx, ny, nz, periodic=tuple(periodic), h=1.0, two_phase=True, rho_liquid=7000.0, rho_air=1.0)
    i = np.arange(g.cells)
    x = i % nx
    y = (i // nx) % ny
    fluid = ((~np.bool_(walls)) | ((x != nx // 2) & ~((x == 1) & (y == 1)))).astype(np.uint8)
    alpha = ((x < nx // 2) ^ bool(reverse)).astype(np.float32)
    return g, fluid, alpha


def manufactured_potential(grid: Grid, binary: bool = False):
    """src/fluid/multigrid/tests.rs:287-296 (f64 sum rounded to f32 per cell)."""
    nx, ny, nz = grid.dims

    def phi(q):
        x, y, z = q
        if binary:
            return np.minimum(x, nx - x).astype(np.float32)
        tau = 2.0 * math.pi
        return (np.cos(tau * (x + 0.5) / nx) + 0.3 * np.sin(tau * (y + 0.5) / ny)
                + 0.2 * np.cos(tau * (z + 0.5) / nz)).astype(np.float32)
    return phi


def manufactured_gradient(grid: Grid, fluid: np.ndarray, beta: np.ndarray, binary: bool = False) -> np.ndarray:
    """u*_a(q) = beta_a(q) (phi(q) - phi(q - e_a)) / h on open faces (float32 ops),
    src/fluid/multigrid/tests.rs:282-329."""
    phi = manufactured_potential(grid, binary)
    d = grid.dims
    out = np.zeros(grid.faces, dtype=np.float32)
    off = 0
    f3 = fluid.reshape(d[2], d[1], d[0]).astype(bool)
    for a in range(3):
        P = grid.plane_dims(a)
        zz, yy, xx = np.meshgrid(np.arange(P[2]), np.arange(P[1]), np.arange(P[0]), indexing="ij")
        q = [xx.ravel().astype(np.int64), yy.ravel().astype(np.int64), zz.ravel().astype(np.int64)]
        for b in range(3):
            if grid.periodic[b]:
                q[b] = np.mod(q[b], d[b])
        left = list(q)
        left[a] = q[a] - 1
        if grid.periodic[a]:
            left[a] = np.mod(left[a], d[a])

        def is_fluid(p):
            ok = (p[0] >= 0) & (p[0] < d[0]) & (p[1] >= 0) & (p[1] < d[1]) & (p[2] >= 0) & (p[2] < d[2])
            res = np.zeros(p[0].shape, dtype=bool)
            res[ok] = f3[p[2][ok], p[1][ok], p[0][ok]]
            return res
        open_ = is_fluid(q) & is_fluid(left)
        fi = q[0] + P[0] * (q[1] + P[1] * q[2])
        b = beta[off + fi]
        val = (b * (phi(q) - phi(left))) / np.float32(grid.h)
        out[off:off + P[0] * P[1] * P[2]] = np.where(open_, val, np.float32(0)).astype(np.float32)
        off += P[0] * P[1] * P[2]
    return out


def channel_mask(dims=(4, 8, 4)):
    """src/fluid/tests.rs:514: the channel with y-row 3 (of each 4-wide x-run) dropped."""
    n = int(np.prod(dims))
    i = np.arange(n)
    return (((i // 4) % 8) != 3).astype(np.uint8)


def channel_alpha(dims=(4, 8, 4)):
    """src/fluid/tests.rs:535-545: alpha = (i % 11) / 10."""
    n = int(np.prod(dims))
    i = np.arange(n)
    return ((i % 11).astype(np.float32) / np.float32(10.0)).astype(np.float32)
