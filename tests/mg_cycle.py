"""numpy restatement of the library's multigrid preconditioner, for its tests.

The cycle of mhdp's fused multigrid path, in FP64: aggregation levels (2x2x2 children per coarse
cell, ceil-half dims), Galerkin coarse graphs as the oracle builds them (lower-slot edges, reaction,
diagonal), a pre-smoother from zero, the residual restricted by summing each aggregate's children,
the coarse correction scaled by ``scale`` and added to active cells, a post-smoother, the coarsest
level solved by 8 weighted-Jacobi sweeps, and level 1 optionally visited twice (the second visit
solves for the residual the first leaves; its result is added damped by ``theta``).

Smoothers per level:
  * "jacobi": weighted Jacobi with weight ``omega`` (the reference's multigrid.wgsl:66-73 with
    omega = 2/3);
  * "red-black": symmetric Gauss-Seidel in red-black order (red: x + y + z even) with the stored
    reciprocal diagonal: red then black before the coarse correction, black then red after it.
    A level uses it when every periodic axis has even length there (else Jacobi).

Fields are 3-D arrays indexed [z, y, x].
"""
from __future__ import annotations

import numpy as np


class Level:
    """One level's graph: lower-slot edge weights per axis (the edge to the cell's lower, wrapped
    neighbour; 0 for a closed face), reaction (0 on the fine level), reciprocal diagonal (0 on
    inactive cells)."""

    def __init__(self, dims, periodic, bx, by, bz, dinv, react=None):
        self.nx, self.ny, self.nz = dims
        self.periodic = tuple(bool(p) for p in periodic)
        shape = (self.nz, self.ny, self.nx)
        self.b = [np.asarray(v, np.float64).reshape(shape) for v in (bx, by, bz)]
        self.dinv = np.asarray(dinv, np.float64).reshape(shape)
        self.react = np.zeros(shape) if react is None else np.asarray(react, np.float64).reshape(shape)
        z, y, x = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx), indexing="ij")
        self.red = (x + y + z) % 2 == 0
        self.active = self.dinv != 0
        self.rb_ok = all(not self.periodic[a] or (self.dims[a] % 2 == 0) for a in range(3))

    @property
    def dims(self):
        return (self.nx, self.ny, self.nz)

    def _shift(self, v, axis, sign):
        """v at the neighbour in direction sign along axis (x=0, y=1, z=2); 0 beyond a closed side."""
        ax = 2 - axis
        if self.periodic[axis]:
            return np.roll(v, -sign, axis=ax)
        out = np.zeros_like(v)
        sl_dst = [slice(None)] * 3
        sl_src = [slice(None)] * 3
        if sign > 0:
            sl_dst[ax], sl_src[ax] = slice(0, -1), slice(1, None)
        else:
            sl_dst[ax], sl_src[ax] = slice(1, None), slice(0, -1)
        out[tuple(sl_dst)] = v[tuple(sl_src)]
        return out

    def faces(self, axis):
        """(lower-face weight, upper-face weight) of every cell along axis."""
        lo = self.b[axis]
        return lo, self._shift(lo, axis, +1)

    def offdiag(self, v):
        """sum over the six faces of weight * v(neighbour)."""
        out = np.zeros_like(v)
        for a in range(3):
            lo, up = self.faces(a)
            out += lo * self._shift(v, a, -1) + up * self._shift(v, a, +1)
        return out

    def apply(self, v):
        """A v = reaction v + sum_faces weight (v - v(neighbour))."""
        out = self.react * v
        for a in range(3):
            lo, up = self.faces(a)
            out += lo * (v - self._shift(v, a, -1)) + up * (v - self._shift(v, a, +1))
        return out


def restrict(fine: Level, coarse: Level, r):
    out = np.zeros((coarse.nz, coarse.ny, coarse.nx))
    np.add.at(out, (np.arange(fine.nz)[:, None, None] // 2, np.arange(fine.ny)[None, :, None] // 2,
                    np.arange(fine.nx)[None, None, :] // 2), r)
    return out


def prolong(fine: Level, zc):
    return zc[np.ix_(np.arange(fine.nz) // 2, np.arange(fine.ny) // 2, np.arange(fine.nx) // 2)]


class Cycle:
    def __init__(self, levels, smoother="red-black", omega=0.9, scale=1.8, visits=1, theta=0.8, sweeps=8):
        self.L = levels
        self.smoother, self.omega, self.scale, self.visits, self.theta, self.sweeps = (
            smoother, omega, scale, visits, theta, sweeps)

    def rb(self, l):
        return self.smoother == "red-black" and self.L[l].rb_ok

    # red-black Gauss-Seidel half-sweeps (old values replaced: unit weight)
    @staticmethod
    def gs_pre(lev: Level, r):
        u = lev.dinv * r
        z = np.where(lev.red, u, 0.0)
        return np.where(lev.red, z, lev.dinv * (r + lev.offdiag(z)))

    @staticmethod
    def gs_post(lev: Level, r, z):
        zb = np.where(lev.red, z, lev.dinv * (r + lev.offdiag(z)))
        return np.where(lev.red, lev.dinv * (r + lev.offdiag(zb)), zb)

    def _level(self, l, r):
        lev = self.L[l]
        if l == len(self.L) - 1:  # coarsest: weighted-Jacobi sweeps from zero
            rhs = np.where(lev.active, r, 0.0)
            z = np.zeros_like(r)
            for _ in range(self.sweeps):
                z = np.where(lev.active, z + self.omega * (rhs - lev.apply(z)) * lev.dinv, 0.0)
            return z
        nxt = self.L[l + 1]
        rb = self.rb(l)
        if rb:
            z1 = self.gs_pre(lev, r)
            # the smoother's diagonal is 1/dinv: the red residual is the flux of the black values
            res = np.where(lev.red & lev.active, lev.offdiag(np.where(lev.red, 0.0, z1)), 0.0)
        else:
            z1 = np.where(lev.active, self.omega * r * lev.dinv, 0.0)
            res = np.where(lev.active if l == 0 else True, r - lev.apply(z1), 0.0)
        rc = restrict(lev, nxt, res)
        zc = self._level(l + 1, rc)
        if l == 0 and self.visits >= 2 and len(self.L) >= 3:
            zc = zc + self.theta * self._level(1, rc - nxt.apply(zc))
        z2 = z1 + np.where(lev.active, self.scale * prolong(lev, zc), 0.0)
        if rb:
            return np.where(lev.active, self.gs_post(lev, r, z2), 0.0)
        return np.where(lev.active, z2 + self.omega * (r - lev.apply(z2)) * lev.dinv, 0.0)

    def __call__(self, r):
        f = self.L[0]
        return self._level(0, np.asarray(r, np.float64).reshape(f.nz, f.ny, f.nx)).ravel()


def hierarchy_from_oracle(orc, g, mask, alpha, meanfree):
    """Levels from the oracle's FP32 coefficients and Galerkin graphs (promoted to FP64)."""
    beta, diag = orc.coefficients(g, mask, alpha, meanfree)
    nx, ny, nz = g.dims
    lower, off = [], 0
    for a in range(3):
        P = g.plane_dims(a)
        plane = beta[off:off + int(np.prod(P))].reshape(P[2], P[1], P[0])
        lower.append(plane[:nz, :ny, :nx].astype(np.float64))
        off += int(np.prod(P))
    active = (mask != 0) if meanfree else ((mask & 3) == 1)
    dinv = np.where(active, 1.0 / np.where(active, diag.astype(np.float32), 1).astype(np.float32), 0.0)
    levels = [Level(g.dims, g.periodic, *lower, dinv)]
    mg = orc.MG(g, mask, alpha, meanfree, 32)
    for l in range(1, len(mg.levels)):
        c4, cdiag = mg.graph(l)
        dims = mg.levels[l][:3]
        cd = np.where(cdiag > 0, (np.float32(1) / cdiag.astype(np.float32)).astype(np.float64), 0.0)
        per = tuple(g.periodic)
        levels.append(Level(dims, per, c4[:, 0], c4[:, 1], c4[:, 2], cd, react=c4[:, 3]))
    return levels
