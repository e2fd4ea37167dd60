"""Golden fixtures of the projection (tests/golden/*.npz, written by make_golden.py).

CPU: the oracle reproduces every fixture (integer and setup arithmetic bit-exact, the FP64
solution to 1e-10), and each fixture is pinned by facts that do not come from the oracle:
  * the FP64 potential equals a dense direct solve of the graph Laplacian assembled here,
    in numpy, from the fixture's face mobilities (gauge / anchor cells removed);
  * the projected faces are divergence-free on every non-anchor fluid cell (discrete
    divergence computed here from the face planes);
  * the diagonal is the sum of the cell's six face mobilities (1 on pinned anchors).
GPU: libmhdp reproduces the fixtures through the C ABI (mask, mobilities, diagonal and FP64
RHS bit-exact; the mixed projection's potential and faces against the FP64 solution).
"""
import glob
import os

import numpy as np
import pytest

from mhd_simulator_amd import configs as C

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
NAMES = [os.path.splitext(os.path.basename(f))[0] for f in FIXTURES]


def load(name):
    z = np.load(os.path.join(HERE, "golden", name + ".npz"))
    g = C.Grid(*[int(v) for v in z["dims"]], periodic=tuple(bool(v) for v in z["periodic"]), h=float(z["h"]),
               two_phase=bool(z["two_phase"]), rho_liquid=float(z["rho"][0]), rho_air=float(z["rho"][1]))
    alpha = z["alpha"] if z["alpha"].size else None
    return g, z, alpha


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def planes(g, faces):
    """face array -> three planes shaped (z, y, x) with the axis' extra entry."""
    out, off = [], 0
    for a in range(3):
        P = g.plane_dims(a)
        n = int(np.prod(P))
        out.append(np.asarray(faces[off:off + n]).reshape(P[2], P[1], P[0]))
        off += n
    return out


def lower_faces(g, beta_faces):
    """face planes -> cell-indexed lower-face values (3, cells)."""
    nx, ny, nz = g.dims
    return np.stack([p[:nz, :ny, :nx].ravel() for p in planes(g, beta_faces)])


def edges(g, beta_faces):
    """(cell, lower neighbour, mobility) for every face with positive mobility."""
    nx, ny, nz = g.dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = [x.ravel(), y.ravel(), z.ravel()]
    low = lower_faces(g, beta_faces)
    out = []
    for a in range(3):
        q = [c.copy() for c in pos]
        q[a] = q[a] - 1
        n = g.dims[a]
        inside = q[a] >= 0
        if g.periodic[a]:
            q[a] = np.mod(q[a], n)
            inside[:] = True
        b = low[a].astype(np.float64)
        assert np.all(b[~inside] == 0), "wall face with non-zero mobility"
        sel = inside & (b > 0)
        cell = np.flatnonzero(sel)
        nb = (q[0] + nx * (q[1] + ny * q[2]))[sel]
        out.append((cell, nb, b[sel]))
    return out


def divergence(g, faces):
    """sum over the six faces of (outflow - inflow), per cell, in FP64."""
    nx, ny, nz = g.dims
    u = [p.astype(np.float64) for p in planes(g, faces)]
    d = np.zeros((nz, ny, nx))
    for a, p in enumerate(u):
        ax = 2 - a
        n = g.dims[a]
        lo = np.take(p, np.arange(n), axis=ax)
        hi = np.take(p, (np.arange(n) + 1) % n if g.periodic[a] else np.arange(1, n + 1), axis=ax)
        d += hi - lo
    return d.ravel()


# ----------------------------------------------------------------------------- CPU


def test_fixtures_present():
    assert len(NAMES) >= 4


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(orc, name):
    g, z, alpha = load(name)
    mask, comps = orc.anchors(z["fluid"], g.dims, g.periodic)
    assert np.array_equal(mask, z["mask"]) and comps == int(z["components"])
    meanfree = comps == 1
    beta, diag = orc.coefficients(g, mask, alpha, meanfree)
    assert np.array_equal(beta, z["beta"]) and np.array_equal(diag, z["diag"])
    assert np.array_equal(orc.rhs(g, mask, z["ustar"], 64, zero_anchor=False), z["b64"])
    rc, phi, u, st = orc.project(g, mask, alpha, z["ustar"], precond=2, precision=64, tol64=1e-12, max_iter=20000)
    assert rc == 0
    assert relerr(phi, z["phi64"]) < 1e-10 and relerr(u, z["u64"]) < 1e-10
    rc32, phi32, _, st32 = orc.project(g, mask, alpha, z["ustar"], rel_tol=1e-5, abs_div_tol=0.0, precond=2,
                                       precision=32, max_iter=20000)
    assert rc32 == int(z["fp32_status"]) and st32["iterations"] == int(z["fp32_iterations"])
    assert relerr(phi32, z["phi32"]) < 1e-6


@pytest.mark.parametrize("name", NAMES)
def test_fixture_diagonal_is_sum_of_face_mobilities(name):
    g, z, _ = load(name)
    n = g.cells
    s = np.zeros(n)
    for cell, nb, b in edges(g, z["beta"]):
        np.add.at(s, cell, b)
        np.add.at(s, nb, b)
    mask = z["mask"]
    pinned = (int(z["components"]) != 1) & ((mask & 3) == 2)
    active = (mask != 0) & ~pinned
    assert np.allclose(z["diag"][active], s[active], rtol=2e-6, atol=0)
    assert np.all(z["diag"][pinned] == 1.0)


@pytest.mark.parametrize("name", [n for n in NAMES])
def test_fixture_potential_is_dense_direct_solution(name):
    g, z, _ = load(name)
    mask = z["mask"]
    fl = np.flatnonzero(mask != 0)
    if fl.size > 5000:
        pytest.skip("dense solve sized for small fixtures")
    gauge = (mask & 3) == 2  # one per component; the potential is zero there
    unknown = fl[~gauge[fl]]
    pos = -np.ones(g.cells, np.int64)
    pos[unknown] = np.arange(unknown.size)
    A = np.zeros((unknown.size, unknown.size))
    for cell, nb, b in edges(g, z["beta"]):
        for i, j in ((cell, nb), (nb, cell)):
            keep = pos[i] >= 0
            np.add.at(A, (pos[i][keep], pos[i][keep]), b[keep])
            both = keep & (pos[j] >= 0)
            np.add.at(A, (pos[i][both], pos[j][both]), -b[both])
    x = np.zeros(g.cells)
    x[unknown] = np.linalg.solve(A, z["b64"][unknown])
    assert np.all(z["phi64"][gauge] == 0)
    assert relerr(z["phi64"][fl], x[fl]) < 1e-9


@pytest.mark.parametrize("name", NAMES)
def test_fixture_projection_is_divergence_free(name):
    g, z, _ = load(name)
    mask = z["mask"]
    check = (mask != 0) & ~((int(z["components"]) != 1) & ((mask & 3) == 2))
    before = np.linalg.norm(z["b64"][check])
    after = np.linalg.norm(divergence(g, z["u64"])[check]) * g.h  # b = -h * (sum of face differences)
    assert before > 0 and after / before < 2e-6  # float32 faces


# ----------------------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_library_reproduces_fixture(name):
    from mhd_simulator_amd.fluid import PressureProjection
    g, z, alpha = load(name)
    print(f"[gpu] golden {name} dims={g.dims} periodic={g.periodic}", flush=True)
    with PressureProjection(g, z["fluid"], alpha) as pr:
        assert np.array_equal(pr.mask, z["mask"]) and pr.components == int(z["components"])
        meanfree = pr.components == 1
        beta, diag, _ = pr.coefficients(meanfree)
        assert np.array_equal(beta, lower_faces(g, z["beta"])) and np.array_equal(diag, z["diag"])
        assert np.array_equal(pr.rhs(z["ustar"], meanfree), z["b64"])
        pr.project(z["ustar"], precision="mixed", precond="mg", refine_tol=1e-8)
        assert pr.stats["rel_residual_f64"] <= 1e-8
        fl = pr.mask != 0
        assert relerr(pr.phi[fl], z["phi64"][fl]) < 1e-6
        assert relerr(pr.velocity, z["u64"]) < 1e-6
