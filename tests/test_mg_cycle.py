"""CPU pins of the multigrid cycle restated in tests/mg_cycle.py (the reference the GPU's
red-black and Jacobi cycles are compared with).

* The red-black half-sweeps are Gauss-Seidel: checked against sparse triangular solves of the
  level operator with the stored reciprocal diagonal (scipy), in red-then-black order before the
  coarse correction and black-then-red after it.
* The whole cycle, assembled column by column on small grids, is symmetric and positive definite
  (red-black and the tuned Jacobi, one and two visits of level 1), as PCG needs; the eigenvalues of
  the level-1 cycle times its operator stay below (1 + theta) / theta = 2.25, the bound that keeps
  the damped second visit positive definite.
* With Jacobi, weight 2/3, scale 1 and one visit the restatement equals the oracle's V-cycle (the
  reference's, multigrid.wgsl:66-144) up to the FP32 rounding of the stored reciprocal diagonal.
* A planted mistake (the post-smoother's sweeps in the same order as the pre-smoother's) makes the
  cycle non-symmetric and fails the symmetry bar.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from mhd_simulator_amd import configs as C
import mg_cycle as M


def problem(name, n=12):
    if name == "plate":
        return C.config5(n, droplets=4)
    if name == "drops":
        return C.droplet_box(n, 4321)
    if name == "periodic":
        g, fluid, alpha = C.mg_fixture((n, n - 2, n - 4), (True,) * 3, False, False)
        return C.Problem("periodic", g, fluid, alpha, C.seeded_faces(g, 3))
    g, fluid, alpha = C.mg_fixture((n, n - 4, n - 2), (False,) * 3, True, True)
    return C.Problem("walls", g, fluid, alpha, C.seeded_faces(g, 3))


def levels_of(orc, p):
    mask, comps = orc.anchors(p.fluid, p.grid.dims, p.grid.periodic)
    return M.hierarchy_from_oracle(orc, p.grid, mask, p.alpha, comps == 1), mask, comps


def tilde_matrix(lev):
    """sparse 1/dinv - N on the active cells (N: the face weights), assembled from unit vectors."""
    n = lev.nx * lev.ny * lev.nz
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        e3 = e.reshape(lev.nz, lev.ny, lev.nx)
        cols.append(-lev.offdiag(e3).ravel())
    A = np.array(cols).T
    act = lev.active.ravel()
    d = np.where(act, 1.0 / np.where(act, lev.dinv.ravel(), 1.0), 0.0)
    A = A + np.diag(d)
    return sp.csr_matrix(A[np.ix_(act, act)]), act


@pytest.mark.parametrize("name", ["plate", "drops", "walls", "periodic"])
def test_red_black_half_sweeps_are_gauss_seidel(orc, name):
    lev = levels_of(orc, problem(name))[0][0]
    At, act = tilde_matrix(lev)
    red = lev.red.ravel()[act]
    rng = np.random.default_rng(5)
    r3 = np.zeros((lev.nz, lev.ny, lev.nx))
    r3[lev.active] = rng.normal(size=int(act.sum()))
    r = r3.ravel()[act]
    # forward sweep, red first, from zero: (D - L) z = r in the order red < black
    order = np.concatenate([np.flatnonzero(red), np.flatnonzero(~red)])
    P = At[order][:, order]
    z = np.zeros_like(r)
    z[order] = spla.spsolve_triangular(sp.tril(P).tocsr(), r[order], lower=True)
    got = M.Cycle.gs_pre(lev, r3).ravel()[act]
    assert np.allclose(got, z, rtol=1e-12, atol=1e-12 * np.abs(z).max())
    # backward sweep from z0, black first: z' = z0 + (D - L_bo)^-1 (r - A z0)
    z0_3 = np.zeros_like(r3)
    z0_3[lev.active] = rng.normal(size=int(act.sum()))
    z0 = z0_3.ravel()[act]
    order = np.concatenate([np.flatnonzero(~red), np.flatnonzero(red)])
    P = At[order][:, order]
    want = z0.copy()
    want[order] += spla.spsolve_triangular(sp.tril(P).tocsr(), (r - At @ z0)[order], lower=True)
    got = M.Cycle.gs_post(lev, r3, z0_3).ravel()[act]
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def assembled(cycle, lev0, active):
    n = lev0.nx * lev0.ny * lev0.nz
    idx = np.flatnonzero(active)
    Mx = np.zeros((len(idx), len(idx)))
    for k, i in enumerate(idx):
        e = np.zeros(n)
        e[i] = 1.0
        Mx[:, k] = cycle(e)[idx]
    return Mx


@pytest.mark.parametrize("smoother,scale", [("red-black", 1.8), ("red-black", 1.9), ("jacobi", 1.6)])
@pytest.mark.parametrize("visits", [1, 2])
@pytest.mark.parametrize("name", ["plate", "drops", "walls", "periodic"])
def test_cycle_is_symmetric_positive_definite(orc, name, smoother, scale, visits):
    levels, mask, comps = levels_of(orc, problem(name, 16))
    cyc = M.Cycle(levels, smoother=smoother, omega=0.9, scale=scale, visits=visits)
    Mx = assembled(cyc, levels[0], levels[0].active.ravel())
    asym = np.abs(Mx - Mx.T).max() / np.abs(Mx).max()
    ev = np.linalg.eigvalsh(0.5 * (Mx + Mx.T))
    print(f"{name} {smoother} visits={visits}: asym={asym:.1e} min eig={ev.min():.3e}")
    assert asym < 1e-12 and ev.min() > 0.0
    if visits == 2 and len(levels) >= 3:  # the level-1 cycle's spectrum against the W-visit bound
        sub = M.Cycle(levels[1:], smoother=smoother, omega=0.9, scale=scale, visits=1)
        l1 = levels[1]
        act1 = l1.active.ravel()
        M1 = assembled(sub, l1, act1)
        A1 = np.array([l1.apply(e.reshape(l1.nz, l1.ny, l1.nx)).ravel()
                       for e in np.eye(l1.nx * l1.ny * l1.nz)]).T[np.ix_(act1, act1)]
        lam = np.linalg.eigvals(M1 @ A1).real.max()
        assert lam < (1 + 0.8) / 0.8, lam


@pytest.mark.parametrize("name", ["plate", "walls"])
def test_reference_jacobi_cycle_equals_oracle_vcycle(orc, name):
    p = problem(name, 16)
    levels, mask, comps = levels_of(orc, p)
    cyc = M.Cycle(levels, smoother="jacobi", omega=2.0 / 3.0, scale=1.0, visits=1)
    mg = orc.MG(p.grid, mask, p.alpha, comps == 1, 64)
    active = levels[0].active.ravel()
    r = np.where(active, np.cos(np.arange(p.grid.cells) * 0.17), 0.0)
    a, b = cyc(r), mg.vcycle(r)
    # the restatement divides by the stored FP32 reciprocal diagonal, the oracle by the diagonal
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6


def test_planted_sweep_order_breaks_symmetry(orc):
    levels, _, _ = levels_of(orc, problem("drops", 12))

    class Planted(M.Cycle):
        @staticmethod
        def gs_post(lev, r, z):  # red first again: not the adjoint of the pre-smoother
            zr = np.where(lev.red, lev.dinv * (r + lev.offdiag(z)), z)
            return np.where(lev.red, zr, lev.dinv * (r + lev.offdiag(zr)))

    Mx = assembled(Planted(levels, visits=1), levels[0], levels[0].active.ravel())
    assert np.abs(Mx - Mx.T).max() / np.abs(Mx).max() > 1e-3
