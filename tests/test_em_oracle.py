"""Pins of the EM Jacobi-CG oracle (oracle/em.inc: the edge curl-curl operator and the solve of
src/em/induction.wgsl + EmState::advance_to, src/em/mod.rs:863-1010).

  * the reference's own known answer: one backward-Euler step of the discrete diffusion
    eigenmode sin(pi y/ny) sin(pi z/nz) decays by 1 / (1 + lambda dt / (mu sigma h^2))
    (src/em/tests.rs:273-333, same grid, field and 2e-5 bound) and refines towards exp(-...);
  * the reference's air-stencil test: where the unit-reluctivity air fast path is used it equals
    the full C^T nu C + gauge operator (src/em/tests.rs:666-737: the three material cases, more
    than a quarter of the edges eligible, relative error < 2e-7 in FP32);
  * algebra: the operator is symmetric on interior edges, C^T nu C v is divergence-free at
    interior nodes (induction.wgsl:50), and the solve satisfies the step equation
    (C^T nu C + G + M) a = M (a_prev + dt E_mot - (As - As_prev)) + C^T (1 - nu) C As on every
    interior edge, checked with the oracle's own operator (pinned above) and numpy.
"""
import math

import numpy as np
import pytest

MU0 = 4e-7 * math.pi


def edge_grid(dims):
    """(axis, x, y, z) of every edge in the packed x|y|z order"""
    out = []
    for a in range(3):
        d = [dims[b] + (0 if b == a else 1) for b in range(3)]
        z, y, x = np.meshgrid(np.arange(d[2]), np.arange(d[1]), np.arange(d[0]), indexing="ij")
        out.append(np.stack([np.full(x.size, a), x.ravel(), y.ravel(), z.ravel()], 1))
    return np.concatenate(out)


def boundary_edges(dims):
    e = edge_grid(dims)
    out = np.zeros(len(e), bool)
    for a in range(3):
        sel = e[:, 0] == a
        for b in ((a + 1) % 3, (a + 2) % 3):
            c = e[sel, 1 + b]
            out[np.flatnonzero(sel)] |= (c == 0) | (c == dims[b])
    return out


def block_materials(dims, lo, hi, mid=1):
    m = np.zeros(int(np.prod(dims)), np.uint32)
    x = np.arange(m.size) % dims[0]
    y = (np.arange(m.size) // dims[0]) % dims[1]
    z = np.arange(m.size) // (dims[0] * dims[1])
    inside = (x >= lo[0]) & (x < hi[0]) & (y >= lo[1]) & (y < hi[1]) & (z >= lo[2]) & (z < hi[2])
    m[inside] = mid
    return m


def test_backward_euler_eigenmode_matches_reference_known_answer(orc):
    n = 16
    dims = (n, n, n)
    h = 16e-3 / n
    sigma = 1.33e6
    props = np.array([[0.0, 1.0], [sigma, 1.0]], np.float32)
    co = orc.em_coefficients(dims, h, np.ones(n ** 3, np.uint32), props)  # conductor everywhere
    e = edge_grid(dims)
    init = np.where(e[:, 0] == 0, 1e-5 * np.sin(np.pi * e[:, 2] / n) * np.sin(np.pi * e[:, 3] / n), 0)
    init = init.astype(np.float32)
    lam = 4 * math.sin(math.pi / (2 * n)) ** 2 * 2
    mu_sigma = MU0 * float(np.float32(sigma))
    total = 2e-5
    centre = (n // 2) + n * ((n // 2) + (n + 1) * (n // 2))
    continuous = math.exp(-lam / (mu_sigma * h * h) * total)
    previous_error = math.inf
    zero = np.zeros(len(e), np.float32)
    for steps in (1, 2, 4):
        dt = total / steps
        a = init.copy()
        for _ in range(steps):
            rc, a, st = orc.em_solve(dims, h, co, zero, zero, a.astype(np.float32), dt=dt, tol=1e-7, max_iter=2000)
            assert rc == 0
        ratio = float(a[centre]) / float(init[centre])
        expected = (1 / (1 + lam * dt / (mu_sigma * h * h))) ** steps
        assert abs(ratio - expected) < 2e-5, (ratio, expected)
        error = abs(ratio - continuous)
        assert error < previous_error
        previous_error = error


@pytest.mark.parametrize("sigma,mu", [(0.0, 1.0), (1.33e6, 1.0), (0.0, 2.0)])
def test_air_stencil_matches_full_operator(orc, sigma, mu):
    dims = (16, 16, 16)
    h = 1e-3
    props = np.array([[0.0, 1.0], [sigma, mu]], np.float32)
    mat = block_materials(dims, (5, 4, 6), (11, 12, 10))
    co = orc.em_coefficients(dims, h, mat, props)
    E = co["sigma"].size
    i = np.arange(E, dtype=np.float32)
    v = ((np.sin(i * np.float32(1.2345)) + np.cos(i * np.float32(0.345))) * np.float32(0.01)).astype(np.float32)
    fast = orc.em_apply(dims, h, co, v)
    eligible = int(np.count_nonzero(co["flags"] & 8))
    assert eligible > E // 4
    full = dict(co, flags=(co["flags"] & ~np.uint8(8)).astype(np.uint8))
    exact = orc.em_apply(dims, h, full, v)
    err = np.linalg.norm(fast.astype(np.float64) - exact) / np.linalg.norm(exact.astype(np.float64))
    assert err < 2e-7, err


def interior_case(orc, dims=(7, 6, 5), mu=3.0, sigma=2e6):
    h = 1e-3
    props = np.array([[0.0, 1.0], [sigma, mu]], np.float32)
    mat = block_materials(dims, (2, 2, 1), (5, 4, 4))
    return h, props, mat, orc.em_coefficients(dims, h, mat, props)


def test_operator_symmetric_on_interior_edges(orc):
    dims = (7, 6, 5)
    h, _, _, co = interior_case(orc, dims)
    bnd = boundary_edges(dims)
    rng = np.random.default_rng(3)
    u = np.where(bnd, 0.0, rng.normal(size=bnd.size))
    w = np.where(bnd, 0.0, rng.normal(size=bnd.size))
    au = orc.em_apply(dims, h, co, u, precision=64)
    aw = orc.em_apply(dims, h, co, w, precision=64)
    assert abs(au @ w - u @ aw) <= 1e-12 * abs(au @ w)
    assert u @ au > 0  # positive definite with the gauge and the conductor's mass term


def test_curl_adjoint_is_divergence_free(orc):
    # conductor everywhere: no air nodes, so (magnetostatic) apply is C^T nu C alone
    dims = (6, 5, 4)
    h = 1e-3
    props = np.array([[0.0, 1.0], [1e6, 2.5]], np.float32)
    mat = np.ones(int(np.prod(dims)), np.uint32)
    mat[::3] = 0  # reluctivity contrast between cells
    props[0] = (5e5, 1.0)
    co = orc.em_coefficients(dims, h, mat, props)
    assert not np.any(co["flags"] & 6)
    bnd = boundary_edges(dims)
    v = np.where(bnd, 0.0, np.random.default_rng(4).normal(size=bnd.size))
    q = orc.em_apply(dims, h, co, v, magnetostatic=True, precision=64)
    q[bnd] = 0.0
    e = edge_grid(dims)
    idx = {tuple(k): j for j, k in enumerate(map(tuple, e))}
    worst = 0.0
    for z in range(2, dims[2] - 1):
        for y in range(2, dims[1] - 1):
            for x in range(2, dims[0] - 1):
                p = (x, y, z)
                div = 0.0
                for a in range(3):
                    lo = list(p)
                    lo[a] -= 1
                    div += q[idx[(a, *p)]] - q[idx[(a, *lo)]]
                worst = max(worst, abs(div))
    assert worst < 1e-9 * np.abs(q).max()


@pytest.mark.parametrize("with_sources", [False, True])
def test_solve_satisfies_the_step_equation(orc, with_sources):
    dims = (7, 6, 5)
    h, props, mat, co = interior_case(orc, dims)
    E = co["sigma"].size
    bnd = boundary_edges(dims)
    rng = np.random.default_rng(7)
    a_prev = (rng.normal(size=E) * 1e-6).astype(np.float32)
    zero = np.zeros(E, np.float32)
    if with_sources:
        as_now = (rng.normal(size=E) * 1e-6).astype(np.float32)
        as_prev = (rng.normal(size=E) * 1e-6).astype(np.float32)
        mot = (rng.normal(size=E) * 1e-3).astype(np.float32)
    else:
        as_now, as_prev, mot = zero, zero, None
    dt = 1e-5
    rc, a, st = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, mot=mot, dt=dt, tol=1e-11, max_iter=5000)
    assert rc == 0 and st["rel_residual"] <= 1e-11
    # mass() (induction.wgsl:72) in FP64 from the FP32 parameters and edge conductivities
    f = lambda v: np.float64(np.float32(v))  # noqa: E731
    mass = f(MU0) * co["sigma"].astype(np.float64) * f(h) * f(h) / f(dt)
    lhs = orc.em_apply(dims, h, co, a, dt=dt, precision=64)
    rhs = mass * (a_prev.astype(np.float64) - (as_now.astype(np.float64) - as_prev))
    if mot is not None:
        rhs += mass * dt * mot
    unit = orc.em_coefficients(dims, h, mat, np.array([[0.0, 1.0], [props[1, 0], 1.0]], np.float32))
    assert np.array_equal(unit["flags"] & 6, co["flags"] & 6)
    contrast = (orc.em_apply(dims, h, unit, as_now, magnetostatic=True, precision=64) -
                orc.em_apply(dims, h, co, as_now, magnetostatic=True, precision=64))
    rhs += contrast
    inner = ~bnd
    err = np.linalg.norm((lhs - rhs)[inner]) / np.linalg.norm(rhs[inner])
    assert err < 1e-9, err
    assert np.array_equal(a[bnd], a_prev[bnd].astype(np.float64))  # held tangential potential
