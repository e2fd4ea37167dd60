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


def sourced_case(orc, seed=7, amp=1e-6):
    dims = (7, 6, 5)
    h, props, mat, co = interior_case(orc, dims)
    E = co["sigma"].size
    rng = np.random.default_rng(seed)
    a_prev = (rng.normal(size=E) * amp).astype(np.float32)
    as_now = (rng.normal(size=E) * amp).astype(np.float32)
    as_prev = (rng.normal(size=E) * amp).astype(np.float32)
    mot = (rng.normal(size=E) * amp * 1e3).astype(np.float32)
    return dims, h, co, a_prev, as_now, as_prev, mot


@pytest.mark.parametrize("amp", [1e-6, 1e-15])
def test_prediction_from_an_exact_history_starts_at_the_solution(orc, amp):
    """begin_step + prepare with a guess (induction.wgsl:116-124, 251-256): when prediction *
    (a_prev - a_before) is the step's exact increment, the start's residual b - A x0 is the exact
    solve's own residual plus FP32 rounding and the solve takes no iteration; the reversed history starts at residual 2 |b|
    (b - A(-x)) and is dropped for the zero start.  amp 1e-15 takes the power-of-two rescale
    (|b|_1 < 1e-8) with the guess in place."""
    dims, h, co, a_prev, as_now, as_prev, mot = sourced_case(orc, amp=amp)
    rc, a, st = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, mot=mot, tol=1e-12, max_iter=5000)
    assert rc == 0 and st["start_residual"] == 1.0 and st["prediction_dropped"] == 0
    assert (st["scale"] == 1.0) == (amp == 1e-6)
    x = a - a_prev
    for f in (1.0, 2.0, 0.5):
        before = (a_prev - x / f).astype(np.float32)
        for prec in (64, 32):
            rc, _, sp = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, mot=mot, tol=1e-5, max_iter=5000,
                                     a_before=before, prediction=f, precision=prec)
            assert rc == 0 and sp["iterations"] == 0 and sp["prediction_dropped"] == 0, (f, prec, sp)
            # the exact solve's own residual (its absolute floor at the tiny scale) + FP32 rounding
            assert sp["start_residual"] < st["rel_residual"] + 3e-7, (f, prec, sp)
            assert sp["scale"] == st["scale"]
        rc, _, sn = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, mot=mot, tol=1e-6, max_iter=5000,
                                 a_before=(a_prev + x / f).astype(np.float32), prediction=f)
        assert rc == 0 and sn["prediction_dropped"] == 1 and abs(sn["start_residual"] - 2.0) < 1e-6


@pytest.mark.parametrize("poison", [1e-3, 1e20])
def test_bad_history_falls_back_to_the_identical_zero_start(orc, poison):
    """The reference's predicted_em_bad_history_falls_back_to_identical_zero_guess_fields
    (src/em/prediction_tests.rs:170-188): a poisoned history is dropped and the step equals the
    zero start bit for bit."""
    dims, h, co, a_prev, as_now, as_prev, mot = sourced_case(orc)
    kw = dict(mot=mot, tol=1e-6, max_iter=5000, precision=32)
    rc0, a0, s0 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, **kw)
    rc1, a1, s1 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, a_before=np.full(a_prev.size, poison, np.float32),
                               prediction=1.0, **kw)
    assert rc0 == rc1 == 0 and s1["prediction_dropped"] == 1 and s1["start_residual"] > 1.0
    assert s1["iterations"] == s0["iterations"]
    assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))


def test_prediction_is_ignored_where_the_reference_does_not_predict(orc):
    """advance_to predicts only on a regular step with a positive finite ratio (mod.rs:869-886):
    the initial (magnetostatic) solve and factors 0, -1, nan start from zero."""
    dims, h, co, a_prev, as_now, as_prev, mot = sourced_case(orc)
    before = (a_prev * np.float32(0.5)).astype(np.float32)
    for ms in (False, True):
        kw = dict(mot=mot, tol=1e-6, max_iter=5000, precision=32, magnetostatic=ms)
        _, a0, s0 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, **kw)
        for f in ((0.0, -1.0, float("nan")) if not ms else (1.0,)):
            _, a1, s1 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, a_before=before, prediction=f, **kw)
            assert s1["start_residual"] == 1.0 and s1["prediction_dropped"] == 0
            assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32)), (ms, f)


def test_predicted_steps_track_timestep_changes(orc):
    """The reference's predicted_em_tracks_timestep_changes_with_the_original_rhs_tolerance
    (src/em/prediction_tests.rs:190-210): a sinusoidally driven conductor stepped with dt 1e-5,
    1e-5, 0.5e-5, 0.5e-5, 2e-5, 1e-5, the factor dt / previous dt; every predicted step meets the
    tolerance relative to its right-hand side and stays within 2e-4 (relative L2) of the zero-start
    steps, with fewer iterations in total."""
    dims = (7, 6, 5)
    h, props, mat, co = interior_case(orc, dims)
    E = co["sigma"].size
    rng = np.random.default_rng(5)
    shape = (rng.normal(size=E) * 1e-5).astype(np.float32)
    shape[boundary_edges(dims)] = 0.0
    drive = lambda t: (shape * np.float32(math.sin(2 * math.pi * 2000.0 * t))).astype(np.float32)  # noqa: E731
    tol = 1e-5
    plain = pred = before = np.zeros(E, np.float32)
    t, prev_dt, its_plain, its_pred = 0.0, 0.0, 0, 0
    for dt in (1e-5, 1e-5, 0.5e-5, 0.5e-5, 2e-5, 1e-5):
        as_prev, as_now = drive(t), drive(t + dt)
        f = dt / prev_dt if prev_dt > 0 else 0.0
        rc0, a0, s0 = orc.em_solve(dims, h, co, as_now, as_prev, plain, dt=dt, tol=tol, max_iter=5000, precision=32)
        rc1, a1, s1 = orc.em_solve(dims, h, co, as_now, as_prev, pred, dt=dt, tol=tol, max_iter=5000, precision=32,
                                   a_before=before, prediction=f)
        assert rc0 == rc1 == 0 and s1["rel_residual"] <= tol and s1["prediction_dropped"] == 0
        if f > 0:
            assert s1["start_residual"] < 1.0
        rel = np.linalg.norm(a1.astype(np.float64) - a0) / np.linalg.norm(a0.astype(np.float64))
        assert rel < 2e-4, (dt, rel)
        its_plain += s0["iterations"]
        its_pred += s1["iterations"]
        plain, before, pred = a0, pred, a1
        t += dt
        prev_dt = dt
    assert its_pred < its_plain, (its_pred, its_plain)
