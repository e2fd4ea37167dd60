"""Pins of the oracle's nodal auxiliary-space correction of the EM Jacobi-CG (oracle/em.inc,
restating src/em/nodal.wgsl and the switch of EmState::advance_to, src/em/mod.rs:887-960).

The correction is z = D^-1 r + G S^-1 G^T r with G the node-to-edge incidence gradient
((G phi)_e = phi(upper node) - phi(lower node), interior nodes only) and S = diag(G^T A G).  The
reference evaluates S in closed form (36 air(p) + the six incident edges' mass + the six neighbours'
air flags); here both are checked against the mathematics they stand for:

  * S equals g_k^T A g_k for every interior node k, with A the full FP64 operator applied by the
    oracle (its air-stencil fast path switched off) and g_k the column of G built here;
  * z equals r / d + G S^-1 G^T r assembled here with scipy (a flipped gradient sign, a dropped
    mass or air term, or a wrong node index fails one of the two);
  * the reference's own test: the preconditioner is symmetric positive (r^T B r > 0, normalized
    symmetry error < 2e-6 on the reference's two integer-pattern vectors,
    src/em/prediction_tests.rs:444-474);
  * the switch: a solve corrected from the start meets the same step equation as plain Jacobi and
    differs from its potential only by a gradient (the gauge; curl A, hence B, agrees: the
    reference compares fields, src/em/prediction_tests.rs:431-442), a short solve under the default
    policy is bit-identical to plain Jacobi (prediction_tests.rs:517-536), and the initial
    magnetostatic solve never switches.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from test_em_oracle import boundary_edges, edge_grid, interior_case, sourced_case

MU0 = 4e-7 * np.pi
DT = 1e-5


def incidence(dims):
    """G (edges x nodes, CSR) restricted to interior nodes, and the interior-node mask"""
    e = edge_grid(dims)
    NX, NY, NZ = dims[0] + 1, dims[1] + 1, dims[2] + 1
    lo = e[:, 1:4].copy()
    up = lo.copy()
    up[np.arange(len(e)), e[:, 0]] += 1
    node = lambda p: p[:, 0] + NX * (p[:, 1] + NY * p[:, 2])  # noqa: E731
    rows = np.concatenate([np.arange(len(e)), np.arange(len(e))])
    cols = np.concatenate([node(up), node(lo)])
    vals = np.concatenate([np.ones(len(e)), -np.ones(len(e))])
    G = sp.csr_matrix((vals, (rows, cols)), shape=(len(e), NX * NY * NZ))
    x, y, z = np.meshgrid(np.arange(NX), np.arange(NY), np.arange(NZ), indexing="ij")
    interior = np.zeros(NX * NY * NZ, bool)
    idx = (x + NX * (y + NY * z)).ravel()
    interior[idx] = ((x > 0) & (x < dims[0]) & (y > 0) & (y < dims[1]) & (z > 0) & (z < dims[2])).ravel()
    G = G[:, np.flatnonzero(interior)]
    return G, interior


def gauge_split(dims, h, co, a0, a1):
    """(part of a1 - a0 outside the gradients of interior nodes, |A (a1 - a0)|), relative to a0"""
    G, _ = incidence(dims)
    d = np.asarray(a1, np.float64) - np.asarray(a0, np.float64)
    phi = spl.lsqr(G, d, atol=1e-14, btol=1e-14)[0]
    full = dict(co, flags=(co["flags"] & ~np.uint8(8)).astype(np.uint8))
    A = lambda v: orc_apply(dims, h, full, v)  # noqa: E731
    return (np.linalg.norm(d - G @ phi) / np.linalg.norm(a0),
            np.linalg.norm(A(d)) / np.linalg.norm(A(np.asarray(a0, np.float64))))


def orc_apply(dims, h, co, v):
    from oracle import oracle as O
    return O.em_apply(dims, h, co, v, dt=DT, precision=64)


def edge_diagonal(co, h, dt=DT):
    """diagonal() (induction.wgsl:73) in FP64 from the FP32 parameters"""
    f = lambda v: np.float64(np.float32(v))  # noqa: E731
    mass = f(MU0) * co["sigma"].astype(np.float64) * f(h) * f(h) / f(dt)
    return np.maximum(co["diag"].astype(np.float64) + mass, 1e-30)


@pytest.fixture(scope="module")
def case(orc):
    dims = (7, 6, 5)
    h, props, mat, co = interior_case(orc, dims)
    G, interior = incidence(dims)
    full = dict(co, flags=(co["flags"] & ~np.uint8(8)).astype(np.uint8))
    S = np.empty(G.shape[1])
    for k in range(G.shape[1]):
        g = G[:, k].toarray().ravel()
        S[k] = g @ orc.em_apply(dims, h, full, g, dt=DT, precision=64)
    return dims, h, co, G, interior, S


def test_node_diagonal_is_diag_of_gradient_galerkin_operator(orc, case):
    dims, h, co, G, interior, S = case
    # the case exercises both terms: conducting edges (mass) and air nodes (gauge)
    assert np.count_nonzero(co["sigma"]) > 0 and np.count_nonzero(co["flags"] & 2) > 0
    _, Sor, _ = orc.em_nodal(dims, h, co, np.zeros(co["sigma"].size), dt=DT, precision=64)
    assert np.all(Sor[~interior] == 0)
    err = np.abs(Sor[interior] - S) / np.abs(S)
    assert err.max() < 1e-12, err.max()


@pytest.mark.parametrize("seed", [1, 2])
def test_correction_is_jacobi_plus_gradient_space_solve(orc, case, seed):
    dims, h, co, G, interior, S = case
    r = np.random.default_rng(seed).normal(size=co["sigma"].size)
    z, _, N = orc.em_nodal(dims, h, co, r, dt=DT, precision=64)
    want = r / edge_diagonal(co, h) + G @ ((G.T @ r) / S)
    err = np.linalg.norm(z - want) / np.linalg.norm(want)
    assert err < 1e-12, err
    assert np.all(N[~interior] == 0)
    # the FP32 restatement (the reference's arithmetic) agrees to FP32 rounding
    z32, _, _ = orc.em_nodal(dims, h, co, r.astype(np.float32), dt=DT, precision=32)
    assert np.linalg.norm(z32 - want) / np.linalg.norm(want) < 1e-6


def test_correction_is_symmetric_positive(orc, case):
    dims, h, co, *_ = case
    n = co["sigma"].size
    r = np.array([((i * 17 % 31) - 15) / 16 for i in range(n)], np.float32)
    s = np.array([((i * 29 % 37) - 18) / 19 for i in range(n)], np.float32)
    br = orc.em_nodal(dims, h, co, r, dt=DT, precision=32)[0].astype(np.float64)
    bs = orc.em_nodal(dims, h, co, s, dt=DT, precision=32)[0].astype(np.float64)
    rbr, sbs = r @ br, s @ bs
    assert rbr > 0 and sbs > 0
    assert abs(r @ bs - s @ br) / np.sqrt(rbr * sbs) < 2e-6


def test_corrected_solve_meets_the_same_step_equation(orc):
    dims, h, co, a_prev, as_now, as_prev, mot = sourced_case(orc)
    tol = 1e-9
    kw = dict(mot=mot, dt=DT, tol=tol, max_iter=5000)
    rc0, a0, s0 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, nodal=-1, **kw)
    rc1, a1, s1 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, nodal=1, nodal_after=0, **kw)
    assert rc0 == rc1 == 0 and s0["nodal"] == 0 and s1["nodal"] == 1
    assert s1["rel_residual"] <= tol
    print(f"iterations plain {s0['iterations']} corrected {s1['iterations']}")
    rest, res = gauge_split(dims, h, co, a0, a1)
    assert rest < 1e-7 and res < 1e-7, (rest, res)
    bnd = boundary_edges(dims)
    assert np.array_equal(a1[bnd], a_prev[bnd].astype(np.float64))


@pytest.mark.parametrize("precision,tol", [(32, 1e-6), (64, 1e-10)])
@pytest.mark.parametrize("after", [1, 40])
def test_switch_at_a_batch_boundary_keeps_the_target(orc, after, precision, tol):
    """The switch happens at the first 32-iteration batch boundary at or after `after`, at a
    true-residual restart; the solve then meets the unchanged target."""
    dims, h, co, a_prev, as_now, as_prev, mot = sourced_case(orc, seed=9)
    kw = dict(mot=mot, dt=DT, tol=tol, max_iter=5000, precision=precision)
    rc0, a0, s0 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, nodal=-1, **kw)
    rc1, a1, s1 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, nodal=1, nodal_after=after, **kw)
    assert rc0 == rc1 == 0 and s1["rel_residual"] <= tol
    switch_at = -(-after // 32) * 32
    assert s0["iterations"] > switch_at  # the case runs past the switch point
    assert s1["nodal"] == 1
    rest, res = gauge_split(dims, h, co, a0, a1)
    print(f"iterations plain {s0['iterations']} switched {s1['iterations']}: {rest:.1e} {res:.1e}")
    assert rest < 100 * tol and res < 100 * tol, (rest, res)


@pytest.mark.parametrize("precision", [32, 64])
def test_default_policy_keeps_short_solves_identical(orc, precision):
    dims, h, co, a_prev, as_now, as_prev, mot = sourced_case(orc)
    kw = dict(mot=mot, dt=DT, tol=1e-7, max_iter=5000, precision=precision)
    rc0, a0, s0 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, nodal=-1, **kw)
    rc1, a1, s1 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, nodal=0, **kw)
    assert rc0 == rc1 == 0 and s1["nodal"] == 0 and s0["iterations"] < 1024
    assert np.array_equal(a0, a1) and s0["iterations"] == s1["iterations"]


def test_magnetostatic_solve_never_switches(orc):
    dims, h, co, a_prev, as_now, as_prev, mot = sourced_case(orc)
    kw = dict(dt=DT, tol=1e-7, max_iter=5000, magnetostatic=True)
    rc0, a0, s0 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, nodal=-1, **kw)
    rc1, a1, s1 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, nodal=1, nodal_after=0, **kw)
    assert rc0 == rc1 == 0 and s1["nodal"] == 0
    assert np.array_equal(a0, a1)
