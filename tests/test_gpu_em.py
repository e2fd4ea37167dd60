"""GPU parity of the EM Jacobi-CG (mhdp_em_* through the C ABI) against the oracle.

Bars: the per-edge coefficients, per-face reluctivities and the operator product are bit-exact
(built without FP contraction, the reference's FP32 operation order); the solve (FP32 vectors,
FP64 reductions) meets its relative tolerance on the actual equation, lies as close to the FP64
oracle's solution as the FP32 oracle (the reference's arithmetic) does - within 3x its distance -
and takes the FP32 oracle's iteration count within 5% + 2.
"""
import numpy as np
import pytest

from mhd_simulator_amd.em import EmSolver
from mhd_simulator_amd.fluid import PressureError

pytestmark = pytest.mark.gpu


def cell_xyz(dims):
    i = np.arange(int(np.prod(dims)))
    return i % dims[0], (i // dims[0]) % dims[1], i // (dims[0] * dims[1])


def case(name):
    if name == "block":
        dims = (12, 10, 8)
        x, y, z = cell_xyz(dims)
        m = (((x >= 3) & (x < 9) & (y >= 2) & (y < 7) & (z >= 2) & (z < 6))).astype(np.uint32)
        return dims, 1e-3, m, np.array([[0, 1], [2e6, 1]], np.float32), False
    if name == "permeable":
        dims = (14, 12, 10)
        x, y, z = cell_xyz(dims)
        m = np.zeros(x.size, np.uint32)
        m[(x >= 2) & (x < 7) & (y >= 3) & (y < 9) & (z >= 2) & (z < 8)] = 1
        m[(x >= 8) & (x < 12) & (y >= 2) & (y < 6) & (z >= 3) & (z < 7)] = 2
        return dims, 2e-3, m, np.array([[0, 1], [1e6, 3], [0, 50]], np.float32), False
    if name == "partial":
        dims = (12, 12, 12)
        x, y, z = cell_xyz(dims)
        r = np.sqrt((x - 5.5) ** 2 + (y - 5.5) ** 2 + (z - 5.5) ** 2)
        m = np.where(r < 3, 1, 0).astype(np.uint32)
        shell = (r >= 3) & (r < 4.2)
        frac = np.clip(4.2 - r, 0, 1)
        m[shell] = (0x80000000 | (np.round(frac[shell] * 8388607).astype(np.uint32) << 8) | 1).astype(np.uint32)
        return dims, 1e-3, m, np.array([[0, 1], [5e6, 2]], np.float32), False
    if name == "vacuum":
        dims = (10, 9, 8)
        return dims, 1e-3, np.zeros(int(np.prod(dims)), np.uint32), np.array([[0, 1]], np.float32), True
    raise KeyError(name)


CASES = ["block", "permeable", "partial", "vacuum"]


@pytest.mark.parametrize("name", CASES)
def test_em_coefficients_and_operator_bit_exact(orc, name):
    dims, h, m, props, vac = case(name)
    print(f"[gpu] em coefficients {name} dims={dims}", flush=True)
    with EmSolver(dims, h, m, props, vac) as em:
        got = em.coefficients()
        want = orc.em_coefficients(dims, h, m, props, vac)
        for k in ("sigma", "diag", "flags", "nu"):
            assert np.array_equal(got[k], want[k]), k
        rng = np.random.default_rng(2)
        v = rng.normal(size=em.edges).astype(np.float32) * np.float32(1e-3)
        for ms in (False, True):
            q = em.apply(v, dt=1e-5, magnetostatic=ms)
            w = orc.em_apply(dims, h, want, v, dt=1e-5, magnetostatic=ms, precision=32)
            assert np.array_equal(q.view(np.uint32), w.view(np.uint32)), f"magnetostatic={ms}"


@pytest.mark.parametrize("name", ["block", "permeable", "partial"])
@pytest.mark.parametrize("sources", [False, True])
def test_em_solve_matches_oracle(orc, name, sources):
    import torch
    dims, h, m, props, vac = case(name)
    tol = 1e-6
    with EmSolver(dims, h, m, props, vac) as em:
        E = em.edges
        rng = np.random.default_rng(11)
        a_prev = (rng.normal(size=E) * 1e-6).astype(np.float32)
        zero = np.zeros(E, np.float32)
        as_now, as_prev, mot = zero, zero, None
        if sources:
            as_now = (rng.normal(size=E) * 1e-6).astype(np.float32)
            as_prev = (rng.normal(size=E) * 1e-6).astype(np.float32)
            mot = (rng.normal(size=E) * 1e-3).astype(np.float32)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        out = torch.zeros(E, dtype=torch.float32, device="cuda")
        print(f"[gpu] em solve {name} sources={sources} dims={dims} edges={E}", flush=True)
        its, res = em.solve(dev(as_now), dev(as_prev), dev(a_prev), out, motional=None if mot is None else dev(mot),
                            dt=1e-5, rel_tol=tol, max_iter=5000)
        a_gpu = out.cpu().numpy().astype(np.float64)
        co = orc.em_coefficients(dims, h, m, props, vac)
        rc32, a32, st32 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, mot=mot, dt=1e-5, tol=tol, max_iter=5000,
                                       precision=32)
        rc64, a64, st64 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, mot=mot, dt=1e-5, tol=1e-12,
                                       max_iter=20000, precision=64)
        assert rc32 == 0 and rc64 == 0
        assert res <= tol
        assert abs(its - st32["iterations"]) <= 2 + 0.05 * st32["iterations"], (its, st32["iterations"])
        ref = np.linalg.norm(a32.astype(np.float64) - a64)
        assert np.linalg.norm(a_gpu - a64) <= 3 * ref + 1e-6 * np.linalg.norm(a64 - a_prev)


def test_em_not_converged_keeps_output(orc):
    import torch
    dims, h, m, props, vac = case("block")
    with EmSolver(dims, h, m, props, vac) as em:
        a_prev = torch.from_numpy((np.random.default_rng(1).normal(size=em.edges) * 1e-6).astype(np.float32)).cuda()
        zero = torch.zeros(em.edges, dtype=torch.float32, device="cuda")
        out = torch.full((em.edges,), 7.0, dtype=torch.float32, device="cuda")
        with pytest.raises(PressureError, match="EM solve did not converge: 3 iterations"):
            em.solve(zero, zero, a_prev, out, rel_tol=1e-7, max_iter=3)
        assert torch.all(out == 7.0)


def test_em_rejects_conductor_on_the_boundary():
    dims = (6, 6, 6)
    m = np.zeros(216, np.uint32)
    m[0] = 1
    with pytest.raises(PressureError, match="A conductor touches the EM outer boundary"):
        EmSolver(dims, 1e-3, m, np.array([[0, 1], [1e6, 1]], np.float32))


def test_em_kernel_timing_counts_working_launches():
    import torch
    dims, h, m, props, vac = case("block")
    with EmSolver(dims, h, m, props, vac) as em:
        E = em.edges
        rng = np.random.default_rng(5)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        zero = torch.zeros(E, dtype=torch.float32, device="cuda")
        a_prev = dev((rng.normal(size=E) * 1e-6).astype(np.float32))
        out = torch.zeros(E, dtype=torch.float32, device="cuda")
        em.set_kernel_timing("apply")
        its, _ = em.solve(zero, zero, a_prev, out, rel_tol=1e-6, max_iter=5000)
        ms, launches = em.kernel_time()
        # every iteration launches the kernel once; on a grid this small the launches queued after
        # convergence inside a batch take about as long as working ones and may be counted too
        assert its > 0 and ms > 0.0 and launches >= its
        em.set_kernel_timing(None)
        em.solve(zero, zero, a_prev, out, rel_tol=1e-6, max_iter=5000)
        assert em.kernel_time() == (0.0, 0)


def sourced(em, seed=11):
    rng = np.random.default_rng(seed)
    E = em.edges
    a_prev = (rng.normal(size=E) * 1e-6).astype(np.float32)
    as_now = (rng.normal(size=E) * 1e-6).astype(np.float32)
    as_prev = (rng.normal(size=E) * 1e-6).astype(np.float32)
    mot = (rng.normal(size=E) * 1e-3).astype(np.float32)
    return a_prev, as_now, as_prev, mot


@pytest.mark.parametrize("name", ["block", "permeable"])
@pytest.mark.parametrize("factor", [1.0, 2.0])
def test_em_predicted_solve_matches_oracle(orc, name, factor):
    """Increment prediction (mhdp_em_solve_predicted_dev): a history whose prediction is 0.9 of the
    step's increment starts near the solution on the GPU as in the oracle (same start residual up
    to FP64 summation order, not dropped), then meets the solve bars of the zero start; an exact
    history starts within FP32 rounding of the solution and takes no iteration."""
    import torch
    dims, h, m, props, vac = case(name)
    tol = 1e-6
    with EmSolver(dims, h, m, props, vac) as em:
        a_prev, as_now, as_prev, mot = sourced(em)
        co = orc.em_coefficients(dims, h, m, props, vac)
        kw = dict(mot=mot, dt=1e-5, max_iter=20000)
        rc64, a64, _ = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, tol=1e-12, precision=64, **kw)
        assert rc64 == 0
        x = a64 - a_prev
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        out = torch.zeros(em.edges, dtype=torch.float32, device="cuda")
        for share, bound in ((0.9, None), (1.0, 3e-7)):
            before = (a_prev - share * x / factor).astype(np.float32)
            print(f"[gpu] em predicted solve {name} factor={factor} share={share} dims={dims}", flush=True)
            its, res = em.solve(dev(as_now), dev(as_prev), dev(a_prev), out, motional=dev(mot), dt=1e-5, rel_tol=tol,
                                max_iter=20000, a_before=dev(before), prediction=factor)
            st = em.stats
            rc32, a32, st32 = orc.em_solve(dims, h, co, as_now, as_prev, a_prev, tol=tol, precision=32,
                                           a_before=before, prediction=factor, **kw)
            assert rc32 == 0 and res <= tol
            assert st["prediction"] == np.float32(factor) and st["prediction_dropped"] == st32["prediction_dropped"] == 0
            assert abs(st["start_residual"] - st32["start_residual"]) <= 1e-5 * st32["start_residual"] + 1e-9
            assert abs(its - st32["iterations"]) <= 2 + 0.05 * st32["iterations"], (its, st32["iterations"])
            a_gpu = out.cpu().numpy().astype(np.float64)
            ref = np.linalg.norm(a32.astype(np.float64) - a64)
            assert np.linalg.norm(a_gpu - a64) <= 3 * ref + 1e-6 * np.linalg.norm(x)
            if bound is not None:
                assert its == 0 and st["start_residual"] < bound


@pytest.mark.parametrize("poison", [1e-3, 1e20])
def test_em_bad_history_gives_the_zero_start_bit_for_bit(poison):
    """src/em/prediction_tests.rs:170-188 on the GPU: a poisoned history is dropped and the step
    equals the zero-start step bit for bit, with the same iteration count."""
    import torch
    dims, h, m, props, vac = case("block")
    with EmSolver(dims, h, m, props, vac) as em:
        a_prev, as_now, as_prev, mot = (torch.from_numpy(a).cuda() for a in sourced(em, 3))
        out0 = torch.zeros(em.edges, dtype=torch.float32, device="cuda")
        out1 = torch.zeros_like(out0)
        print(f"[gpu] em poisoned history {poison} dims={dims}", flush=True)
        its0, _ = em.solve(as_now, as_prev, a_prev, out0, motional=mot, rel_tol=1e-6, max_iter=5000)
        assert em.stats["start_residual"] == 1.0 and em.stats["prediction"] == 0.0
        poisoned = torch.full_like(out0, poison)
        its1, _ = em.solve(as_now, as_prev, a_prev, out1, motional=mot, rel_tol=1e-6, max_iter=5000,
                           a_before=poisoned, prediction=1.0)
        assert em.stats["prediction_dropped"] == 1 and em.stats["start_residual"] > 1.0
        assert its0 == its1
        assert torch.equal(out0.view(torch.int32), out1.view(torch.int32))
