"""Launch shapes and the autotuner on the GPU (256^3 two-phase sphere + cylinder, where the
z-marching kernels run): the brick kernels' occupancy target changes register allocation only,
so the projection is bit-identical; the marching chunk length and block cap change the
reduction partials, and the projection still meets the FP64 1e-8 bar and agrees with the
default shapes' result; the tuner times every candidate and its winner reproduces."""
import numpy as np
import pytest

from mhd_simulator_amd import configs as C
from mhd_simulator_amd import tune
from mhd_simulator_amd.fluid import PressureProjection

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    import torch
    p = C.config2(256)
    pr = PressureProjection(p.grid, p.fluid, p.alpha)
    u = torch.from_numpy(p.ustar).cuda()
    yield pr, u
    pr.close()


@pytest.fixture
def restore():
    before = tune.current()
    yield before
    tune.set_launch(before)


def solve(pr, u, **kw):
    import torch
    uo = torch.empty_like(u)
    phi = torch.empty(pr.grid.cells, dtype=torch.float64, device="cuda")
    its, res = pr.project(u, precision="mixed", refine_tol=1e-8, u_out=uo, phi_out=phi, **kw)
    torch.cuda.synchronize()
    return its, pr.stats["rel_residual_f64"], uo.cpu().numpy(), phi.cpu().numpy()


def test_launch_shapes_keep_the_projection(problem, restore):
    pr, u = problem
    print("[gpu] tune: 256^3 default shapes", flush=True)
    its0, r0, u0, phi0 = solve(pr, u)
    assert r0 <= 1e-8
    tune.set_launch({"brick_occ": 3})
    print("[gpu] tune: brick occupancy 3", flush=True)
    its1, r1, u1, phi1 = solve(pr, u)
    assert its1 == its0 and np.array_equal(u1.view(np.uint32), u0.view(np.uint32))
    assert np.array_equal(phi1.view(np.uint64), phi0.view(np.uint64))
    for z, cap in ((16, 2048), (64, 1024)):  # z 64 leaves too few columns: the brick path
        tune.set_launch({"brick_occ": 2, "march_z": z, "march_cap": cap})
        print(f"[gpu] tune: march z {z} cap {cap}", flush=True)
        its2, r2, u2, phi2 = solve(pr, u)
        assert r2 <= 1e-8
        assert np.linalg.norm(u2.astype(np.float64) - u0) <= 1e-6 * np.linalg.norm(u0.astype(np.float64))
        assert abs(its2 - its0) <= 2


def test_tuner_winner_reproduces(problem, restore):
    pr, u = problem
    cands = tune.candidates(march_z=(16, 32), march_cap=(1024, 2048), brick_occ=(2,), coarse_visits=(1, 2))
    print(f"[gpu] tune: {len(cands)} candidates", flush=True)
    best = tune.tune_projection(pr, u, refine_tol=1e-8, repeats=1, cands=cands)
    assert len(best["table"]) == len(cands) and all(r["ok"] for r in best["table"])
    assert best["ms"] == min(r["ms"] for r in best["table"]) and best["ms"] > 0
    assert tune.current() == restore  # the search leaves the shapes as it found them
    opts = tune.apply(best)
    its, r, _, _ = solve(pr, u, **opts)
    assert r <= 1e-8


def test_two_stage_tuner_over_smoothers(problem, restore):
    """Schedules first (Jacobi and red-black at their tuned weights and scales), then launch shapes
    at the winner; every candidate converges with the cycle it asked for and the winner reproduces."""
    pr, u = problem
    sched = tune.schedules(smooth_weight=(0.0,), coarse_scale=(0.0,), coarse_visits=(0,))
    print(f"[gpu] tune: two stages, {len(sched)} schedules", flush=True)
    best = tune.tune_projection(pr, u, refine_tol=1e-8, repeats=1, schedule=sched,
                                launch_axes={"march_z": (32,), "march_cap": (1024, 2048), "brick_occ": (2,)})
    rows = best["table"]
    assert len(rows) == len(sched) + 2 and all(r["ok"] for r in rows)
    assert {r["opts"]["smoother"] for r in rows[:len(sched)]} == {"jacobi", "red-black"}
    assert all(r["opts"] == rows[len(sched)]["opts"] for r in rows[len(sched):])
    assert best["ms"] == min(r["ms"] for r in rows)
    assert tune.current() == restore
    opts = tune.apply(best)
    its, r, _, _ = solve(pr, u, **opts)
    assert r <= 1e-8
