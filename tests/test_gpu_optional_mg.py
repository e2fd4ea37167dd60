"""The optional multigrid preconditioner and the PCG breakdown guard, through the C ABI.

* Lazy, accounted, reused and possibly unavailable storage: the reference's
  ``optional_workspace_is_lazy_accounted_reused_and_can_be_unavailable``
  (src/fluid/multigrid/tests.rs:620-660; src/fluid/mod.rs:1031-1043).  A zero divergence never
  allocates it; the first nonzero residual does, once, and the context's device bytes grow by
  exactly its size; a refused allocation (forced by the test hook part-way through, or the cached
  refusal state) leaves the bytes where they were and the projection solves with Jacobi.
* Breakdown guard: a cycle that is not positive definite (forced by the test hook: a negative
  smoother weight) makes r.z <= 0; the projection notices on the device, solves again with the
  reference's V-cycle and reports it, single or batched.
"""
import numpy as np
import pytest

from mhd_simulator_amd import configs as C
from mhd_simulator_amd.fluid import FluidConfig, PressureProjection

pytestmark = pytest.mark.gpu


def fixture():
    # the reference test's fixture: [17, 13, 9], every axis periodic, no internal wall
    g, fluid, alpha = C.mg_fixture((17, 13, 9), (True,) * 3, False, False)
    return C.Problem("optional-mg", g, fluid, alpha, C.seeded_faces(g, 21))


def make(p):
    cfg = FluidConfig(relative_tolerance=1e-4, absolute_divergence_tolerance=0.0, max_iterations=20000)
    print(f"[gpu] optional multigrid {p.grid.dims}", flush=True)
    return PressureProjection(p.grid, p.fluid, p.alpha, cfg)


class Device:
    """Device-resident faces and outputs (the host entry point would add its staging buffers to the
    context's device bytes)."""

    def __init__(self, p):
        import torch
        self.torch = torch
        self.u = torch.from_numpy(p.ustar).cuda()
        self.zero = torch.zeros_like(self.u)
        self.u_out = torch.zeros_like(self.u)
        self.phi = torch.zeros(p.grid.cells, dtype=torch.float64, device="cuda")

    def project(self, pr, u, **kw):
        it, res = pr.project(u, u_out=self.u_out, phi_out=self.phi, **kw)
        self.torch.cuda.synchronize()
        return it, res


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_optional_workspace_is_lazy_accounted_reused_and_can_be_unavailable(orc, precision):
    p = fixture()
    with make(p) as pr:
        d = Device(p)
        base = pr.device_bytes
        assert pr.multigrid_state() == ("none", 0)
        # zero divergence never needs the optional allocation
        it, _ = d.project(pr, d.zero, precision=precision, precond="mg")
        assert it == 0 and pr.stats["multigrid"] == 0
        assert pr.multigrid_state() == ("none", 0) and pr.device_bytes == base
        # the first nonzero residual allocates it, accounted exactly
        d.project(pr, d.u, precision=precision, precond="mg")
        state, mg_bytes = pr.multigrid_state()
        assert state == "allocated" and mg_bytes > 0 and pr.stats["multigrid"] == 1
        assert pr.device_bytes == base + mg_bytes
        u1 = d.u_out.cpu().numpy()
        # reused: no further allocation
        d.project(pr, d.u, precision=precision, precond="mg")
        assert pr.device_bytes == base + mg_bytes and np.array_equal(d.u_out.cpu().numpy(), u1)
        # the cached refusal state: the projection still works with the original fine equation
        assert pr.multigrid_state(action=2) == ("refused", 0)
        assert pr.device_bytes == base
        it, res = d.project(pr, d.u, precision=precision, precond="mg")
        assert pr.stats["multigrid"] == 0 and pr.device_bytes == base
        if precision == "mixed":
            assert orc.true_residual(p.grid, pr.mask, p.alpha, p.ustar, d.phi.cpu().numpy()) <= 1e-8
        else:
            assert res <= 1e-4


@pytest.mark.parametrize("arrays", [0, 5])
@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_refused_allocation_falls_back_to_jacobi(orc, precision, arrays):
    """The allocation attempt fails after `arrays` of its arrays (as an exhausted device would): what
    it had allocated is freed, the attempt is not repeated, and the projection meets its bar with
    Jacobi."""
    p = fixture()
    with make(p) as pr:
        d = Device(p)
        base = pr.device_bytes
        pr.multigrid_state(action=1, arg=arrays)
        it, res = d.project(pr, d.u, precision=precision, precond="mg")
        assert pr.stats["multigrid"] == 0 and it > 0
        assert pr.multigrid_state() == ("refused", 0) and pr.device_bytes == base
        if precision == "mixed":
            assert res <= 1e-8
            assert orc.true_residual(p.grid, pr.mask, p.alpha, p.ustar, d.phi.cpu().numpy()) <= 1e-8
        else:
            assert res <= 1e-4
        d.project(pr, d.u, precision=precision, precond="mg")  # attempted once per context
        assert pr.multigrid_state() == ("refused", 0) and pr.device_bytes == base


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
@pytest.mark.parametrize("name", ["plate-64", "walls-pinned"])
def test_breakdown_is_detected_and_solved_with_reference_cycle(orc, precision, name):
    """A negative smoother weight makes the preconditioner negative definite: r.z < 0 on the first
    cycle.  The device guard stops the solve, the projection is solved again with the reference's
    V-cycle (weight 2/3, scale 1, one visit), equals that projection bit for bit and says so."""
    if name == "plate-64":
        p = C.config5(64, droplets=24)
    else:
        g, fluid, alpha = C.mg_fixture((16, 12, 8), (False,) * 3, True, True)
        p = C.Problem("walls-pinned", g, fluid, alpha, C.seeded_faces(g, 11))
    with make(p) as pr:
        ref_cycle = dict(smooth_weight=2.0 / 3.0, coarse_scale=1.0, coarse_visits=1, smoother="jacobi")
        pr.project(p.ustar, precision=precision, precond="mg", **ref_cycle)
        assert pr.stats["cycle_fallback"] == 0
        u_ref, phi_ref, it_ref = pr.velocity.copy(), pr.phi.copy(), pr.stats["iterations"]
        pr.next_cycle(-0.9, 1.6, 1)
        pr.project(p.ustar, precision=precision, precond="mg")
        assert pr.stats["cycle_fallback"] == 1 and pr.stats["multigrid"] == 1
        assert pr.stats["iterations"] == it_ref
        assert np.array_equal(pr.velocity, u_ref) and np.array_equal(pr.phi, phi_ref)
        if precision == "mixed":
            assert orc.true_residual(p.grid, pr.mask, p.alpha, p.ustar, pr.phi) <= 1e-8
        pr.project(p.ustar, precision=precision, precond="mg")  # the hook is spent
        assert pr.stats["cycle_fallback"] == 0


def tall_layers():
    """20x18x70 two-phase grid of alternating 3- and 4-cell liquid and air layers."""
    g = C.Grid(20, 18, 70, two_phase=True, rho_liquid=C.RHO_LIQUID, rho_air=C.RHO_AIR)
    alpha = ((np.arange(g.cells) // (g.nx * g.ny)) % 7 < 3).astype(np.float32)
    return C.Problem("tall", g, np.ones(g.cells, np.uint8), alpha, C.seeded_faces(g, 5))


def test_diverging_cycle_is_detected_and_solved_with_reference_cycle(orc):
    """A red-black cycle over-corrected to scale 1.9 on a tall layered grid with odd coarse levels
    keeps r.z > 0 but is not positive definite there: the inner residual grows or wanders without
    converging.  The device guard stops the solve once the residual exceeds 1e5 times the solve's
    right-hand side, the host once 100 iterations set no new least residual; the projection is
    solved again with the reference's V-cycle and meets the 1e-8 bar."""
    p = tall_layers()
    with make(p) as pr:
        print("[gpu] diverging red-black cycle on the tall grid", flush=True)
        it, res = pr.project(p.ustar, precision="mixed", precond="mg", smoother="red-black", coarse_scale=1.9,
                             refine_tol=1e-8)
        print(f"[gpu]   it={it} res64={res:.2e} fallback={pr.stats['cycle_fallback']}", flush=True)
        assert pr.stats["cycle_fallback"] == 1
        assert res <= 1e-8 and orc.true_residual(p.grid, pr.mask, p.alpha, p.ustar, pr.phi) <= 1e-8


def test_batch_breakdown_is_solved_again_with_reference_cycle(orc):
    import torch
    from mhd_simulator_amd.fluid import project_batch
    ps = [C.droplet_box(side, 300 + side) for side in (16, 24, 32)]
    prs = [make(p) for p in ps]
    try:
        us = [torch.from_numpy(p.ustar).cuda() for p in ps]
        uo = [torch.zeros_like(u) for u in us]
        po = [torch.zeros(p.grid.cells, dtype=torch.float64, device="cuda") for p in ps]
        prs[1].next_cycle(-0.9, 1.6, 1)
        res = project_batch(prs, us, precision="mixed", precond="mg", u_outs=uo, phi_outs=po)
        torch.cuda.synchronize()
        assert [pr.stats["cycle_fallback"] for pr in prs] == [0, 1, 0]
        for k, (p, pr) in enumerate(zip(ps, prs)):
            assert res[k][1] <= 1e-8
            assert orc.true_residual(p.grid, pr.mask, p.alpha, p.ustar, po[k].cpu().numpy()) <= 1e-8
        one_u, one_phi = torch.zeros_like(uo[1]), torch.zeros_like(po[1])
        prs[1].project(us[1], precision="mixed", precond="mg", smooth_weight=2.0 / 3.0, coarse_scale=1.0,
                       coarse_visits=1, smoother="jacobi", u_out=one_u, phi_out=one_phi)
        torch.cuda.synchronize()
        assert torch.equal(one_u, uo[1]) and torch.equal(one_phi, po[1])
    finally:
        for pr in prs:
            pr.close()


def test_batch_diverging_cycle_is_solved_again_with_reference_cycle(orc):
    """The tall layered grid between two droplet boxes in one batch with the over-corrected
    red-black cycle: only its solve is stopped (device guard or no new least residual in 100
    iterations) and solved again on its own with the reference's V-cycle; all three meet 1e-8."""
    import torch
    from mhd_simulator_amd.fluid import project_batch
    ps = [C.droplet_box(16, 316), tall_layers(), C.droplet_box(24, 324)]
    prs = [make(p) for p in ps]
    try:
        us = [torch.from_numpy(p.ustar).cuda() for p in ps]
        po = [torch.zeros(p.grid.cells, dtype=torch.float64, device="cuda") for p in ps]
        print("[gpu] batch with the diverging red-black cycle", flush=True)
        res = project_batch(prs, us, precision="mixed", precond="mg", smoother="red-black", coarse_scale=1.9,
                            phi_outs=po)
        torch.cuda.synchronize()
        print(f"[gpu]   fallback={[pr.stats['cycle_fallback'] for pr in prs]}", flush=True)
        assert [pr.stats["cycle_fallback"] for pr in prs] == [0, 1, 0]
        for k, (p, pr) in enumerate(zip(ps, prs)):
            assert res[k][1] <= 1e-8
            assert orc.true_residual(p.grid, pr.mask, p.alpha, p.ustar, po[k].cpu().numpy()) <= 1e-8
    finally:
        for pr in prs:
            pr.close()
