"""GPU parity of the momentum step (mhdp_predict_dev, mhdp_viscous_dev through the C ABI) against
the oracle: both kernels are built without FP contraction and follow the oracle's FP32 operation
order, so every face must be equal bit for bit (predict with and without advection; viscous in
one phase with an EM force, in two phases with capillary forces), and the chain predict ->
viscous -> mixed projection meets the projection's 1e-8 bar against the oracle."""
import numpy as np
import pytest

from mhd_simulator_amd import configs as C
from mhd_simulator_amd.fluid import PressureError, PressureProjection

pytestmark = pytest.mark.gpu


def walls_pinned():
    g, fluid, alpha = C.mg_fixture((16, 12, 8), (False,) * 3, True, True)
    return C.Problem("walls-pinned", g, fluid, alpha, C.seeded_faces(g, 11))


def periodic_odd():
    g, fluid, alpha = C.mg_fixture((17, 13, 9), (True,) * 3, False, False)
    return C.Problem("periodic-odd", g, fluid, alpha, C.seeded_faces(g, 12))


CASES = {
    "config1-32": lambda: C.config1(32),
    "droplets-32": lambda: C.droplet_box(32, 1234),
    "walls-pinned": walls_pinned,
    "periodic-odd": periodic_odd,
}


def packets(p, seed):
    rng = np.random.default_rng(seed)
    cells = np.zeros((p.grid.cells, 16), np.float32)
    cells[:, 14] = p.alpha if p.alpha is not None else 1.0
    cells[:, 15] = rng.normal(scale=3.0, size=p.grid.cells).astype(np.float32)  # curvature
    return cells


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("name", list(CASES))
def test_predict_and_viscous_match_oracle(orc, name):
    import torch
    p = CASES[name]()
    g = p.grid
    dt = 0.4 * g.h  # departure points up to 0.4 cells away (|u| < 1)
    with PressureProjection(g, p.fluid, p.alpha) as pr:
        print(f"[gpu] step {name} dims={g.dims} two_phase={g.two_phase}", flush=True)
        u = torch.from_numpy(p.ustar).cuda()
        out = torch.empty_like(u)
        for advect in (False, True):
            pr.predict(u, out, dt=dt, advect=advect)
            want = orc.predict(g, pr.mask, p.ustar, dt=dt, advect=advect)
            assert np.array_equal(bits(out.cpu().numpy()), bits(want)), f"predict advect={advect}"
        pred = out.clone()
        pred_h = pred.cpu().numpy()
        cells = packets(p, 5)
        em = np.zeros((g.cells, 16), np.float32)
        em[:, 8:11] = np.random.default_rng(6).normal(size=(g.cells, 3)).astype(np.float32)
        kw = dict(dt=dt, viscosity=1e-3, acceleration=(0.25, -0.5, -9.81), surface_tension=0.07)
        ustar = torch.empty_like(u)
        for use_em in (False, True):
            pr.viscous(pred, ustar, cells=torch.from_numpy(cells).cuda(),
                       em=torch.from_numpy(em).cuda() if use_em else None, **kw)
            want = orc.viscous(g, pr.mask, pred_h, cells, em=em if use_em else None, **kw)
            assert np.array_equal(bits(ustar.cpu().numpy()), bits(want)), f"viscous em={use_em}"
        # the step's u* through the projection
        uo = torch.zeros_like(u)
        phi = torch.zeros(g.cells, dtype=torch.float64, device="cuda")
        pr.project(ustar, precision="mixed", precond="mg", refine_tol=1e-8, u_out=uo, phi_out=phi)
        assert orc.true_residual(g, pr.mask, p.alpha, ustar.cpu().numpy(), phi.cpu().numpy()) <= 1e-8


def test_in_place_predict_and_rejected_viscous_alias(orc):
    import torch
    p = C.droplet_box(16, 5)
    g = p.grid
    with PressureProjection(g, p.fluid, p.alpha) as pr:
        u = torch.from_numpy(p.ustar).cuda()
        pr.predict(u, u, dt=0.4 * g.h)  # predict reads its own open-face copy: in place is allowed
        want = orc.predict(g, pr.mask, p.ustar, dt=0.4 * g.h)
        assert np.array_equal(bits(u.cpu().numpy()), bits(want))
        with pytest.raises(PressureError):
            pr.viscous(u, u, dt=0.1, cells=torch.from_numpy(packets(p, 1)).cuda())
