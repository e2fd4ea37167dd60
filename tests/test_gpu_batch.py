"""The batched entry point (mhdp_project_batch) over uneven droplet problems (different grid
sizes and hierarchy depths, brick and non-brick coarse levels): every problem's potential and
faces equal, bit for bit, those of its own single projection, and each meets the FP64 1e-8 bar
against the oracle's operator."""
import numpy as np
import pytest

from mhd_simulator_amd import configs as C
from mhd_simulator_amd.fluid import PressureError, PressureProjection, project_batch

pytestmark = pytest.mark.gpu


def problems():
    rng = C.SplitMix64(0x6D686404)
    out = []
    for k in range(7):
        side = (16, 24, 32, 40)[int(rng.uniform() * 4)]
        out.append(C.droplet_box(side, 777 + k))
    return out


@pytest.mark.parametrize("visits", [0, 2])
@pytest.mark.parametrize("mode", ["lockstep", "threads"])
def test_batch_matches_single_projections(orc, mode, visits, monkeypatch):
    """lockstep: one launch per kernel for all problems; threads: per-context streams.  visits 0:
    the automatic cycle (V on these small grids); 2: the level-1 W-cycle, whose extra kernels the
    lockstep merge interleaves with the other problems' launches."""
    import torch
    monkeypatch.setenv("MHDP_BATCH_MODE", mode)
    ps = problems()
    prs = [PressureProjection(p.grid, p.fluid, p.alpha) for p in ps]
    try:
        us = [torch.from_numpy(p.ustar).cuda() for p in ps]
        uo = [torch.zeros_like(u) for u in us]
        po = [torch.zeros(p.grid.cells, dtype=torch.float64, device="cuda") for p in ps]
        print(f"[gpu] batch ({mode}, visits {visits}) of {len(ps)}: sides {[p.grid.nx for p in ps]}", flush=True)
        res = project_batch(prs, us, precision="mixed", precond="mg", coarse_visits=visits, u_outs=uo, phi_outs=po)
        torch.cuda.synchronize()
        for p, pr, u, phi, (it, r64) in zip(ps, prs, uo, po, res):
            assert r64 <= 1e-8 and it > 0
            one_u = torch.zeros_like(u)
            one_phi = torch.zeros_like(phi)
            pr.project(torch.from_numpy(p.ustar).cuda(), precision="mixed", precond="mg", coarse_visits=visits,
                       u_out=one_u, phi_out=one_phi)
            torch.cuda.synchronize()
            assert pr.stats["iterations"] == it
            assert torch.equal(one_u, u) and torch.equal(one_phi, phi)
            true = orc.true_residual(p.grid, pr.mask, p.alpha, p.ustar, phi.cpu().numpy())
            assert true <= 1e-8
    finally:
        for pr in prs:
            pr.close()


def test_batch_rejects_repeated_context():
    import torch
    p = C.droplet_box(16, 5)
    with PressureProjection(p.grid, p.fluid, p.alpha) as pr:
        u = torch.from_numpy(p.ustar).cuda()
        with pytest.raises(PressureError):
            project_batch([pr, pr], [u, u])
