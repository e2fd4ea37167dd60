"""mhdp_fields_noalloc_dev (the fields + diagnostics entry point that allocates no device memory)
against mhdp_fields_dev through the C ABI: packets bit for bit and diagnostics exactly, on two calls
in a row, the second on a smaller grid (the static scratch then holds the first call's larger
records), and the context's device bytes unchanged by the no-allocation calls."""
import numpy as np
import pytest

from mhd_simulator_amd import configs as C
from mhd_simulator_amd.fluid import PressureProjection

pytestmark = pytest.mark.gpu


def packets(p, seed):
    rng = np.random.default_rng(seed)
    cells = rng.uniform(-1, 1, size=(p.grid.cells, 16)).astype(np.float32)
    if p.grid.two_phase and p.alpha is not None:
        cells[:, 14] = p.alpha
    return cells


def test_fields_noalloc_matches_fields_on_two_calls():
    import torch
    for k, p in enumerate([C.droplet_box(48, 4321), C.droplet_box(24, 99)]):
        g = p.grid
        print(f"[gpu] no-alloc fields call {k + 1}: dims {g.dims}", flush=True)
        with PressureProjection(g, p.fluid, p.alpha) as pr:
            u = torch.zeros(g.faces, dtype=torch.float32, device="cuda")
            phi = torch.zeros(g.cells, dtype=torch.float64, device="cuda")
            pr.project(torch.from_numpy(p.ustar).cuda(), precision="mixed", precond="mg", u_out=u, phi_out=phi)
            cells0 = torch.from_numpy(packets(p, 7 + k)).cuda()
            em = torch.from_numpy(np.random.default_rng(3).normal(size=(g.cells, 16)).astype(np.float32)).cuda()
            kw = dict(dt=2e-3, acceleration=(0.0, 0.5, -9.81), mushy=True, em=em)
            before = pr.device_bytes
            a = cells0.clone()
            d_noalloc = pr.fields(u, phi, a, no_alloc=True, **kw)
            torch.cuda.synchronize()
            assert pr.device_bytes == before
            b = cells0.clone()
            d_ref = pr.fields(u, phi, b, **kw)
            torch.cuda.synchronize()
            assert torch.equal(a.view(torch.int32), b.view(torch.int32))
            assert d_noalloc == d_ref
            assert d_ref["max_speed"] > 0 and d_ref["nonfinite_cells"] == 0
