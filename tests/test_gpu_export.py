"""Frame export of the GPU's FluidCell packets (mhd_simulator_amd/export.py; RunState::export_fields,
src/run/export.rs:9-60): after a projection and mhdp_fields_dev, the packets exported straight from
the device tensor read back bit for bit, with the grid's metadata."""
import numpy as np
import pytest

from mhd_simulator_amd import configs as C
from mhd_simulator_amd.export import export_frame, read_frame
from mhd_simulator_amd.fluid import PressureProjection

pytestmark = pytest.mark.gpu


def test_exported_fluid_packets_round_trip(tmp_path):
    import torch
    p = C.droplet_box(24, 99)
    g = p.grid
    with PressureProjection(g, p.fluid, p.alpha) as pr:
        u = torch.zeros(g.faces, dtype=torch.float32, device="cuda")
        phi = torch.zeros(g.cells, dtype=torch.float64, device="cuda")
        print(f"[gpu] export dims={g.dims}", flush=True)
        pr.project(torch.from_numpy(p.ustar).cuda(), precision="mixed", precond="mg", u_out=u, phi_out=phi)
        cells = torch.zeros(16 * g.cells, dtype=torch.float32, device="cuda")
        if p.alpha is not None:
            cells.view(g.cells, 16)[:, 14] = torch.from_numpy(p.alpha).cuda()
        dt = 1e-4
        pr.fields(u, phi, cells, dt=dt)
        torch.cuda.synchronize()
        frame = tmp_path / "frame-00000001"
        export_frame(str(frame), g.dims, g.h, fields={"fluid": cells}, step=1, time_s=dt,
                     metadata={"fluid": {"relative_tolerance": 1e-5}})
        meta, got = read_frame(str(frame))
        assert meta["dims"] == list(g.dims) and meta["step"] == 1
        want = cells.cpu().numpy().reshape(g.cells, 16)
        assert np.array_equal(got["fluid"].view(np.uint32), want.view(np.uint32))
