"""The z-marching operator apply (pair columns, LDS planes) against the brick apply it replaces
on large grids, through the C ABI: the same projections with MHDP_APPLY_MARCH=1 (the z-march on
every eligible grid) in a second process give the same iteration counts, potentials and faces
equal to rounding (the per-cell operator is the brick pass's bit for bit; only the partial sums of
p.Ap run in another order), and the mixed solves meet the FP64 oracle's 1e-8.  Covers two-phase
and single-phase grids, periodic and walled axes, pinned anchors and grids of several tiles in x
and y with partial last tiles, in both precisions."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from mhd_simulator_amd import configs as C
from mhd_simulator_amd.fluid import FluidConfig, PressureProjection

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def fixture(dims, periodic, walls, reverse=False, seed=11):
    g, fluid, alpha = C.mg_fixture(dims, periodic, walls, reverse)
    return C.Problem(f"fixture{dims}", g, fluid, alpha, C.seeded_faces(g, seed))


CASES = {
    "plate-32": lambda: C.config5(32, droplets=6),
    "walls-pinned": lambda: fixture((16, 12, 8), (False,) * 3, True, True),
    "periodic": lambda: fixture((16, 12, 8), (True,) * 3, False),
    "single-phase-24": lambda: C.config1(24),
    "droplets-32": lambda: C.droplet_box(32, 99),
    "wide-136x68x20": lambda: fixture((136, 68, 20), (True, False, True), True),
}
RUNS = [("mixed", "auto"), ("fp32", "auto"), ("mixed", "pinned")]


def run_all():
    """Every case and run in this process: {case/run: (iterations, phi, faces)}."""
    out = {}
    cfg = FluidConfig(relative_tolerance=1e-4, absolute_divergence_tolerance=0.0, max_iterations=20000)
    for name, make in CASES.items():
        p = make()
        with PressureProjection(p.grid, p.fluid, p.alpha, cfg) as pr:
            for precision, gauge in RUNS:
                it, _ = pr.project(p.ustar, precision=precision, precond="mg", gauge=gauge)
                out[f"{name}/{precision}/{gauge}"] = (it, pr.phi.copy(), pr.velocity.copy())
    return out


def test_march_apply_equals_brick_apply(tmp_path, orc):
    runs = {}
    for mode in ("0", "1"):
        dst = tmp_path / f"m{mode}.npz"
        env = dict(os.environ, MHDP_APPLY_MARCH=mode)
        code = ("import sys, json, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r); "
                "import test_gpu_apply_march as T; r = T.run_all(); "
                "np.savez(%r, **{k + '|phi': v[1] for k, v in r.items()}, **{k + '|u': v[2] for k, v in r.items()}); "
                "print(json.dumps({k: v[0] for k, v in r.items()}))") % (ROOT, os.path.join(ROOT, "tests"), str(dst))
        print(f"[gpu] MHDP_APPLY_MARCH={mode} run in its own process", flush=True)
        res = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
        assert res.returncode == 0, res.stderr[-2000:]
        runs[mode] = (json.loads(res.stdout.strip().splitlines()[-1]), np.load(dst))
    (its0, v0), (its1, v1) = runs["0"], runs["1"]
    problems = {name: make() for name, make in CASES.items()}
    for key in its0:
        name = key.split("/")[0]
        phi0, phi1 = v0[key + "|phi"], v1[key + "|phi"]
        u0, u1 = v0[key + "|u"], v1[key + "|u"]
        dphi = float(np.abs(phi1 - phi0).max() / max(np.abs(phi0).max(), 1e-30))
        du = float(np.abs(u1 - u0).max() / max(np.abs(u0).max(), 1e-30))
        print(f"[gpu]   {key}: iterations brick {its0[key]} march {its1[key]}; phi {dphi:.1e} faces {du:.1e}",
              flush=True)
        assert its0[key] == its1[key], key
        assert dphi <= 1e-6 and du <= 1e-6, key
        if key.endswith("/mixed/auto"):
            p = problems[name]
            mask, _ = orc.anchors(p.fluid, p.grid.dims, p.grid.periodic)
            assert orc.true_residual(p.grid, mask, p.alpha, p.ustar, phi1) <= 1e-8, key
