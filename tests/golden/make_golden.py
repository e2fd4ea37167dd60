"""Write the golden fixtures of the projection (inputs + the oracle's outputs).

    python tests/golden/make_golden.py

Each fixture is an .npz with the seeded inputs (grid, fluid mask, liquid fraction, face
velocities) and the CPU oracle's results: the encoded gauge mask, face mobilities, diagonal,
FP64 divergence RHS, the FP64 solution (PCG to 1e-12), its projected faces, and the FP32
reference-order solve (iterations, reported residual, potential).  The reference holds no
golden vectors for this path; these pin the oracle against regressions and give the GPU tests
fixed vectors.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mhd_simulator_amd import configs as C  # noqa: E402
from oracle import oracle as O  # noqa: E402


def cases():
    yield "shell16", C.config1(16)
    g, fluid, alpha = C.mg_fixture((12, 10, 6), (True, False, True), True, False)
    yield "odd_periodic_12x10x6", C.Problem("odd", g, fluid, alpha, C.seeded_faces(g, 4242))
    yield "droplets16", C.droplet_box(16, 99, droplets_lo=2, droplets_hi=2, r_lo=0.2, r_hi=0.3)
    g, fluid, alpha = C.mg_fixture((9, 7, 5), (False,) * 3, True, True)
    yield "disconnected_9x7x5", C.Problem("disc", g, fluid, alpha, C.seeded_faces(g, 7))


def make(name, p):
    g = p.grid
    mask, comps = O.anchors(p.fluid, g.dims, g.periodic)
    meanfree = comps == 1
    beta, diag = O.coefficients(g, mask, p.alpha, meanfree)
    b64 = O.rhs(g, mask, p.ustar, 64, zero_anchor=False)
    rc, phi64, u64, st64 = O.project(g, mask, p.alpha, p.ustar, precond=2, precision=64, tol64=1e-12, max_iter=20000)
    assert rc == 0, rc
    rc32, phi32, u32, st32 = O.project(g, mask, p.alpha, p.ustar, rel_tol=1e-5, abs_div_tol=0.0, precond=2,
                                       precision=32, max_iter=20000)
    out = dict(
        dims=np.array(g.dims, np.int64), periodic=np.array(g.periodic, bool), h=np.float32(g.h),
        two_phase=np.bool_(g.two_phase), rho=np.array([g.rho_liquid, g.rho_air], np.float32),
        fluid=p.fluid.astype(np.uint8), alpha=(p.alpha if p.alpha is not None else np.zeros(0, np.float32)),
        ustar=p.ustar, mask=mask, components=np.int64(comps), beta=beta, diag=diag, b64=b64,
        phi64=phi64, u64=u64, fp32_status=np.int64(rc32), fp32_iterations=np.int64(st32["iterations"]),
        fp32_residual=np.float64(st32["rel_residual"]), phi32=phi32)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(f"{name}: cells={g.cells} comps={comps} fp64 it={st64['iterations']} fp32 it={st32['iterations']} "
          f"rc32={rc32}")


def load(path):
    z = np.load(path)
    g = C.Grid(*[int(v) for v in z["dims"]], periodic=tuple(bool(v) for v in z["periodic"]), h=float(z["h"]),
               two_phase=bool(z["two_phase"]), rho_liquid=float(z["rho"][0]), rho_air=float(z["rho"][1]))
    alpha = z["alpha"] if z["alpha"].size else None
    return g, z, alpha


if __name__ == "__main__":
    O.build()
    for name, p in cases():
        make(name, p)
