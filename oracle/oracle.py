"""ctypes wrapper of the C oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Imported only by tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg, always as the checker / CPU baseline, never as the product
path.  See mhdp_oracle.h for the reference file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class _Grid(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("periodic", C.c_uint32),
                ("h", C.c_float), ("two_phase", C.c_int), ("rho_liquid", C.c_float), ("rho_air", C.c_float)]


class _Opts(C.Structure):
    _fields_ = [("rel_tol", C.c_float), ("abs_div_tol", C.c_float), ("max_iter", C.c_uint32),
                ("precond", C.c_int), ("gauge", C.c_int), ("precision", C.c_int), ("tol64", C.c_double),
                ("threads", C.c_int)]


class _Stats(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("levels", C.c_uint32), ("components", C.c_uint32),
                ("fluid_cells", C.c_uint32), ("meanfree", C.c_int), ("multigrid", C.c_int),
                ("rel_residual", C.c_double), ("seconds", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_anchors.restype = C.c_uint32
        L.orc_anchors.argtypes = [P, P, C.c_uint32]
        L.orc_hierarchy.restype = C.c_int
        L.orc_hierarchy.argtypes = [P, P, C.c_int, P]
        L.orc_coefficients.argtypes = [P, P, P, C.c_int, P, P]
        L.orc_diagonal_uncached.argtypes = [P, P, P, C.c_int, P]
        for nm in ("orc_rhs_f32", "orc_rhs_f64"):
            getattr(L, nm).argtypes = [P, P, C.c_int, P, P]
        for nm in ("orc_apply_f32", "orc_apply_f64"):
            getattr(L, nm).argtypes = [P, P, C.c_int, P, P, P]
        for nm in ("orc_laplacian_f32", "orc_laplacian_f64"):
            getattr(L, nm).argtypes = [P, P, P, P, P]
        L.orc_true_residual.restype = C.c_double
        L.orc_true_residual.argtypes = [P, P, P, P, P]
        L.orc_correct_f64.argtypes = [P, P, P, P, P, P]
        L.orc_project.restype = C.c_int
        L.orc_project.argtypes = [P, P, P, P, P, P, P, P]
        L.orc_fields.argtypes = [P, P, P, P, P, P, P, P]
        L.orc_mg_new.restype = P
        L.orc_mg_new.argtypes = [P, P, P, C.c_int, C.c_int]
        L.orc_mg_free.argtypes = [P]
        L.orc_mg_levels.restype = C.c_int
        L.orc_mg_levels.argtypes = [P, P]
        L.orc_mg_action.argtypes = [P, C.c_int, P, P]
        L.orc_mg_vcycle.argtypes = [P, P, P]
        L.orc_mg_coarse_solve.argtypes = [P, P, P]
        L.orc_mg_coarse_sweeps.argtypes = [P, P, P, C.c_int]
        L.orc_mg_graph.argtypes = [P, C.c_int, P, P]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _grid(g) -> _Grid:
    return _Grid(g.nx, g.ny, g.nz, g.periodic_bits, g.h, int(g.two_phase), g.rho_liquid, g.rho_air)


def anchors(fluid: np.ndarray, dims, periodic=(False, False, False)):
    """src/fluid/mod.rs:1447-1493 -> (encoded uint32 mask, components)."""
    mask = np.ascontiguousarray((np.asarray(fluid) != 0).astype(np.uint32))
    d = np.array(dims, dtype=np.uint32)
    bits = sum(1 << a for a in range(3) if periodic[a])
    comps = lib().orc_anchors(_ptr(mask), _ptr(d), bits)
    return mask, int(comps)


def hierarchy(dims4):
    """src/fluid/multigrid.rs:25-65 -> (list of level dims (nx,ny,nz,count), coarse cells) or None."""
    d = np.array(dims4, dtype=np.uint32)
    out = np.zeros(64 * 4, dtype=np.uint32)
    coarse = np.zeros(1, dtype=np.uint64)
    n = lib().orc_hierarchy(_ptr(d), _ptr(out), 64, _ptr(coarse))
    if n < 0:
        return None
    return [tuple(int(v) for v in out[4 * l:4 * l + 4]) for l in range(n)], int(coarse[0])


def coefficients(g, mask, alpha, meanfree: bool):
    beta = np.zeros(g.faces, dtype=np.float32)
    diag = np.zeros(g.cells, dtype=np.float32)
    gg = _grid(g)
    lib().orc_coefficients(C.byref(gg), _ptr(mask), _ptr(alpha), int(meanfree), _ptr(beta), _ptr(diag))
    return beta, diag


def diagonal_uncached(g, mask, alpha, meanfree: bool):
    diag = np.zeros(g.cells, dtype=np.float32)
    gg = _grid(g)
    lib().orc_diagonal_uncached(C.byref(gg), _ptr(mask), _ptr(alpha), int(meanfree), _ptr(diag))
    return diag


def rhs(g, mask, ustar, precision=64, zero_anchor=True):
    dt = np.float64 if precision == 64 else np.float32
    b = np.zeros(g.cells, dtype=dt)
    gg = _grid(g)
    fn = lib().orc_rhs_f64 if precision == 64 else lib().orc_rhs_f32
    fn(C.byref(gg), _ptr(mask), int(zero_anchor), _ptr(np.ascontiguousarray(ustar, dtype=np.float32)), _ptr(b))
    return b


def apply(g, mask, meanfree, beta, v, precision=64):
    dt = np.float64 if precision == 64 else np.float32
    v = np.ascontiguousarray(v, dtype=dt)
    out = np.zeros(g.cells, dtype=dt)
    gg = _grid(g)
    fn = lib().orc_apply_f64 if precision == 64 else lib().orc_apply_f32
    fn(C.byref(gg), _ptr(mask), int(meanfree), _ptr(beta), _ptr(v), _ptr(out))
    return out


def laplacian(g, mask, alpha, v, precision=64):
    dt = np.float64 if precision == 64 else np.float32
    v = np.ascontiguousarray(v, dtype=dt)
    out = np.zeros(g.cells, dtype=dt)
    gg = _grid(g)
    fn = lib().orc_laplacian_f64 if precision == 64 else lib().orc_laplacian_f32
    fn(C.byref(gg), _ptr(mask), _ptr(alpha), _ptr(v), _ptr(out))
    return out


def true_residual(g, mask, alpha, ustar, phi) -> float:
    gg = _grid(g)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    return float(lib().orc_true_residual(C.byref(gg), _ptr(mask), _ptr(alpha),
                                         _ptr(np.ascontiguousarray(ustar, dtype=np.float32)), _ptr(phi)))


def correct(g, mask, alpha, ustar, phi):
    gg = _grid(g)
    out = np.zeros(g.faces, dtype=np.float32)
    lib().orc_correct_f64(C.byref(gg), _ptr(mask), _ptr(alpha), _ptr(np.ascontiguousarray(ustar, dtype=np.float32)),
                          _ptr(np.ascontiguousarray(phi, dtype=np.float64)), _ptr(out))
    return out


class _FieldsParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("acceleration", C.c_float * 3), ("mushy", C.c_int), ("meanfree", C.c_int)]


def fields(g, mask, faces, phi, cells, dt=1e-3, acceleration=(0.0, 0.0, -9.81), mushy=False, meanfree=None, em=None):
    """fields() + diagnostics() (solver.wgsl:256-291): the 16-float FluidCell packet per cell
    (cells: reserved lanes read, all overwritten; returned as a copy) and the 10 totals."""
    gg = _grid(g)
    if meanfree is None:
        meanfree = int(np.count_nonzero((mask & 3) == 2)) == 1
    prm = _FieldsParams(dt, (C.c_float * 3)(*acceleration), int(bool(mushy)), int(bool(meanfree)))
    out = np.ascontiguousarray(np.array(cells, dtype=np.float32).reshape(-1))
    assert out.size == 16 * g.cells
    totals = np.zeros(10, dtype=np.float64)
    lib().orc_fields(C.byref(gg), _ptr(mask), _ptr(np.ascontiguousarray(faces, dtype=np.float32)),
                     _ptr(np.ascontiguousarray(phi, dtype=np.float64)),
                     None if em is None else _ptr(np.ascontiguousarray(em, dtype=np.float32).reshape(-1)),
                     C.byref(prm), _ptr(out), _ptr(totals))
    return out.reshape(g.cells, 16), totals


class _StepParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("viscosity", C.c_float), ("acceleration", C.c_float * 3),
                ("surface_tension", C.c_float), ("advect", C.c_int)]


def _step_params(dt, viscosity, acceleration, surface_tension, advect):
    return _StepParams(dt, viscosity, (C.c_float * 3)(*acceleration), surface_tension, int(bool(advect)))


def predict(g, mask, faces, dt=1e-3, advect=True):
    """predict (solver.wgsl:127-140): semi-Lagrangian RK2 advection of the face velocities."""
    out = np.zeros(g.faces, dtype=np.float32)
    prm = _step_params(dt, 0.0, (0.0, 0.0, 0.0), 0.0, advect)
    lib().orc_predict(C.byref(_grid(g)), _ptr(mask), _ptr(np.ascontiguousarray(faces, dtype=np.float32)),
                      C.byref(prm), _ptr(out))
    return out


def viscous(g, mask, pred, cells=None, dt=1e-3, viscosity=0.0, acceleration=(0.0, 0.0, -9.81), surface_tension=0.0,
            em=None):
    """viscous (solver.wgsl:143-163): explicit diffusion, body/EM/capillary acceleration, mobility."""
    out = np.zeros(g.faces, dtype=np.float32)
    prm = _step_params(dt, viscosity, acceleration, surface_tension, False)
    cl = None if cells is None else np.ascontiguousarray(cells, dtype=np.float32).reshape(-1)
    lib().orc_viscous(C.byref(_grid(g)), _ptr(mask), _ptr(np.ascontiguousarray(pred, dtype=np.float32)),
                      None if cl is None else _ptr(cl),
                      None if em is None else _ptr(np.ascontiguousarray(em, dtype=np.float32).reshape(-1)),
                      C.byref(prm), _ptr(out))
    return out


def stress_bound(g, mask, cells, nu, mu_air):
    """viscosity_bound + reduce (src/fluid/viscosity.wgsl:83-104)"""
    L = lib()
    L.orc_stress_bound.restype = C.c_float
    cl = np.ascontiguousarray(cells, dtype=np.float32).reshape(-1)
    return float(L.orc_stress_bound(C.byref(_grid(g)), _ptr(mask), _ptr(cl), C.c_float(nu), C.c_float(mu_air)))


def stress_euler(g, mask, cells, u, nu, mu_air, dt_sub):
    """one explicit substep of stress_diffuse (src/fluid/viscosity.wgsl:105-112)"""
    out = np.zeros(g.faces, np.float32)
    cl = np.ascontiguousarray(cells, dtype=np.float32).reshape(-1)
    lib().orc_stress_euler(C.byref(_grid(g)), _ptr(mask), _ptr(cl), C.c_float(nu), C.c_float(mu_air),
                           C.c_float(dt_sub), _ptr(np.ascontiguousarray(u, dtype=np.float32)), _ptr(out))
    return out


def rkc_schedule(dt, euler_steps):
    """stage rows of the damped RKC2 schedule (src/fluid/stabilized.rs:13-53), shape (stages, 8)"""
    rows = np.zeros((256, 8), np.float32)
    n = lib().orc_rkc_schedule(C.c_double(dt), C.c_uint32(euler_steps), _ptr(rows), 256)
    assert n > 0
    return rows[:n].copy()


def surface_stress(g, mask, cells, u, dt, nu, mu_air, max_substeps=4096):
    """diffuse_surface_stress (src/fluid/mod.rs:1160-1234) -> (status, u_new, substeps, euler count)"""
    out = np.ascontiguousarray(np.array(u, dtype=np.float32))
    sub = C.c_uint32(0)
    cnt = C.c_double(0)
    cl = np.ascontiguousarray(cells, dtype=np.float32).reshape(-1)
    rc = lib().orc_surface_stress(C.byref(_grid(g)), _ptr(mask), _ptr(cl), C.c_double(dt), C.c_float(nu),
                                  C.c_float(mu_air), C.c_uint32(max_substeps), _ptr(out), C.byref(sub), C.byref(cnt))
    return rc, out, int(sub.value), float(cnt.value)


class _EmGrid(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("h", C.c_float)]


class _EmParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("mu0", C.c_float), ("tol2", C.c_float), ("max_iter", C.c_uint32),
                ("magnetostatic", C.c_int), ("nodal", C.c_int), ("nodal_after", C.c_uint32)]


class _EmStats(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("rel_residual", C.c_double), ("scale", C.c_double),
                ("start_residual", C.c_double), ("prediction_dropped", C.c_uint32), ("nodal", C.c_uint32)]


def em_counts(dims):
    """edges per axis (dims + 1 except n_a on axis a) and faces per axis (dims + e_a)"""
    n = list(dims)
    edges = [int(np.prod([n[b] + (0 if b == a else 1) for b in range(3)])) for a in range(3)]
    faces = [int(np.prod([n[b] + (1 if b == a else 0) for b in range(3)])) for a in range(3)]
    return edges, faces


def em_coefficients(dims, h, materials, props, vacuum=False):
    """coefficients (induction.wgsl:82-99) -> dict(sigma, diag, flags, nu, check)"""
    edges, faces = em_counts(dims)
    E, Fn = sum(edges), sum(faces)
    out = dict(sigma=np.zeros(E, np.float32), diag=np.zeros(E, np.float32), flags=np.zeros(E, np.uint8),
               nu=np.zeros(Fn, np.float32), check=np.zeros(4, np.float32))
    g = _EmGrid(dims[0], dims[1], dims[2], h)
    lib().orc_em_coefficients(C.byref(g), _ptr(np.ascontiguousarray(materials, dtype=np.uint32)),
                              _ptr(np.ascontiguousarray(props, dtype=np.float32).reshape(-1)), int(bool(vacuum)),
                              _ptr(out["sigma"]), _ptr(out["diag"]), _ptr(out["flags"]), _ptr(out["nu"]),
                              _ptr(out["check"]))
    return out


def _em_params(dt, mu0, tol, max_iter, magnetostatic, nodal=0, nodal_after=0):
    # Params.physical.w = relative_tolerance.powi(2) with an f32 tolerance (src/em/mod.rs:888)
    return _EmParams(dt, mu0, float(np.float32(tol) * np.float32(tol)), max_iter, int(bool(magnetostatic)),
                     int(nodal), int(nodal_after))


def em_apply(dims, h, co, v, dt=1e-5, mu0=4e-7 * np.pi, magnetostatic=False, precision=32):
    """apply_operator (induction.wgsl:160-170) on v (one value per edge)"""
    dt_ = np.float64 if precision == 64 else np.float32
    v = np.ascontiguousarray(v, dtype=dt_)
    q = np.zeros_like(v)
    g = _EmGrid(dims[0], dims[1], dims[2], h)
    prm = _em_params(dt, mu0, 1e-5, 1, magnetostatic)
    fn = lib().orc_em_apply_f64 if precision == 64 else lib().orc_em_apply_f32
    fn(C.byref(g), C.byref(prm), _ptr(co["sigma"]), _ptr(co["diag"]), _ptr(co["flags"]), _ptr(co["nu"]),
       _ptr(v), _ptr(q))
    return q


def em_solve(dims, h, co, as_now, as_prev, a_prev, mot=None, dt=1e-5, mu0=4e-7 * np.pi, tol=1e-5,
             max_iter=2000, magnetostatic=False, precision=64, a_before=None, prediction=0.0, nodal=0,
             nodal_after=0):
    """One step's Jacobi-CG solve (src/em/mod.rs:863-1010) -> (status, a_out, stats dict); with
    a_before and prediction > 0 the increment starts from prediction * (a_prev - a_before).
    nodal: 0 the reference's switch to the nodal correction (after 1024 iterations of an induction
    solve allowed more), -1 never, 1 after nodal_after iterations (0: from the start)."""
    edges, _ = em_counts(dims)
    E = sum(edges)
    dt_ = np.float64 if precision == 64 else np.float32
    a_out = np.zeros(E, dt_)
    st = _EmStats()
    g = _EmGrid(dims[0], dims[1], dims[2], h)
    prm = _em_params(dt, mu0, tol, max_iter, magnetostatic, nodal, nodal_after)
    f32 = lambda a: _ptr(np.ascontiguousarray(a, dtype=np.float32))  # noqa: E731
    keep = [np.ascontiguousarray(a, dtype=np.float32) for a in (as_now, as_prev, a_prev)]
    mot_ = None if mot is None else np.ascontiguousarray(mot, dtype=np.float32)
    before = None if a_before is None else np.ascontiguousarray(a_before, dtype=np.float32)
    fn = lib().orc_em_solve_f64 if precision == 64 else lib().orc_em_solve_f32
    rc = fn(C.byref(g), C.byref(prm), _ptr(co["sigma"]), _ptr(co["diag"]), _ptr(co["flags"]), _ptr(co["nu"]),
            _ptr(keep[0]), _ptr(keep[1]), None if mot_ is None else _ptr(mot_), _ptr(keep[2]),
            None if before is None else _ptr(before), C.c_float(prediction), _ptr(a_out), C.byref(st))
    del f32
    return rc, a_out, {k: getattr(st, k) for k, _ in _EmStats._fields_}


def em_nodal(dims, h, co, r, dt=1e-5, mu0=4e-7 * np.pi, magnetostatic=False, precision=64):
    """The nodal correction applied to r (nodal.wgsl) -> (z, node diag(G^T A G), node values);
    nodes (n + 1)^3, x fastest."""
    dt_ = np.float64 if precision == 64 else np.float32
    r = np.ascontiguousarray(r, dtype=dt_)
    nodes = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)
    z, S, N = np.zeros_like(r), np.zeros(nodes, dt_), np.zeros(nodes, dt_)
    g = _EmGrid(dims[0], dims[1], dims[2], h)
    prm = _em_params(dt, mu0, 1e-5, 1, magnetostatic)
    fn = lib().orc_em_nodal_f64 if precision == 64 else lib().orc_em_nodal_f32
    fn(C.byref(g), C.byref(prm), _ptr(co["sigma"]), _ptr(co["diag"]), _ptr(co["flags"]), _ptr(r), _ptr(z),
       _ptr(S), _ptr(N))
    return z, S, N


def project(g, mask, alpha, ustar, rel_tol=1e-5, abs_div_tol=0.0, max_iter=1000, precond=0, gauge=0,
            precision=32, tol64=0.0, threads=0):
    """FluidState::project(false) (src/fluid/mod.rs:1015-1115) -> (status, phi f64, u f32, stats dict)."""
    gg = _grid(g)
    o = _Opts(rel_tol, abs_div_tol, max_iter, precond, gauge, precision, tol64, threads)
    st = _Stats()
    phi = np.zeros(g.cells, dtype=np.float64)
    u = np.zeros(g.faces, dtype=np.float32)
    rc = lib().orc_project(C.byref(gg), _ptr(mask), _ptr(alpha), _ptr(np.ascontiguousarray(ustar, dtype=np.float32)),
                           C.byref(o), _ptr(phi), _ptr(u), C.byref(st))
    stats = {k: getattr(st, k) for k, _ in _Stats._fields_}
    return rc, phi, u, stats


class MG:
    """Multigrid probes (src/fluid/multigrid_probes.wgsl and the test helpers around it)."""

    def __init__(self, g, mask, alpha, meanfree: bool, precision=32):
        self.g = g
        gg = _grid(g)
        self._keep = (mask, alpha)
        self.h = lib().orc_mg_new(C.byref(gg), _ptr(mask), _ptr(alpha), int(meanfree), precision)
        d = np.zeros(64 * 4, dtype=np.uint32)
        n = lib().orc_mg_levels(self.h, _ptr(d))
        self.levels = [tuple(int(v) for v in d[4 * l:4 * l + 4]) for l in range(n)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mg_free(self.h)
            self.h = None

    def action(self, level, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.levels[level][3], dtype=np.float64)
        lib().orc_mg_action(self.h, level, _ptr(v), _ptr(out))
        return out

    def vcycle(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.g.cells, dtype=np.float64)
        lib().orc_mg_vcycle(self.h, _ptr(r), _ptr(z))
        return z

    def coarse(self, rhs, sweeps=None):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        z = np.zeros(self.levels[-1][3], dtype=np.float64)
        if sweeps is None:
            lib().orc_mg_coarse_solve(self.h, _ptr(rhs), _ptr(z))
        else:
            lib().orc_mg_coarse_sweeps(self.h, _ptr(rhs), _ptr(z), sweeps)
        return z

    def graph(self, level):
        n = self.levels[level][3]
        c4 = np.zeros(4 * n, dtype=np.float64)
        diag = np.zeros(n, dtype=np.float64)
        lib().orc_mg_graph(self.h, level, _ptr(c4), _ptr(diag))
        return c4.reshape(n, 4), diag
