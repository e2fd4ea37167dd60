"""Host-side mirror of the reference's pressure-projection interface.

The reference exposes the projection as ``FluidState::project(&mut self,
warm_start) -> Result<(u32, f32), String>`` (src/fluid/mod.rs:1015) over GPU
state that ``FluidState`` owns, configured by ``FluidConfig``
(src/fluid/mod.rs:14-82).  ``PressureProjection`` plays the same role around a
``libmhdp.so`` context: it owns the device buffers, labels components with
``anchors()`` when the cells are set, and ``project()`` returns
``(iterations, relative residual)`` or raises ``PressureError`` carrying the
reference's error texts.  Inputs may be numpy arrays on the host or torch
tensors already resident on the GPU (the HBM-resident path).
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from .configs import Grid


class PressureError(RuntimeError):
    """Err(String) of FluidState::project; ``status`` is the mhdp_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


@dataclasses.dataclass
class FluidConfig:
    """The pressure-solver fields of FluidConfig (src/fluid/mod.rs:14-49) and their
    defaults; validate() keeps the reference's ranges (src/fluid/mod.rs:58-82)."""
    relative_tolerance: float = 1e-5
    absolute_divergence_tolerance: float = 1e-5
    max_iterations: int = 1000

    def validate(self) -> None:
        if not (1e-7 <= self.relative_tolerance <= 1e-2) or not (1 <= self.max_iterations <= 20000) \
                or not np.isfinite(self.absolute_divergence_tolerance) or self.absolute_divergence_tolerance < 0:
            raise ValueError("Invalid fluid timestep, initial condition or solver limits/tolerances")


def anchors(fluid: np.ndarray, dims, periodic=(False, False, False)):
    """anchors() (src/fluid/mod.rs:1447-1493) -> (encoded uint32 mask, components), on the host."""
    lib = _lib.load()
    mask = np.ascontiguousarray((np.asarray(fluid).ravel() != 0).astype(np.uint32))
    d = np.ascontiguousarray(np.array(dims, dtype=np.uint32))
    bits = sum(1 << a for a in range(3) if periodic[a])
    comps = lib.mhdp_anchors(mask.ctypes.data_as(C.c_void_p), d.ctypes.data_as(C.c_void_p), bits)
    return mask, int(comps)


def hierarchy(dims4):
    """hierarchy() (src/fluid/multigrid.rs:25-65) -> ([level dims], coarse cells) or None on error."""
    lib = _lib.load()
    d = np.ascontiguousarray(np.array(dims4, dtype=np.uint32))
    out = np.zeros(64 * 4, dtype=np.uint32)
    coarse = np.zeros(1, dtype=np.uint64)
    n = lib.mhdp_hierarchy(d.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), 64,
                           coarse.ctypes.data_as(C.c_void_p))
    if n < 0:
        return None
    return [tuple(int(v) for v in out[4 * l:4 * l + 4]) for l in range(n)], int(coarse[0])


def project_batch(projections, u_stars, *, precision="mixed", precond="auto", gauge="auto", refine_tol=1e-8,
                  max_outer=0, inner_tol=0.0, smooth_weight=0.0, coarse_scale=0.0, coarse_visits=0, u_outs=None,
                  phi_outs=None, smoother="auto"):
    """Many independent projections in one call (mhdp_project_batch): device tensors only.
    The options of the first projection's FluidConfig apply to all.  Returns a list of
    (iterations, relative residual); raises PressureError for the first failing problem."""
    n = len(projections)
    if n == 0:
        return []
    if len(u_stars) != n or (u_outs is not None and len(u_outs) != n) or (phi_outs is not None and len(phi_outs) != n):
        raise ValueError("one u_star (and output) per projection")
    for pr, u in zip(projections, u_stars):
        if not hasattr(u, "data_ptr") or u.numel() != pr.grid.faces:
            raise ValueError("u_star must be a device tensor holding the three face planes")
    lib = projections[0].lib
    o = projections[0]._opts(precision, precond, gauge, refine_tol, max_outer, inner_tol, smooth_weight, coarse_scale,
                             coarse_visits=coarse_visits, smoother=smoother)
    P = C.c_void_p * n
    ctxs = P(*[pr.handle for pr in projections])
    us = P(*[u.data_ptr() for u in u_stars])
    uo = P(*[t.data_ptr() for t in u_outs]) if u_outs is not None else None
    po = P(*[t.data_ptr() for t in phi_outs]) if phi_outs is not None else None
    st = (_lib.Stats * n)()
    rc = lib.mhdp_project_batch(ctxs, n, us, C.byref(o), uo, po, st)
    out = []
    for k, pr in enumerate(projections):
        pr.stats = st[k].as_dict()
        out.append((int(st[k].iterations), float(st[k].rel_residual if precision == "fp32" else st[k].rel_residual_f64)))
    if rc != _lib.OK:
        bad = next((pr for pr in projections if lib.mhdp_last_error(pr.handle)), projections[0])
        raise PressureError(rc, lib.mhdp_last_error(bad.handle).decode() or "batch projection failed")
    return out


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


class PressureProjection:
    """The projection half of FluidState: one libmhdp context per grid."""

    def __init__(self, grid: Grid, fluid: np.ndarray, alpha: np.ndarray | None = None,
                 config: FluidConfig | None = None):
        self.lib = _lib.load()
        self.grid = grid
        self.config = config or FluidConfig()
        self.config.validate()
        g = _lib.Grid(grid.nx, grid.ny, grid.nz, grid.periodic_bits, grid.h)
        ph = _lib.Phase(int(grid.two_phase), grid.rho_liquid, grid.rho_air)
        h = C.c_void_p()
        rc = self.lib.mhdp_create(C.byref(h), C.byref(g), C.byref(ph))
        if rc != _lib.OK:
            raise PressureError(rc, "Fluid allocation exceeds device buffer limits" if rc != _lib.INVALID
                                else "Invalid pressure grid dimensions")
        self.handle = h
        self.stats: dict = {}
        self.set_cells(fluid, alpha)

    # -- state
    def _check(self, rc):
        if rc != _lib.OK:
            raise PressureError(rc, self.lib.mhdp_last_error(self.handle).decode())

    def set_cells(self, fluid: np.ndarray, alpha: np.ndarray | None = None):
        f = np.ascontiguousarray(np.asarray(fluid).ravel().astype(np.uint8))
        a = None if alpha is None else np.ascontiguousarray(np.asarray(alpha, dtype=np.float32).ravel())
        if f.size != self.grid.cells or (a is not None and a.size != self.grid.cells):
            raise ValueError("cell arrays must have one entry per cell")
        self._check(self.lib.mhdp_set_cells(self.handle, _ptr(f), _ptr(a)))
        self.mask = np.zeros(self.grid.cells, dtype=np.uint32)
        self._check(self.lib.mhdp_get_mask(self.handle, _ptr(self.mask)))
        self.components = int(np.count_nonzero((self.mask & 3) == 2))
        self.fluid_cells = int(np.count_nonzero(self.mask))

    @property
    def device_bytes(self) -> int:
        return int(self.lib.mhdp_device_bytes(self.handle))

    def _opts(self, precision, precond, gauge, refine_tol, max_outer, inner_tol, smooth_weight=0.0,
              coarse_scale=0.0, warm_start=False, coarse_visits=0, smoother="auto"):
        c = self.config
        return _lib.Opts(c.relative_tolerance, c.absolute_divergence_tolerance,
                         c.max_iterations, _lib.PRECOND[precond], _lib.PRECISION[precision], _lib.GAUGE[gauge],
                         refine_tol, max_outer, inner_tol, smooth_weight, coarse_scale, int(bool(warm_start)),
                         int(coarse_visits), _lib.SMOOTHER[smoother])

    def set_potential(self, phi):
        """The previous potential a warm start begins from (numpy float64 or a device tensor)."""
        if hasattr(phi, "data_ptr"):
            if phi.numel() != self.grid.cells:
                raise ValueError("one potential per cell")
            self._check(self.lib.mhdp_set_potential(self.handle, _ptr(phi), 1))
        else:
            a = np.ascontiguousarray(np.asarray(phi, dtype=np.float64).ravel())
            if a.size != self.grid.cells:
                raise ValueError("one potential per cell")
            self._check(self.lib.mhdp_set_potential(self.handle, _ptr(a), 0))

    # -- FluidState::project(false)
    def project(self, u_star, *, precision="fp32", precond="auto", gauge="auto", refine_tol=1e-8,
                max_outer=0, inner_tol=0.0, smooth_weight=0.0, coarse_scale=0.0, warm_start=False, u_out=None,
                phi_out=None, coarse_visits=0, smoother="auto"):
        """Project u_star; returns (iterations, relative residual) like the reference.
        warm_start: FluidState::project(true) - begin from the previous potential.

        numpy input: host path, results in ``self.velocity`` / ``self.phi``.
        torch CUDA input: device path; ``u_out`` (float32, faces) and ``phi_out``
        (float64, cells) are device tensors the caller provides or None."""
        o = self._opts(precision, precond, gauge, refine_tol, max_outer, inner_tol, smooth_weight, coarse_scale,
                       warm_start, coarse_visits, smoother)
        st = _lib.Stats()
        if hasattr(u_star, "data_ptr"):
            if u_star.numel() != self.grid.faces:
                raise ValueError("u_star must hold the three face planes")
            rc = self.lib.mhdp_project_dev(self.handle, _ptr(u_star), C.byref(o), _ptr(u_out), _ptr(phi_out),
                                           C.byref(st))
        else:
            u = np.ascontiguousarray(np.asarray(u_star, dtype=np.float32).ravel())
            if u.size != self.grid.faces:
                raise ValueError("u_star must hold the three face planes")
            vel = np.zeros(self.grid.faces, dtype=np.float32)
            phi = np.zeros(self.grid.cells, dtype=np.float64)
            rc = self.lib.mhdp_project(self.handle, _ptr(u), C.byref(o), _ptr(vel), _ptr(phi), C.byref(st))
            if rc == _lib.OK:
                self.velocity, self.phi = vel, phi
        self._check(rc)
        self.stats = st.as_dict()
        return int(st.iterations), float(st.rel_residual if precision == "fp32" else st.rel_residual_f64)

    # -- the momentum step that produces u* (predict + viscous, solver.wgsl:127-163)
    def _faces_dev(self, *ts):
        for t in ts:
            if not hasattr(t, "data_ptr") or t.numel() != self.grid.faces:
                raise ValueError("the step takes device tensors holding the three face planes")

    def predict(self, u, u_out, *, dt, advect=True):
        """u_out = u advected along itself (semi-Lagrangian RK2); closed faces 0."""
        self._faces_dev(u, u_out)
        sp = _lib.StepParams(dt, 0.0, (C.c_float * 3)(0.0, 0.0, 0.0), 0.0, int(bool(advect)))
        self._check(self.lib.mhdp_predict_dev(self.handle, _ptr(u), C.byref(sp), _ptr(u_out)))

    def viscous(self, u_pred, u_out, *, dt, viscosity=0.0, acceleration=(0.0, 0.0, 0.0), surface_tension=0.0,
                cells=None, em=None):
        """u_out = explicit viscous update + body/EM/capillary acceleration, times the mobility."""
        self._faces_dev(u_pred, u_out)
        for t in (cells, em):
            if t is not None and t.numel() != 16 * self.grid.cells:
                raise ValueError("cells / em hold 16 floats per cell")
        sp = _lib.StepParams(dt, viscosity, (C.c_float * 3)(*acceleration), surface_tension, 0)
        self._check(self.lib.mhdp_viscous_dev(self.handle, _ptr(u_pred), _ptr(cells), _ptr(em), C.byref(sp),
                                              _ptr(u_out)))

    def surface_stress(self, u, cells, *, dt, viscosity, mu_air, max_substeps=4096):
        """diffuse_surface_stress (two-phase): u (device faces) in place; returns the substep count."""
        self._faces_dev(u)
        if not hasattr(cells, "data_ptr") or cells.numel() != 16 * self.grid.cells:
            raise ValueError("cells holds 16 floats per cell")
        sp = _lib.StressParams(dt, viscosity, mu_air, max_substeps)
        n = C.c_uint32(0)
        self._check(self.lib.mhdp_surface_stress_dev(self.handle, _ptr(u), _ptr(cells), C.byref(sp), C.byref(n)))
        return int(n.value)

    # -- fields() + diagnostics() (solver.wgsl:256-291, read_diagnostics mod.rs:1116-1146)
    def fields(self, u_faces, phi, cells, *, dt, acceleration=(0.0, 0.0, 0.0), mushy=False, gauge="auto",
               em=None, no_alloc=False):
        """Rewrite ``cells`` (device float32, 16 per cell: the FluidCell packet, whose reserved words
        are read first) from the projected faces and the potential (device tensors); returns the
        diagnostics as a dict.  Raises PressureError ("Nonfinite fluid fields; state restored") when a
        total is not finite or a cell is flagged - ``cells`` has then been written.  ``no_alloc``: the
        entry point that allocates no device memory (its partial records live in the library's static
        per-GPU scratch; calls overlapping on one GPU race)."""
        for t, n in ((u_faces, self.grid.faces), (phi, self.grid.cells), (cells, 16 * self.grid.cells)):
            if not hasattr(t, "data_ptr") or t.numel() != n:
                raise ValueError("fields takes device tensors of the grid's sizes")
        if em is not None and em.numel() != 16 * self.grid.cells:
            raise ValueError("em holds 16 floats per cell")
        fp = _lib.FieldsParams(dt, (C.c_float * 3)(*acceleration), int(bool(mushy)), _lib.GAUGE[gauge])
        d = _lib.Diagnostics()
        fn = self.lib.mhdp_fields_noalloc_dev if no_alloc else self.lib.mhdp_fields_dev
        self._check(fn(self.handle, _ptr(u_faces), _ptr(phi), _ptr(em), C.byref(fp), _ptr(cells), C.byref(d)))
        return d.as_dict()

    # -- test hooks (the reference's cfg(test) probes)
    def coefficients(self, meanfree: bool):
        n = self.grid.cells
        beta = np.zeros(3 * n, dtype=np.float32)
        diag = np.zeros(n, dtype=np.float32)
        dinv = np.zeros(n, dtype=np.float32)
        self._check(self.lib.mhdp_probe_coefficients(self.handle, int(meanfree), _ptr(beta), _ptr(diag), _ptr(dinv)))
        return beta.reshape(3, n), diag, dinv

    def apply(self, p: np.ndarray, meanfree: bool):
        p = np.ascontiguousarray(p, dtype=np.float32)
        q = np.zeros(self.grid.cells, dtype=np.float32)
        self._check(self.lib.mhdp_probe_apply(self.handle, int(meanfree), _ptr(p), _ptr(q)))
        return q

    def vcycle(self, r: np.ndarray, meanfree: bool, precision="fp32", smooth_weight=2.0 / 3.0,
               coarse_scale=1.0, coarse_visits=1, smoother="jacobi"):
        """One multigrid cycle; the defaults reproduce the reference's V-cycle (coarse_visits=2:
        the W-cycle at level 1 the solver uses on grids of 2^24 cells and more)."""
        r = np.ascontiguousarray(r, dtype=np.float32)
        z = np.zeros(self.grid.cells, dtype=np.float32)
        self._check(self.lib.mhdp_probe_vcycle(self.handle, int(meanfree), _lib.PRECISION[precision],
                                               smooth_weight, coarse_scale, int(coarse_visits),
                                               _lib.SMOOTHER[smoother], _ptr(r), _ptr(z)))
        return z

    def mg_graph(self, level: int, meanfree: bool):
        dims = np.zeros(4, dtype=np.uint32)
        self._check(self.lib.mhdp_probe_mg_graph(self.handle, int(meanfree), level, None, _ptr(dims)))
        g5 = np.zeros(5 * int(dims[3]), dtype=np.float32)
        self._check(self.lib.mhdp_probe_mg_graph(self.handle, int(meanfree), level, _ptr(g5), _ptr(dims)))
        return tuple(int(v) for v in dims), g5.reshape(5, int(dims[3]))

    def multigrid_state(self, action=0, arg=0):
        """The optional multigrid storage: (state, bytes) with state "none", "allocated" or
        "refused".  action 1: the next allocation attempt fails after ``arg`` arrays; action 2: free
        it and record a refused attempt (test hooks of the reference's lazy-workspace test)."""
        st = C.c_int(0)
        nb = C.c_uint64(0)
        self._check(self.lib.mhdp_probe_multigrid(self.handle, int(action), int(arg), C.byref(st), C.byref(nb)))
        return ("none", "allocated", "refused")[st.value], int(nb.value)

    def next_cycle(self, smooth_weight, coarse_scale=1.0, coarse_visits=1):
        """Test hook: the next projection's first attempt uses this cycle unvalidated."""
        self._check(self.lib.mhdp_probe_next_cycle(self.handle, float(smooth_weight), float(coarse_scale),
                                                   int(coarse_visits)))

    def rhs(self, u_star: np.ndarray, meanfree: bool):
        u = np.ascontiguousarray(np.asarray(u_star, dtype=np.float32).ravel())
        b = np.zeros(self.grid.cells, dtype=np.float64)
        self._check(self.lib.mhdp_probe_rhs(self.handle, int(meanfree), _ptr(u), _ptr(b)))
        return b

    def close(self):
        if getattr(self, "handle", None):
            self.lib.mhdp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
