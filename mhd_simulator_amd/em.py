"""The EM step's Jacobi-CG (EmState's solve, src/em/mod.rs:863-1010) over libmhdp.so.

    with EmSolver(dims, h, materials, props) as em:
        iterations, residual = em.solve(as_now, as_prev, a_prev, a_out, dt=1e-5, rel_tol=1e-5)

Edge arrays are device tensors (float32, one value per edge, packed x | y | z; edges of axis a
have dims n + 1 except n_a on axis a).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .fluid import PressureError, _ptr

MU0 = 4e-7 * math.pi


def edge_count(dims) -> int:
    nx, ny, nz = dims
    return nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz


def face_count(dims) -> int:
    nx, ny, nz = dims
    return (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)


class EmSolver:
    """One EM grid: materials (one id per cell), props (count, 2): conductivity S/m and relative
    permeability per material id; vacuum ignores the materials (Params.control.z)."""

    def __init__(self, dims, h, materials, props, vacuum=False):
        self.lib = _lib.load()
        self.dims = tuple(int(v) for v in dims)
        self.h = float(h)
        handle = C.c_void_p()
        g = _lib.Grid(self.dims[0], self.dims[1], self.dims[2], 0, self.h)
        rc = self.lib.mhdp_em_create(C.byref(handle), C.byref(g))
        if rc != _lib.OK:
            raise PressureError(rc, "EM grid indexing exceeds u32" if rc == _lib.INVALID else "EM allocation failed")
        self.handle = handle
        self.edges = int(self.lib.mhdp_em_edges(handle))
        self.set_materials(materials, props, vacuum)

    def _check(self, rc):
        if rc != _lib.OK:
            raise PressureError(rc, self.lib.mhdp_em_last_error(self.handle).decode())

    def set_materials(self, materials, props, vacuum=False):
        m = np.ascontiguousarray(np.asarray(materials, dtype=np.uint32).ravel())
        p = np.ascontiguousarray(np.asarray(props, dtype=np.float32).reshape(-1, 2))
        if m.size != int(np.prod(self.dims)):
            raise ValueError("one material id per cell")
        rng = (C.c_float * 2)()
        self._check(self.lib.mhdp_em_set_materials(self.handle, _ptr(m), _ptr(p), p.shape[0], int(bool(vacuum)), rng))
        self.sigma_range = (rng[0], rng[1])

    def _opts(self, dt, rel_tol, max_iter, magnetostatic, mu0, nodal=0, nodal_after=0):
        return _lib.EmOpts(dt, mu0, rel_tol, max_iter, int(bool(magnetostatic)), int(nodal), int(nodal_after))

    def solve(self, as_now, as_prev, a_prev, a_out, *, motional=None, dt=1e-5, rel_tol=1e-5, max_iter=1000,
              magnetostatic=False, mu0=MU0, a_before=None, prediction=0.0, nodal=0, nodal_after=0):
        """a_out = a_prev + the step's increment; returns (iterations, relative residual).  With
        a_before (the potential accepted one step before a_prev) and prediction > 0 (the ratio of
        this step's dt to the last one's) the increment starts from prediction * (a_prev - a_before)
        (mhdp_em_solve_predicted_dev); stats["prediction_dropped"] tells whether that start was
        rejected for the zero start.  nodal: the nodal auxiliary-space correction's policy (0 the
        reference's: switch after 1024 iterations of an induction solve allowed more; -1 never; 1
        after nodal_after iterations, 0 = from the start); stats["nodal"] tells whether it ran."""
        extra = tuple(t for t in (motional, a_before) if t is not None)
        for t in (as_now, as_prev, a_prev, a_out) + extra:
            if not hasattr(t, "data_ptr") or t.numel() != self.edges:
                raise ValueError("the EM solve takes device tensors with one value per edge")
        o = self._opts(dt, rel_tol, max_iter, magnetostatic, mu0, nodal, nodal_after)
        st = _lib.EmStats()
        self._check(self.lib.mhdp_em_solve_predicted_dev(self.handle, C.byref(o), _ptr(as_now), _ptr(as_prev),
                                                         _ptr(motional), _ptr(a_prev), _ptr(a_before),
                                                         float(prediction), _ptr(a_out), C.byref(st)))
        self.stats = st.as_dict()
        return int(st.iterations), float(st.rel_residual)

    # -- test hooks
    def coefficients(self):
        out = dict(sigma=np.zeros(self.edges, np.float32), diag=np.zeros(self.edges, np.float32),
                   flags=np.zeros(self.edges, np.uint8), nu=np.zeros(face_count(self.dims), np.float32))
        self._check(self.lib.mhdp_em_get_coefficients(self.handle, _ptr(out["sigma"]), _ptr(out["diag"]),
                                                      _ptr(out["flags"]), _ptr(out["nu"])))
        return out

    def apply(self, v, *, dt=1e-5, magnetostatic=False, mu0=MU0):
        v = np.ascontiguousarray(v, dtype=np.float32)
        q = np.zeros_like(v)
        o = self._opts(dt, 1e-5, 1, magnetostatic, mu0)
        self._check(self.lib.mhdp_em_probe_apply(self.handle, C.byref(o), _ptr(v), _ptr(q)))
        return q

    def nodal(self, r, *, dt=1e-5, magnetostatic=False, mu0=MU0):
        """The nodal correction applied to r (host vector) -> (z, node values)."""
        r = np.ascontiguousarray(r, dtype=np.float32)
        z = np.zeros_like(r)
        nodes = np.zeros((self.dims[0] + 1) * (self.dims[1] + 1) * (self.dims[2] + 1), np.float32)
        o = self._opts(dt, 1e-5, 1, magnetostatic, mu0)
        self._check(self.lib.mhdp_em_probe_nodal(self.handle, C.byref(o), _ptr(r), _ptr(z), _ptr(nodes)))
        return z, nodes

    def set_kernel_timing(self, kernel=None):
        """Time every launch of one CG kernel ("curl", "apply", "update", "search"; None: off)
        during the following solves with HIP events (mhdp_em_set_kernel_timing)."""
        self._check(self.lib.mhdp_em_set_kernel_timing(self.handle, -1 if kernel is None else _lib.EM_KERNELS[kernel]))

    def kernel_time(self):
        """(total ms, launches) of the timed kernel's working launches in the last solve."""
        ms, n = C.c_double(), C.c_uint32()
        self._check(self.lib.mhdp_em_kernel_time(self.handle, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.mhdp_em_destroy(self.handle)
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
