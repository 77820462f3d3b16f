"""The SIMP loop on the GPU: density filter, sensitivities and the OC update
around the batched MG-PCG solve (liblpmi.so lp_simp_*).

Mirrors:
- the top3d loop ``optimize_reference(p, "top3d")`` (benchmarks/reference.py
  :214-257): SIMP p = 3, Emin 1e-9, density filter H/Hs, OC bisection with
  l2 = 1e9 and relative tolerance 1e-3, stop when change <= 0.01 or 200
  iterations, returning ``x, iterations, compliance_reported, seconds``;
- ``density_filter(dims, radius, pad)`` / ``GridFilter`` (simp.py:1143-1189):
  ``h @ v``, ``h.T``, ``hs``, on the GPU.

Several load cases are solved together and their compliances summed.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from .fem import hex8_element, hex8_ke


def cone_taps(radius: float, reach: int):
    taps = []
    for dx in range(-reach, reach + 1):
        for dy in range(-reach, reach + 1):
            for dz in range(-reach, reach + 1):
                w = radius - np.sqrt(dx * dx + dy * dy + dz * dz)
                if w > 0:
                    taps.append(((dx, dy, dz), float(w)))
    return taps


class GridFilter:
    """simp.GridFilter on a full box (unit pitch): ``h @ v`` runs the cone
    stencil on the GPU; symmetric, so ``h.T is h``."""

    def __init__(self, dims, radius: float, reach: int | None = None):
        self.dims = tuple(int(d) for d in dims)
        self.radius = float(radius)
        self.reach = int(np.floor(radius)) if reach is None else int(reach)
        self.taps = cone_taps(self.radius, self.reach)
        self.n = int(np.prod(self.dims))
        self.shape = (self.n, self.n)
        self.T = self

    def __matmul__(self, v):
        v = np.ascontiguousarray(np.asarray(v, np.float64).reshape(self.dims))
        out = np.empty(self.dims)
        _lib.check(_lib.lib().lp_filter_apply(*self.dims, self.radius, self.reach, _lib.ptr(v), _lib.ptr(out), 0),
                   "filter")
        return out.ravel()

    def sum(self, axis=1):
        return self @ np.ones(self.n)


def _project_call(xt, beta, eta, want_value, want_deriv):
    x = np.ascontiguousarray(np.asarray(xt, float))
    out = np.empty_like(x) if want_value else None
    dout = np.empty_like(x) if want_deriv else None
    _lib.check(_lib.lib().lp_project(x.size, _lib.ptr(x), float(beta), float(eta),
                                     _lib.ptr(out) if want_value else None,
                                     _lib.ptr(dout) if want_deriv else None, 0), "project")
    return out, dout


def project(xt, beta, eta):
    """Heaviside projection simp._project (simp.py:1443-1446), on the GPU."""
    return _project_call(xt, beta, eta, True, False)[0]


def dproject(xt, beta, eta):
    """Its derivative simp._dproject (simp.py:1448-1450), on the GPU."""
    return _project_call(xt, beta, eta, False, True)[1]


def density_filter(dims, radius: float, pad: bool = True):
    """(h, hs) as simp.density_filter (simp.py:1181-1189) on a full box."""
    h = GridFilter(dims, radius)
    hs = np.full(h.n, sum(w for _, w in h.taps)) if pad else h.sum()
    return h, hs


class GridSimp:
    """One SIMP study on the device: the grid hierarchy, the load cases and
    the design fields stay in HBM between iterations."""

    def __init__(self, problem, *, precision="fp32", smoother="chebyshev", omega=0.6, degree=2, rtol=1e-8,
                 maxiter=500, coarse_cells=2000, max_levels=0, element=None, nu=0.3, penal=3.0, emin=1e-9,
                 move=0.2, filter="top3d", warm_start=True, optimizer="oc"):
        from .mg import GridElasticity
        from .fem import box_mesh

        _lib.require_gpu()
        if problem.present is not None:
            raise NotImplementedError("removed cells: use GridElasticity directly")
        self.problem = problem
        mesh = box_mesh(*problem.dims)
        fixed = problem.fixed_grid()
        fdof = np.empty(mesh.n_dof, bool)
        flat = fixed.reshape(3, -1)
        for i in range(3):
            fdof[i::3] = flat[i]
        loads = problem.load_grids()
        self.nrhs = loads.shape[0]
        self.grid = GridElasticity(mesh, nu, fdof, rtol=rtol, maxiter=maxiter, coarse_cells=coarse_cells,
                                   degree=degree, dtype=np.float64 if precision == "fp64" else np.float32,
                                   smoother=smoother, omega=omega, max_rhs=self.nrhs, max_levels=max_levels,
                                   element=element or hex8_element(), precision=precision)
        h = C.c_void_p()
        self._loads = np.ascontiguousarray(loads)
        _lib.check(_lib.lib().lp_simp_create(C.byref(h), self.grid._h, self.nrhs, _lib.ptr(self._loads), None,
                                             problem.volfrac, problem.rmin, _lib.FILTER[filter], penal, emin, move,
                                             int(warm_start)), "lp_simp_create")
        self._h = h
        self.iterations = 0
        if optimizer not in ("oc", "mma"):
            raise ValueError(f"optimizer must be 'oc' or 'mma', not {optimizer!r}")
        self.optimizer = optimizer
        _lib.check(_lib.lib().lp_simp_set_optimizer(h, 0 if optimizer == "oc" else 1), "simp_set_optimizer")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().lp_simp_destroy(h)
            except Exception:
                pass
            self._h = None

    def iterate(self, n: int = 1):
        stats = (_lib.SimpStats * max(n, 1))()
        _lib.check(_lib.lib().lp_simp_iterate(self._h, n, stats), "simp_iterate")
        self.iterations += n
        # the hierarchy was rebuilt on the device: refresh the level table
        self.grid._refresh(lmax=False)
        out = []
        for k in range(n):
            s = stats[k]
            out.append(dict(compliance=s.compliance, change=s.change, volume=s.volume, cg_iters=s.cg_iters,
                            cg_iters_total=s.cg_iters_total, oc_evals=s.oc_evals, seconds=s.seconds))
        return out

    def _get(self, what):
        out = np.empty(self.problem.dims)
        _lib.check(_lib.lib().lp_simp_get(self._h, what, _lib.ptr(out)), "simp_get")
        return out

    @property
    def x(self):
        return self._get(0)

    @property
    def xphys(self):
        return self._get(1)

    @property
    def dc(self):
        return self._get(2)

    @property
    def ce(self):
        return self._get(3)

    def displacements(self):
        out = np.empty((self.nrhs, 3) + self.problem.node_shape)
        _lib.check(_lib.lib().lp_simp_displacements(self._h, _lib.ptr(out)), "simp_displacements")
        return out


def optimize_top3d(problem, iterations=None, maxit=200, on_iter=None, **kw):
    """benchmarks/reference.optimize_reference(p, "top3d") on the GPU
    (reference.py:214-257): stop at change <= 0.01 (or after `iterations`)."""
    s = GridSimp(problem, **kw)
    t0 = time.time()
    loop, change, c = 0, 1.0, 0.0
    hist = []
    while (iterations is None and change > 0.01 and loop < maxit) or (iterations is not None and loop < iterations):
        st = s.iterate(1)[0]
        loop += 1
        change, c = st["change"], st["compliance"]
        hist.append(st)
        if on_iter is not None:
            on_iter(loop, c, change)
    return dict(x=s.xphys, iterations=loop, compliance_reported=c, seconds=time.time() - t0, history=hist,
                solver="top3d")
