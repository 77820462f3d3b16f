"""p-mean von Mises stress aggregate and its adjoint gradient on the GPU.

Mirrors the non-averaged stress branch of loadpath's constrained.py
(constrained.py:607-664), which is the p-norm stress objective of a
minimum-mass study:

- ``hex8_cb(nu, points, element, pitch)``: C B at the sample points, the
  centre (fem.hex8_b_center) or the 8 corners with the bubbles
  (constrained.py:383-387), from liblpmi's host routine lp_hex8_cb;
- ``StressPNorm(grid, points, p, q, limit, counted)``: the aggregate on a
  ``GridElasticity``'s fine grid. ``evaluate(us, x)`` returns per load case
  g = kappa pm, s_max, the adjoint right-hand side dg/du (zero on fixed
  dofs) and the explicit dg/dx; ``gradient(us, x, de)`` adds the adjoint
  solve (the grid's MG-PCG, one batched solve for all cases) and returns
  dg/dx = explicit - de * lambda_e^T Ke u_e.

x is the material share per element (constrained.material with no share:
x itself); de = dE/dx of the interpolation in use.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

P_DEFAULT = 8.0     # constrained._P
Q_DEFAULT = 1.0     # constrained._Q
_POINTS = {"center": 0, "corners": 1}


def hex8_cb(nu: float, points: str = "center", element: str = "im", pitch: float = 1.0) -> np.ndarray:
    if points not in _POINTS:
        raise ValueError(f"points must be one of {tuple(_POINTS)}, not {points!r}")
    if element not in ("im", "plain"):
        raise ValueError(f"element must be 'im' or 'plain', not {element!r}")
    npts = 1 if points == "center" else 8
    out = np.empty((npts, 6, 24))
    _lib.check(_lib.lib().lp_hex8_cb(float(nu), 0 if element == "im" else 1, _POINTS[points], _lib.ptr(out)),
               "lp_hex8_cb")
    return out / pitch


class StressPNorm:
    """The p-mean stress aggregate on a GridElasticity (constrained.py:607-664)."""

    def __init__(self, grid, points: str = "center", p: float = P_DEFAULT, q: float = Q_DEFAULT,
                 limit: float = 1.0, counted=None):
        self.grid = grid
        self.points = points
        self.p, self.q, self.limit = float(p), float(q), float(limit)
        self.cb = np.ascontiguousarray(hex8_cb(grid.nu, points, grid.element, grid.mesh.pitch))
        ne = len(grid.elem_ijk)
        cnt = np.ones(ne, bool) if counted is None else np.asarray(counted, bool)
        if cnt.shape != (ne,):
            raise ValueError("counted must be (n_elem,) bool")
        self._counted = np.ascontiguousarray(grid.cell_grid(cnt.astype(float)) > 0.5).astype(np.uint8)
        self.nc = float(max(int(cnt.sum()), 1) * len(self.cb))

    def evaluate(self, us, x, kappa: float = 1.0, need_grad: bool = True) -> dict:
        us = [np.asarray(u, float) for u in us]
        nr = len(us)
        g = self.grid
        ug = np.ascontiguousarray(np.stack([g.to_grid(u) for u in us]))
        xg = np.ascontiguousarray(g.cell_grid(np.asarray(x, float)))
        pm = np.empty(nr)
        smax = np.empty(nr)
        rhs = np.empty_like(ug) if need_grad else None
        gx = np.empty((nr,) + g.cells) if need_grad else None
        _lib.check(_lib.lib().lp_stress_pnorm(
            g._h, nr, _lib.ptr(ug), _lib.ptr(xg), _lib.ptr(self._counted), _lib.ptr(self.cb), len(self.cb),
            self.p, self.q, self.limit, float(kappa), self.nc, _lib.ptr(pm), _lib.ptr(smax),
            _lib.ptr(rhs) if need_grad else None, _lib.ptr(gx) if need_grad else None, 0), "stress_pnorm")
        out = dict(g=kappa * pm, pm=pm, smax=smax)
        if need_grad:
            at = tuple(g.elem_ijk.T)
            out["rhs"] = [g.from_grid(r) for r in rhs]
            out["gx"] = [v[at] for v in gx]
        return out

    def gradient(self, us, x, de, kappa: float = 1.0):
        """(g per case, dg/dx per case) with one batched adjoint solve."""
        ev = self.evaluate(us, x, kappa)
        lams = self.grid.solve_many(ev["rhs"])
        de = np.asarray(de, float)
        grads = [gx - de * self.grid.element_product(lam, u) for gx, lam, u in zip(ev["gx"], lams, us)]
        return ev["g"], grads
