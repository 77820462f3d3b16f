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

- ``StressAveraged(mesh, nu, radius, readable, p, q, limit, bond)``: the
  averaged branch (constrained._Averaged, constrained.py:146-252, read at
  517-605, no thin walls): each readable node's stress averaged over the
  material within radius, relaxed by the density round it, the P-mean over
  the nodes with Le's normalization, the adjoint right-hand side and the
  explicit dg/dxe (lp_stress_averaged).

x is the material share per element (constrained.material with no share:
x itself); de = dE/dx of the interpolation in use.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

# the node (fem.CORNERS order) each corner sample sits on, samples x outermost
# (constrained._CORNER_OF, constrained.py:106)
CORNER_OF = (0, 4, 2, 6, 1, 5, 3, 7)
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
                 limit: float = 1.0, counted=None, point_counted=None):
        self.grid = grid
        self.points = points
        self.p, self.q, self.limit = float(p), float(q), float(limit)
        self.cb = np.ascontiguousarray(hex8_cb(grid.nu, points, grid.element, grid.mesh.pitch))
        ne = len(grid.elem_ijk)
        npts = len(self.cb)
        if point_counted is not None:
            # per corner, fem.CORNERS order (constrained.py:349-358): bit k
            # of a cell's mask is its k-th sample point (x outermost)
            if points != "corners":
                raise ValueError("point_counted needs the corner samples")
            pc = np.asarray(point_counted, bool)
            if pc.shape != (ne, 8):
                raise ValueError("point_counted must be (n_elem, 8) bool")
            bits = np.zeros(ne, np.int64)
            for k, corner in enumerate(CORNER_OF):
                bits |= pc[:, corner].astype(np.int64) << k
            self.nc = float(max(int(pc.sum()), 1))
        else:
            cnt = np.ones(ne, bool) if counted is None else np.asarray(counted, bool)
            if cnt.shape != (ne,):
                raise ValueError("counted must be (n_elem,) bool")
            bits = np.where(cnt, (1 << npts) - 1, 0)
            self.nc = float(max(int(cnt.sum()), 1) * npts)
        self._counted = np.ascontiguousarray(grid.cell_grid(bits.astype(float))).astype(np.uint8)

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


class StressAveraged:
    """constrained._Averaged(mesh, nu, radius, readable) and the stress
    branch that reads it (constrained.py:517-605) on the device, for a
    HexMesh with no thin walls. bond: Limits.bond, (ratio, normal, k), or
    None. fixed: the dofs the adjoint right-hand side is zero on."""

    def __init__(self, mesh, nu: float, radius: float, readable=None, p: float = P_DEFAULT, q: float = Q_DEFAULT,
                 limit: float = 1.0, bond=None, element: str = "im", fixed=None):
        from .verify import _grid_of

        if element not in ("im", "plain"):
            raise ValueError(f"element must be 'im' or 'plain', not {element!r}")
        self.mesh = mesh
        self.nu, self.radius = float(nu), float(radius)
        self.p, self.q, self.limit = float(p), float(q), float(limit)
        self.element = element
        self.bond = None if bond is None else np.array([float(bond[0]), *map(float, bond[1]), float(bond[2])])
        self._dims, _, self._nodes, self._cells = _grid_of(mesh)
        nn = int(np.prod([d + 1 for d in self._dims]))
        rd = np.ones(mesh.n_nodes, bool) if readable is None else np.asarray(readable, bool)
        if rd.shape != (mesh.n_nodes,):
            raise ValueError("readable must be (n_nodes,) bool")
        self._readable = np.zeros(nn, np.uint8)
        self._readable[self._nodes] = rd
        self.fixed = None if fixed is None else np.asarray(fixed)

    def evaluate(self, us, xe, kappa=None, it: int = 1, need_grad: bool = True) -> list:
        """Per load case dict(g, pm, smax, kappa, rhs, gx): kappa is
        constrained._normalize of the previous one (kappa) at iteration it."""
        _lib.require_gpu()
        us = [np.asarray(u, float) for u in us]
        nr = len(us)
        nn = len(self._readable)
        ncell = int(np.prod(self._dims))
        ug = np.zeros((nr, nn, 3))
        for i, u in enumerate(us):
            ug[i, self._nodes] = u.reshape(-1, 3)
        xg = np.zeros(ncell)
        xg[self._cells] = np.asarray(xe, float)
        kap = np.ones(nr) if kappa is None else np.array(np.broadcast_to(np.asarray(kappa, float), (nr,)))
        pm, smax = np.empty(nr), np.empty(nr)
        rhs = np.empty((nr, nn, 3)) if need_grad else None
        gx = np.empty((nr, ncell)) if need_grad else None
        _lib.check(_lib.lib().lp_stress_averaged(
            *self._dims, float(self.mesh.pitch), self.nu, 0 if self.element == "im" else 1, self.radius, nr,
            _lib.ptr(ug), _lib.ptr(xg), _lib.ptr(self._readable), self.p, self.q, self.limit,
            _lib.ptr(self.bond) if self.bond is not None else None, _lib.ptr(kap), int(it), _lib.ptr(pm),
            _lib.ptr(smax), _lib.ptr(rhs) if need_grad else None, _lib.ptr(gx) if need_grad else None),
            "stress_averaged")
        out = []
        for i in range(nr):
            rec = dict(g=kap[i] * pm[i], pm=pm[i], smax=smax[i], kappa=kap[i])
            if need_grad:
                r = rhs[i, self._nodes].ravel()
                if self.fixed is not None:
                    r[self.fixed] = 0.0
                rec["rhs"] = r
                rec["gx"] = gx[i, self._cells]
            out.append(rec)
        return out

    def gradient(self, grid, us, xe, de, kappa=None, it: int = 1):
        """(records, dg/dxe per case): the adjoint solve on grid (a device
        GridElasticity over the same mesh), dg/dxe = explicit - de lambda_e^T
        Ke u_e (constrained.py:603-605)."""
        ev = self.evaluate(us, xe, kappa, it)
        lams = grid.solve_many([e["rhs"] for e in ev])
        de = np.asarray(de, float)
        return ev, [e["gx"] - de * grid.element_product(lam, u) for e, lam, u in zip(ev, lams, us)]
