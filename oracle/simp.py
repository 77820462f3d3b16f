"""Oracle: density filters, sensitivities and the top3d OC loop (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates:
- top3d/top88 filter: reach ceil(rmin)-1, cone weights rmin - dist > 0,
  row-sum normalised                               benchmarks/reference.py:188-211
- loadpath GridFilter stencil: reach floor(r/pitch), weights r - dist*pitch,
  zero outside the grid, hs = full kernel sum (pad) or row sums
                                                   simp.py:1143-1189
- evaluator: assembled sparse K = sum E_c Ke, free-dof direct solve
                                                   benchmarks/evaluate.py:88-117
- the top3d OC loop (SIMP p=3, Emin 1e-9, density filter, OC bisection
  l2 = 1e9 / rel. tol 1e-3, stop change <= 0.01, at most 200 iterations)
                                                   benchmarks/reference.py:214-257
- textbook problems cantilever_3d, bridge_3d        benchmarks/problems.py:107-136

Pinned by (tests/test_oracle_simp.py):
- the reference's golden designs benchmarks/reference/cantilever-3d.top3d.npz
  (133 iterations, c = 2418.152775040522, x grid) reproduced, committed here as
  tests/golden/cantilever-3d.top3d.npz (data only);
- filter: brute-force dense H against the stencil, tap counts 19/27/81 for
  r = 1.5/2.0/2.5, symmetry of H;
- sensitivities: central finite differences of f.u(E) on tiny grids.
"""
from __future__ import annotations

import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .fem import box_mesh, hex8_ke

E_VOID = 1e-9
NU = 0.3


# --------------------------------------------------------------------------
# filters
# --------------------------------------------------------------------------

def cone_taps(radius: float, reach: int, pitch: float = 1.0):
    """[((dx, dy, dz), w)] with w = radius - |d| pitch > 0, |d_i| <= reach."""
    taps = []
    for dx in range(-reach, reach + 1):
        for dy in range(-reach, reach + 1):
            for dz in range(-reach, reach + 1):
                w = radius - np.sqrt(dx * dx + dy * dy + dz * dz) * pitch
                if w > 0:
                    taps.append(((dx, dy, dz), float(w)))
    return taps


def top3d_filter(dims, rmin: float, active=None):
    """(H csr, Hs) over the cells of `dims` (C order), reach ceil(rmin)-1
    (benchmarks/reference.py:188-211)."""
    dims = tuple(int(d) for d in dims)
    ii, jj, kk = np.meshgrid(*(np.arange(d) for d in dims), indexing="ij")
    ijk = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)
    if active is not None:
        ijk = ijk[np.asarray(active, bool).ravel()]
    lut = np.full(dims, -1, np.int64)
    lut[tuple(ijk.T)] = np.arange(len(ijk))
    reach = int(np.ceil(rmin)) - 1
    rows, cols, vals = [], [], []
    for (dx, dy, dz), w in cone_taps(rmin, reach):
        if (dx and dims[0] == 1) or (dy and dims[1] == 1) or (dz and dims[2] == 1):
            continue
        nb = ijk + np.array([dx, dy, dz])
        ok = np.all((nb >= 0) & (nb < np.array(dims)), axis=1)
        j = np.full(len(ijk), -1, np.int64)
        j[ok] = lut[tuple(nb[ok].T)]
        hit = j >= 0
        rows.append(np.flatnonzero(hit))
        cols.append(j[hit])
        vals.append(np.full(int(hit.sum()), w))
    n = len(ijk)
    h = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return h, np.asarray(h.sum(axis=1)).ravel()


class GridFilter:
    """loadpath's stencil filter (simp.py:1143-1189): h @ v over the cells
    ijk (default: the full box) of a box of dims cells of side pitch, zero
    outside them; hs = full kernel weight (pad) or row sums."""

    def __init__(self, dims, radius: float, pitch: float = 1.0, pad: bool = True, ijk=None):
        self.dims = tuple(int(d) for d in dims)
        self.reach = int(np.floor(radius / pitch))
        self.taps = cone_taps(radius, self.reach, pitch)
        self.T = self
        self.ijk = None if ijk is None else tuple(np.asarray(ijk, np.int64).T)
        n = int(np.prod(self.dims)) if ijk is None else len(ijk)
        self.hs = np.full(n, sum(w for _, w in self.taps)) if pad else self @ np.ones(n)

    def __matmul__(self, v):
        r = self.reach
        nx, ny, nz = self.dims
        g = np.zeros((nx + 2 * r, ny + 2 * r, nz + 2 * r))
        inner = g[r:r + nx, r:r + ny, r:r + nz]
        if self.ijk is None:
            inner[...] = np.asarray(v, float).reshape(self.dims)
        else:
            inner[self.ijk] = v
        out = np.zeros(self.dims)
        for (dx, dy, dz), w in self.taps:
            out += w * g[r + dx:r + dx + nx, r + dy:r + dy + ny, r + dz:r + dz + nz]
        return out.ravel() if self.ijk is None else out[self.ijk]


# --------------------------------------------------------------------------
# problems and the assembled evaluator
# --------------------------------------------------------------------------

def nodes(i, j, k):
    ii, jj, kk = np.meshgrid(np.atleast_1d(i), np.atleast_1d(j), np.atleast_1d(k), indexing="ij")
    return np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.int64)


class Problem:
    """A box problem: fixed [(nodes (N,3), "xyz")], loads [[(nodes, comp, value)], ...]
    (one list per load case)."""

    def __init__(self, name, dims, volfrac, rmin, fixed, loads, void=None):
        self.name = name
        self.dims = tuple(int(d) for d in dims)
        self.volfrac = float(volfrac)
        self.rmin = float(rmin)
        self.fixed = fixed
        self.loads = loads if loads and isinstance(loads[0], list) else [loads]
        self.void = void

    @property
    def active(self):
        a = np.ones(self.dims, bool)
        if self.void is not None:
            a &= ~self.void
        return a


def cantilever_3d(nx=60, ny=4, nz=20, volfrac=0.3, rmin=1.5):
    """top3d's default (benchmarks/problems.py:107-119)."""
    return Problem("cantilever-3d", (nx, ny, nz), volfrac, rmin,
                   [(nodes(0, range(ny + 1), range(nz + 1)), "xyz")],
                   [(nodes(nx, range(ny + 1), 0), "z", -1.0)])


def bridge_3d(nx=48, ny=12, nz=16, volfrac=0.15, rmin=1.5):
    """benchmarks/problems.py:122-136."""
    top = nodes(range(nx + 1), range(ny + 1), nz)
    return Problem("bridge-3d", (nx, ny, nz), volfrac, rmin,
                   [(nodes([0, nx], range(ny + 1), 0), "xyz")],
                   [(top, "z", -100.0 / len(top))])


_AX = {"x": 0, "y": 1, "z": 2}


class Evaluator:
    """Assembled FP64 solve on a full box (benchmarks/evaluate.py:88-117)."""

    def __init__(self, p: Problem, element="im", nu=NU):
        self.p = p
        self.mesh = box_mesh(*p.dims)
        self.ke = hex8_ke(nu, element)
        self.edof = self.mesh.edof
        n = self.mesh.n_dof
        fixed = np.zeros(n, bool)
        for nd, comps in p.fixed:
            for c in comps:
                fixed[3 * self.mesh.node_id(nd) + _AX[c]] = True
        self.fixed = fixed
        self.free = np.flatnonzero(~fixed)
        self.fs = []
        for case in p.loads:
            f = np.zeros(n)
            for nd, c, v in case:
                np.add.at(f, 3 * self.mesh.node_id(nd) + _AX[c], v)
            self.fs.append(f)
        self.f = self.fs[0]
        self._rows = np.repeat(self.edof, 24, axis=1).ravel()
        self._cols = np.tile(self.edof, (1, 24)).ravel()

    def solve(self, e, f=None):
        f = self.f if f is None else f
        data = (np.asarray(e, float)[:, None, None] * self.ke[None]).ravel()
        n = self.mesh.n_dof
        k = sp.csr_matrix((data, (self._rows, self._cols)), shape=(n, n))
        kf = k[self.free][:, self.free]
        u = np.zeros(n)
        u[self.free] = spla.spsolve(kf.tocsc(), f[self.free])
        return u

    def compliance(self, x, penal=3.0):
        u = self.solve(E_VOID + np.asarray(x, float).ravel() ** penal * (1 - E_VOID))
        return float(self.f @ u)


# --------------------------------------------------------------------------
# sensitivities and OC
# --------------------------------------------------------------------------

def element_energy(u, edof, ke):
    """ce_c = u_e^T Ke u_e (femops.py:236-242 / reference.py:230)."""
    ue = u[edof]
    return np.einsum("ij,jk,ik->i", ue, ke, ue)


def simp_modulus(xphys, penal=3.0, emin=E_VOID):
    return emin + xphys ** penal * (1 - emin)


def compliance_sensitivity(xphys, ce, penal=3.0, emin=E_VOID):
    """dc = -p (1 - Emin) x^(p-1) ce (reference.py:232)."""
    return -penal * (1 - emin) * xphys ** (penal - 1) * ce


def oc_bisect(x, dc, dv, h, hs, volfrac, move=0.2, passive=None, l1=0.0, l2=1e9, tol=1e-3):
    """OC update with the volume multiplier by bisection (reference.py:240-250).
    Returns (xnew, xphys, n_evaluations)."""
    n = len(x)
    evals = 0
    while (l2 - l1) / (l1 + l2) > tol:
        lmid = 0.5 * (l2 + l1)
        xnew = np.maximum(0.0, np.maximum(x - move, np.minimum(1.0, np.minimum(x + move, x * np.sqrt(-dc / dv / lmid)))))
        if passive is not None:
            xnew[passive] = 0.0
        xphys = (h @ xnew) / hs
        if passive is not None:
            xphys = np.where(passive, 0.0, xphys)
        evals += 1
        if xphys.sum() > volfrac * n:
            l1 = lmid
        else:
            l2 = lmid
    return xnew, xphys, evals


def project(xt, beta, eta):
    """Heaviside projection (simp._project, simp.py:1443-1446)."""
    den = np.tanh(beta * eta) + np.tanh(beta * (1 - eta))
    return (np.tanh(beta * eta) + np.tanh(beta * (np.asarray(xt, float) - eta))) / den


def dproject(xt, beta, eta):
    """Its derivative (simp._dproject, simp.py:1448-1450)."""
    den = np.tanh(beta * eta) + np.tanh(beta * (1 - eta))
    return beta * (1 - np.tanh(beta * (np.asarray(xt, float) - eta)) ** 2) / den


def optimize_mma(p: Problem, iterations, solve=None, penal=3.0, move=0.2, element="im", filt=None):
    """Compliance with the volume constraint under MMA: f0 = 10 c / c_first,
    con = sum xphys / (vf n) - 1, both gradients through the filter (the
    normalization of loadpath's MMA loops on compliance-type objectives,
    thermal.py:1035-1043), xphys = H x / hs. Returns the history of
    (compliance, volume, change) and the final (x, xphys)."""
    from .mma import MMA

    ev = Evaluator(p, element)
    if solve is None:
        solve = lambda e: [ev.solve(e, f) for f in ev.fs]
    h, hs = filt if filt is not None else top3d_filter(p.dims, p.rmin)
    n = len(hs)
    x = np.full(n, p.volfrac)
    xphys = x.copy()
    mma = MMA(n, 1, move=move)
    c_ref = None
    hist = []
    for _ in range(iterations):
        e = simp_modulus(xphys, penal)
        us = solve(e)
        ce = sum(element_energy(u, ev.edof, ev.ke) for u in us)
        c = float(np.sum(e * ce))
        dc = h @ (compliance_sensitivity(xphys, ce, penal) / hs)
        dv = h @ (np.ones(n) / hs)
        if c_ref is None:
            c_ref = abs(c) or 1.0
        vol = float(xphys.sum())
        con = vol / (p.volfrac * n) - 1.0
        xnew = mma.update(x, 10.0 * c / c_ref, 10.0 * dc / c_ref, np.array([con]), (dv / (p.volfrac * n))[None])
        xphys = (h @ xnew) / hs
        change = float(np.abs(xnew - x).max())
        x = xnew
        hist.append((c, vol / n, change))
    return hist, x, xphys


def optimize_top3d(p: Problem, solve=None, maxit=200, iterations=None, penal=3.0, move=0.2,
                   element="im", on_iter=None, filt=None, on_design=None):
    """The top3d loop (benchmarks/reference.py:214-257) for the 3-D case.

    solve(e) -> list of displacement vectors (mesh dof numbering), one per
    load case; defaults to the assembled direct solve. `iterations` runs a
    fixed count instead of the change <= 0.01 stop (the configs' "N OC
    iterations"). filt: (h, hs) or None for the top3d filter. on_design(loop,
    xphys): the filtered design after each iteration."""
    ev = Evaluator(p, element)
    if solve is None:
        solve = lambda e: [ev.solve(e, f) for f in ev.fs]
    h, hs = filt if filt is not None else top3d_filter(p.dims, p.rmin)
    n = len(hs)
    passive = ~p.active.ravel()
    x = np.full(n, p.volfrac)
    x[passive] = 0.0
    xphys = x.copy()
    t0 = time.time()
    change, loop, c = 1.0, 0, 0.0
    hist = []
    while (iterations is None and change > 0.01 and loop < maxit) or (iterations is not None and loop < iterations):
        loop += 1
        e = simp_modulus(xphys, penal)
        us = solve(e)
        ce = sum(element_energy(u, ev.edof, ev.ke) for u in us)
        c = float(np.sum(e * ce))
        dc = compliance_sensitivity(xphys, ce, penal)
        dv = np.ones(n)
        dc = h @ (dc / hs)
        dv = h @ (dv / hs)
        xnew, xphys, _ = oc_bisect(x, dc, dv, h, hs, p.volfrac, move, passive if passive.any() else None)
        change = float(np.abs(xnew - x).max())
        x = xnew
        hist.append((loop, c, change))
        if on_iter is not None:
            on_iter(loop, c, change)
        if on_design is not None:
            on_design(loop, xphys)
    return dict(x=xphys.reshape(p.dims), iterations=loop, compliance_reported=c, seconds=time.time() - t0,
                history=hist)
