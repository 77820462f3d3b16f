"""Oracle: the constrained minimum-mass loop on a box problem (FP64, numpy,
direct solves).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows loadpath/constrained.py `_constrained_run` (constrained.py:295-854)
for a box design domain with no add-ons (no buckling, vibration, AM filter,
thin walls, mirror, extrusion or checkpointing):
- blueprint xp = P(H x / hs, beta, eta), eroded xe = P(., beta, eta +
  erode), preserved cells 1; E = E_MIN + xe^3 (1 - E_MIN)  (:467-481)
- deflection per case: C / (sum_n |f_n| E_ref), its limit as defl / limit - 1
  <= 0, gradient -dE ce / (sum |f| E_ref) through the eroded chain  (:508-516)
- stress per case: the corner-sampled P-mean over the counted cells (cells
  touching no support or load node) or, with a radius, the material-averaged
  reading at the readable nodes (oracle/stress.py), each with Le's
  normalization and the adjoint solve, as g - 1 <= 0  (:517-664)
- objective 10 x mean(xp over design), through the blueprint chain  (:768-774)
- MMA on the design cells, move min(opts.move, 0.1): halved (>= 0.02) and the
  asymptotes tightened when the worst constraint exceeds 0.02 and grew;
  grown by 1.2 (<= the cap) when feasible  (:652-664, :783-795)
- beta doubling once feasible after beta_every iterations (or a change under
  2 tol), or after 3 beta_every regardless; the cap shrinks by 0.6  (:811-823)
- at beta_max: the lightest feasible design, a stop when feasible with a
  change under tol or when the best volume gained under STALL_GAIN over
  stall_window iterations  (:824-833); the delivered blueprint is the best
  design's (:836-848)

Pinned (tests/test_oracle_constrained.py) by its parts' own pins (the
stress readings against central differences and closed forms, MMA against
Svanberg's known answer), by the loop's whole constraint gradients against
central differences of the constraint functions through the direct solve,
and by its invariants on a small clamped bar (every limit met at the end,
beta only rising, the move never above its cap).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import stress as S
from .loop import E_MIN, PENALTY, LoopOptions, column_average, mirror_average
from .mma import MMA
from .simp import Evaluator, GridFilter, dproject, element_energy, project

STALL_GAIN = 5e-3     # constrained.py:68
HELD_SIZES = 1e3      # constrained.py:93
FEASIBLE = 0.02       # the slack the loop calls feasible (constrained.py:785, 808)


@dataclass
class Limits:
    """constrained.Limits (constrained.py:116-135), the fields read here."""
    stress: float | None
    deflection: float | None
    modulus: float
    bond: tuple | None = None
    radius: float | None = None
    contact_band: float | None = None


CORNER_OF = (0, 4, 2, 6, 1, 5, 3, 7)   # constrained._CORNER_OF (constrained.py:106)


def corner_cbs(nu, element="im", pitch=1.0):
    """C B / pitch at the 8 corners, x outermost (constrained.py:383-385)."""
    return S.sample_cb(nu, "corners", element, pitch)


class Constrained:
    """The loop's per-iteration pieces on one box problem."""

    def __init__(self, p, opts: LoopOptions, limits: Limits, nu=0.3, preserve=None, element="im"):
        self.p, self.opts, self.limits, self.nu = p, opts, limits, nu
        self.ev = Evaluator(p, element, nu)
        mesh = self.ev.mesh
        self.mesh = mesh
        n = mesh.n_elem
        self.n = n
        self.pres = np.zeros(n, bool) if preserve is None else np.asarray(preserve, bool).ravel()
        self.design = ~self.pres
        self.free = np.flatnonzero(self.design)
        self.filt = GridFilter(p.dims, opts.filter_radius, pad=opts.pad)
        self.hs = self.filt.hs
        self.eta_e = opts.eta + opts.erode
        self.element = element
        fs = self.ev.fs
        self.load_size = [float(np.linalg.norm(f.reshape(-1, 3), axis=1).sum()) for f in fs]
        self.size = max(float(np.ptp(mesh.node_ijk * mesh.pitch, axis=0).max()), 1e-9)
        self.defl_cases = [c for c in range(len(fs)) if limits.deflection is not None and self.load_size[c] > 0]
        self.stress_cases = list(range(len(fs))) if limits.stress is not None else []
        if not self.defl_cases and not self.stress_cases:
            raise ValueError("a stress or deflection limit is needed")
        support = self.ev.fixed.reshape(-1, 3).any(axis=1)
        loaded = np.zeros(mesh.n_nodes, bool)
        for f in fs:
            loaded |= f.reshape(-1, 3).any(axis=1)
        contact = (support | loaded)[mesh.enode].any(axis=1)
        self.counted = ~contact
        # with a contact band: every corner within that distance of a loaded
        # or supported node reads no stress (constrained.py:349-358)
        self.pt_counted = None
        if limits.contact_band and (support | loaded).any():
            from scipy.spatial import cKDTree

            xyz = mesh.node_ijk * mesh.pitch
            near = cKDTree(xyz[support | loaded]).query(xyz, k=1, distance_upper_bound=float(limits.contact_band))[0] \
                < np.inf
            self.pt_counted = ~near[mesh.enode]
            self.counted = self.pt_counted.any(axis=1)
        self.readable = None
        if limits.stress is not None and limits.radius:
            if self.pt_counted is not None:
                rd = np.zeros(mesh.n_nodes, bool)
                rd[mesh.enode[self.pt_counted]] = True
                rd[mesh.enode[~self.pt_counted]] = False
            else:
                rd = np.ones(mesh.n_nodes, bool)
                rd[mesh.enode[contact].ravel()] = False
            self.readable = rd
        elif limits.bond is not None:
            raise NotImplementedError("the layer bond is restated for the averaged reading only")
        self.cbs = corner_cbs(nu, element, mesh.pitch)
        self.m = len(self.defl_cases) + len(self.stress_cases)

    def designs(self, x, beta):
        xt = (self.filt @ x) / self.hs
        xp = project(xt, beta, self.opts.eta)
        xp[self.pres] = 1.0
        xe = project(xt, beta, self.eta_e)
        xe[self.pres] = 1.0
        dpb = dproject(xt, beta, self.opts.eta)
        dpb[self.pres] = 0.0
        dpe = dproject(xt, beta, self.eta_e)
        dpe[self.pres] = 0.0
        return xt, xp, xe, dpb, dpe

    def solve_k(self, e):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla

        ev = self.ev
        nd = self.mesh.n_dof
        data = (np.asarray(e, float)[:, None, None] * ev.ke[None]).ravel()
        k = sp.csr_matrix((data, (ev._rows, ev._cols)), shape=(nd, nd))
        return spla.splu(k[ev.free][:, ev.free].tocsc())

    def constraints(self, x, beta, kappa, it):
        """(cons, grads in x, records, kappa): every constraint of one
        iteration at design x (constrained.py:467-664)."""
        xt, xp, xe, dpb, dpe = self.designs(x, beta)
        e = E_MIN + xe ** PENALTY * (1 - E_MIN)
        de = PENALTY * xe ** (PENALTY - 1) * (1 - E_MIN)
        lu = self.solve_k(e)
        free = self.ev.free
        nd = self.mesh.n_dof
        us = []
        for f in self.ev.fs:
            u = np.zeros(nd)
            u[free] = lu.solve(f[free])
            us.append(u)
        lim = self.limits
        edof = self.ev.edof

        def to_x(g_xe):
            return self.filt.T @ (g_xe * dpe / self.hs)

        cons, grads, rec = [], [], {}
        for c in self.defl_cases:
            comp = float(self.ev.fs[c] @ us[c])
            defl = comp / (self.load_size[c] * lim.modulus)
            if not np.isfinite(defl) or defl > HELD_SIZES * self.size:
                raise ValueError("the loads are not held")
            ce = element_energy(us[c], edof, self.ev.ke)
            cons.append(defl / lim.deflection - 1.0)
            grads.append(to_x(-de * ce / (self.load_size[c] * lim.modulus)) / lim.deflection)
            rec[f"deflection.{c}"] = defl
        kappa = list(kappa)
        if self.stress_cases:
            parts = []
            if self.readable is not None:
                got = S.averaged_pnorm(self.mesh, self.nu, float(lim.radius), self.readable,
                                       [us[c] for c in self.stress_cases], xe, limit=lim.stress, bond=lim.bond,
                                       kappa_prev=[kappa[c] for c in self.stress_cases], it=it,
                                       element=self.element, fixed=np.flatnonzero(self.ev.fixed))
                for c, r in zip(self.stress_cases, got):
                    parts.append((c, r["g"], r["kappa"], r["smax"], r["rhs"], r["gx"]))
                    rec[f"pm.{c}"] = r["pm"]
            else:
                if self.pt_counted is None:
                    nc = max(int(self.counted.sum()), 1) * len(self.cbs)
                else:
                    nc = max(int(self.pt_counted.sum()), 1)
                act = np.flatnonzero(self.counted)
                for c in self.stress_cases:
                    moved = us[c].reshape(-1, 3).any(axis=1)[self.mesh.enode].any(axis=1)
                    a = act[moved[act]]
                    ue = us[c][edof[a]]
                    pmask = None if self.pt_counted is None else self.pt_counted[a][:, list(CORNER_OF)]
                    _, pm, smax, _ = S.pnorm_forward(ue, xe[a], self.cbs, lim=lim.stress, nc=nc, pmask=pmask)
                    k = S.normalize_kappa(kappa[c], smax, pm, it)
                    g, pm, smax, gx_a, rhs_e = S.pnorm_partials(ue, xe[a], self.cbs, lim=lim.stress, kappa=k, nc=nc,
                                                                pmask=pmask)
                    gx = np.zeros(self.n)
                    gx[a] = gx_a
                    rhs = S.assemble_rhs(edof[a], rhs_e, nd, self.ev.fixed)
                    parts.append((c, g, k, smax, rhs, gx))
                    rec[f"pm.{c}"] = pm
            for c, g, k, smax, rhs, gx in parts:
                lam = np.zeros(nd)
                lam[free] = lu.solve(rhs[free])
                kappa[c] = k
                cons.append(g - 1.0)
                grads.append(to_x(gx - de * S.element_product(lam, us[c], edof, self.ev.ke)))
                rec[f"stress.{c}"] = smax * lim.stress
        return cons, grads, rec, kappa, xp, dpb


def optimize_constrained(p, opts: LoopOptions, limits: Limits, nu=0.3, preserve=None, element="im", x0=None,
                         on_iter=None, mirror=None, extrusion=None):
    """The loop of constrained._constrained_run on a box problem: dict(x,
    xphys, volume_fraction, iterations, converged, history, delivered)."""
    L = Constrained(p, opts, limits, nu, preserve, element)
    n, design, free = L.n, L.design, L.free
    x = np.full(n, opts.volume_fraction) if x0 is None else np.clip(np.asarray(x0, float), 0.0, 1.0)
    x[L.pres] = 1.0
    mma = MMA(len(free), L.m, xmin=0.0, xmax=1.0, move=min(opts.move, 0.1))
    move_cap = mma.move
    worst_prev = np.inf
    kappa = [1.0] * len(L.ev.fs)
    beta = float(opts.beta_start)
    since_beta = 0
    history, trail = [], []
    best = None
    converged = False
    it = 0
    for it in range(1, opts.max_iterations + 1):
        x_now = x
        cons, grads, rec, kappa, xp, dpb = L.constraints(x, beta, kappa, it)
        f0 = 10.0 * float(xp[design].mean())
        g0 = 10.0 * design / design.sum()
        df0 = L.filt.T @ (g0 * dpb / L.hs)
        df0 = column_average(mirror_average(df0, p.dims, mirror), p.dims, extrusion, design)
        grads = [column_average(mirror_average(g, p.dims, mirror), p.dims, extrusion, design) for g in grads]
        worst = max(cons)
        if worst > FEASIBLE and worst > worst_prev:
            mma.move = max(0.02, 0.5 * mma.move)
            mma.tighten(0.5)
        elif worst <= FEASIBLE:
            mma.move = min(move_cap, 1.2 * mma.move)
        worst_prev = worst
        xnew = x.copy()
        xnew[free] = mma.update(x[free], f0, df0[free], np.array(cons), np.array([g[free] for g in grads]))
        xnew = column_average(mirror_average(xnew, p.dims, mirror), p.dims, extrusion, design)
        change = float(np.abs(xnew - x)[design].max())
        x = xnew
        since_beta += 1
        r = {"iteration": it, "volume": float(xp[design].mean()), "beta": beta, "change": change,
             "constraints": [float(v) for v in cons], "move": mma.move, **rec}
        history.append(r)
        if on_iter is not None:
            on_iter(r, xp)
        feasible = max(cons) <= FEASIBLE
        if beta < opts.beta_max and ((feasible and (since_beta >= opts.beta_every or change < 2 * opts.tol))
                                     or since_beta >= 3 * opts.beta_every):
            beta = min(beta * 2, opts.beta_max)
            since_beta = 0
            move_cap = max(0.02, move_cap * 0.6)
            mma.move = min(mma.move, move_cap)
        elif beta >= opts.beta_max:
            if feasible and (best is None or r["volume"] < best[1]["volume"]):
                best = (x_now, r)
            trail.append(best[1]["volume"] if best is not None else np.inf)
            w = opts.stall_window
            if best is not None and ((feasible and change < opts.tol)
                                     or (len(trail) > w and trail[-1] > (1.0 - STALL_GAIN) * trail[-1 - w])):
                converged = True
                break
    if best is not None:
        x = best[0]
    _, xp, _, _, _ = L.designs(x, beta)
    return dict(x=x, xphys=xp, volume_fraction=float(xp[design].mean()), iterations=it, converged=converged,
                history=history, delivered=best[1] if best is not None else None)
