"""The constrained minimum-mass loop on the device: loadpath's
constrained.optimize_constrained (constrained.py:295-854) for a box design
domain with no add-ons (no buckling, vibration, AM filter, thin walls,
mirror, extrusion, contact band or checkpointing).

Each iteration: the blueprint and eroded designs through the device filter
and projection, the batched multigrid solve of every load case, the
deflection constraints from the compliance and the element energies, the
stress constraints from the corner-sampled P-mean over the counted cells
(``StressPNorm``) or, with ``limits.radius``, the material-averaged reading
(``StressAveraged``), one batched adjoint solve, and the device MMA update
with the loop's move control, beta continuation and best-design tracking.
Everything above the kernels is the reference's control flow:

- E = E_MIN + xe^3 (1 - E_MIN) on the eroded design (constrained.py:467-481)
- deflection C / (sum_n |f_n| E_ref) <= limit (constrained.py:508-516)
- stress kappa pm - 1 <= 0 with Le's normalization (constrained.py:517-664)
- objective 10 mean(xp over design) (constrained.py:768-774)
- MMA move halved (>= 0.02) with the asymptotes tightened while the worst
  constraint grows above 0.02, grown by 1.2 to the cap once feasible
  (constrained.py:783-795)
- beta doubling once feasible (or after 3 beta_every), the cap times 0.6
  (constrained.py:811-823); the lightest feasible design at beta_max and
  the stall stop (constrained.py:824-833); the delivered blueprint is the
  best design's (constrained.py:836-848)
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .mma import MMA
from .simp import E_MIN, PENALTY, SimpOptions, SimpResult, density_filter, dproject, project
from .stress import StressAveraged, StressPNorm
from .ties import tie

STALL_GAIN = 5e-3     # constrained.STALL_GAIN (constrained.py:68)
HELD_SIZES = 1e3      # constrained.HELD_SIZES (constrained.py:93)
FEASIBLE = 0.02       # the slack the loop counts as inside the limits (constrained.py:785, 808)


@dataclass
class Limits:
    """constrained.Limits (constrained.py:116-135): stress the allowable
    peak von Mises (MPa), deflection at the loads (mm), modulus (MPa) to
    turn unit-E displacements into mm, bond (ratio, layer normal, k) of a
    part built in layers, radius the material length of the averaged
    reading, contact_band the distance from a loaded or supported node
    within which a corner reads no stress. buckling is not supported here."""
    stress: float | None
    deflection: float | None
    modulus: float
    buckling: float | None = None
    bond: tuple | None = None
    contact_band: float | None = None
    radius: float | None = None


class _Run:
    """One constrained study's device state (grid, filter, aggregates)."""

    def __init__(self, problem, nu, opts: SimpOptions, limits: Limits, preserve, precision, element, grid_kw,
                 autotune=False):
        from .mg import GridElasticity

        if getattr(problem, "present", None) is not None:
            raise NotImplementedError("the constrained loop here runs on a full box")
        if limits.buckling is not None:
            raise NotImplementedError("buckling limits are not supported")
        if limits.bond is not None and not limits.radius:
            raise NotImplementedError("the layer bond is read on the averaged stress (limits.radius)")
        self.opts, self.limits = opts, limits
        self.dims = tuple(int(d) for d in problem.dims)
        mesh = problem.mesh()
        self.mesh = mesh
        fixed, loads = problem.dof_vectors(mesh)
        self.fixed, self.loads = fixed, list(loads)
        n = mesh.n_elem
        self.n = n
        self.pres = np.zeros(n, bool) if preserve is None else np.asarray(preserve, bool).ravel()
        self.design = ~self.pres
        self.free = np.flatnonzero(self.design)
        self.h, self.hs = density_filter(self.dims, opts.filter_radius)
        self.eta_e = opts.eta + opts.erode
        kw = dict(rtol=opts.solve_rtol, maxiter=2000, precision=precision, element=element)
        kw.update(grid_kw)
        self.grid = GridElasticity(mesh, nu, fixed, **kw)
        self.load_size = [float(np.linalg.norm(f.reshape(-1, 3), axis=1).sum()) for f in self.loads]
        self.size = max(float(np.ptp(mesh.node_ijk * mesh.pitch, axis=0).max()), 1e-9)
        ncase = len(self.loads)
        self.defl_cases = [c for c in range(ncase) if limits.deflection is not None and self.load_size[c] > 0]
        self.stress_cases = list(range(ncase)) if limits.stress is not None else []
        if not self.defl_cases and not self.stress_cases:
            raise ValueError("optimize_constrained needs a stress or deflection limit")
        # stress is not read on the cells the loads and supports touch
        support = fixed.reshape(-1, 3).any(axis=1)
        loaded = np.zeros(mesh.n_nodes, bool)
        for f in self.loads:
            loaded |= f.reshape(-1, 3).any(axis=1)
        contact = (support | loaded)[mesh.enode].any(axis=1)
        # with a contact band: every corner within that distance of a loaded
        # or supported node reads no stress (constrained.py:349-358)
        pt_counted = None
        if limits.contact_band and (support | loaded).any():
            from scipy.spatial import cKDTree

            xyz = mesh.node_ijk * mesh.pitch
            near = cKDTree(xyz[support | loaded]).query(xyz, k=1, distance_upper_bound=float(limits.contact_band))[0] \
                < np.inf
            pt_counted = ~near[mesh.enode]
        self.agg = self.avg = None
        if self.stress_cases:
            if limits.radius:
                if pt_counted is not None:
                    readable = np.zeros(mesh.n_nodes, bool)
                    readable[mesh.enode[pt_counted]] = True
                    readable[mesh.enode[~pt_counted]] = False
                else:
                    readable = np.ones(mesh.n_nodes, bool)
                    readable[mesh.enode[contact].ravel()] = False
                self.avg = StressAveraged(mesh, nu, float(limits.radius), readable, limit=limits.stress,
                                          bond=limits.bond, element=element, fixed=np.flatnonzero(fixed))
            else:
                self.agg = StressPNorm(self.grid, "corners", limit=limits.stress,
                                       counted=None if pt_counted is not None else ~contact, point_counted=pt_counted)
        self.m = len(self.defl_cases) + len(self.stress_cases)
        self.autotune = autotune

    def designs(self, x, beta):
        xt = (self.h @ x) / self.hs
        xp = project(xt, beta, self.opts.eta)
        xp[self.pres] = 1.0
        xe = project(xt, beta, self.eta_e)
        xe[self.pres] = 1.0
        dpb = dproject(xt, beta, self.opts.eta)
        dpb[self.pres] = 0.0
        dpe = dproject(xt, beta, self.eta_e)
        dpe[self.pres] = 0.0
        return xp, xe, dpb, dpe

    def constraints(self, x, beta, kappa, it):
        xp, xe, dpb, dpe = self.designs(x, beta)
        e = E_MIN + xe ** PENALTY * (1 - E_MIN)
        de = PENALTY * xe ** (PENALTY - 1) * (1 - E_MIN)
        g = self.grid
        g.set_moduli(e)
        if self.autotune:
            # the fine operator's x-chunk for this grid, timed once (a small
            # grid wants more, shorter chunks than the default)
            g.autotune(len(self.loads))
            self.autotune = False
        us = g.solve_many(self.loads)
        lim = self.limits

        def to_x(g_xe):
            return self.h.T @ (g_xe * dpe / self.hs)

        cons, grads, rec = [], [], {}
        for c in self.defl_cases:
            comp = float(self.loads[c] @ us[c])
            defl = comp / (self.load_size[c] * lim.modulus)
            if not np.isfinite(defl) or defl > HELD_SIZES * self.size:
                raise ValueError(f"load case {c} is not held: deflection {defl:.3g} against a part {self.size:.3g}")
            ce = g.strain_energy(us[c])
            cons.append(defl / lim.deflection - 1.0)
            grads.append(to_x(-de * ce / (self.load_size[c] * lim.modulus)) / lim.deflection)
            rec[f"deflection.{c}"] = defl
        kappa = list(kappa)
        if self.stress_cases:
            sus = [us[c] for c in self.stress_cases]
            parts = []
            if self.avg is not None:
                ev = self.avg.evaluate(sus, xe, kappa=[kappa[c] for c in self.stress_cases], it=it)
                for c, r in zip(self.stress_cases, ev):
                    parts.append((c, r["g"], r["kappa"], r["smax"], r["rhs"], r["gx"]))
            else:
                # the partials are linear in kappa: read at kappa 1, then scale
                # by this iteration's normalization (constrained.py:642-653)
                ev = self.agg.evaluate(sus, xe, kappa=1.0)
                for i, c in enumerate(self.stress_cases):
                    pm, smax = float(ev["pm"][i]), float(ev["smax"][i])
                    kr = smax / max(pm, 1e-30)
                    k = 0.5 * (kappa[c] + kr) if it > 1 else kr
                    parts.append((c, k * pm, k, smax, k * ev["rhs"][i], k * ev["gx"][i]))
            lams = g.solve_many([p[4] for p in parts])
            for (c, gv, k, smax, _, gx), lam in zip(parts, lams):
                kappa[c] = k
                cons.append(gv - 1.0)
                grads.append(to_x(gx - de * g.element_product(lam, us[c])))
                rec[f"stress.{c}"] = smax * lim.stress
        rec["cgIterations"] = int(np.max(np.atleast_1d(g.last_iters)))
        return cons, grads, rec, kappa, xp, dpb


def optimize_constrained(problem, nu: float, opts: SimpOptions, limits: Limits, *, preserve=None,
                         precision: str = "fp64", element: str = "im", x0=None, on_iter=None,
                         autotune: bool = False, mirror=None, extrusion=None, **grid_kw) -> SimpResult:
    """constrained.optimize_constrained(domain, nu, opts, limits, on_iter)
    (constrained.py:276-854) for a box problem (a BoxProblem with no present
    mask): the lightest design within the stress and deflection limits.
    Returns a SimpResult (compliance nan, as the reference's) whose history
    records carry iteration, volume, beta, change, constraints, seconds,
    cgIterations, move and the per-case readings; ``.delivered`` is the
    record of the design delivered (None when no design at beta_max was
    inside the limits). autotune: time the fine operator's launch shape for
    the grid once, at the first solve (a speed setting only). mirror: the
    mirror planes through the box centre, as axes ("y", "xz"; domain.mirror
    on a box); extrusion: the axis of a 2.5D part (SimpOptions.extrusion);
    both tied as the reference ties them (constrained.py:775-780, 796-799;
    loadpath_amd.ties)."""
    _lib.require_gpu()
    t0 = time.time()
    R = _Run(problem, nu, opts, limits, preserve, precision, element, grid_kw, autotune)
    n, design, free = R.n, R.design, R.free
    x = np.full(n, opts.volume_fraction) if x0 is None else np.clip(np.asarray(x0, float).ravel(), 0.0, 1.0)
    x[R.pres] = 1.0
    mma = MMA(len(free), R.m, xmin=0.0, xmax=1.0, move=min(opts.move, 0.1))
    move_cap = mma.move
    worst_prev = np.inf
    kappa = [1.0] * len(R.loads)
    beta = float(opts.beta_start)
    since_beta = 0
    history, trail = [], []
    best = None
    converged = False
    it = 0
    g0 = 10.0 * design / design.sum()
    for it in range(1, opts.max_iterations + 1):
        x_now = x
        cons, grads, rec, kappa, xp, dpb = R.constraints(x, beta, kappa, it)
        f0 = 10.0 * float(xp[design].mean())
        df0 = R.h.T @ (g0 * dpb / R.hs)
        df0 = tie(df0, R.dims, mirror, extrusion, design)
        grads = [tie(gr, R.dims, mirror, extrusion, design) for gr in grads]
        worst = max(cons)
        if worst > FEASIBLE and worst > worst_prev:
            mma.move = max(0.02, 0.5 * mma.move)
            mma.tighten(0.5)
        elif worst <= FEASIBLE:
            mma.move = min(move_cap, 1.2 * mma.move)
        worst_prev = worst
        xnew = x.copy()
        xnew[free] = mma.update(x[free], f0, df0[free], np.array(cons), np.array([gr[free] for gr in grads]))
        xnew = tie(xnew, R.dims, mirror, extrusion, design)
        change = float(np.abs(xnew - x)[design].max())
        x = xnew
        since_beta += 1
        r = {"iteration": it, "volume": float(xp[design].mean()), "beta": beta, "change": change,
             "constraints": [float(v) for v in cons], "seconds": time.time() - t0, "move": mma.move, **rec}
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
    xp, _, _, _ = R.designs(x, beta)
    res = SimpResult(x=x, xphys=xp, compliance=float("nan"), volume_fraction=float(xp[design].mean()),
                     iterations=it, converged=converged, seconds=time.time() - t0, history=history)
    res.delivered = best[1] if best is not None else None
    return res
