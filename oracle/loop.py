"""Oracle: loadpath's production SIMP loop on a box (FP64, numpy).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates simp._optimize_run (simp.py:1640-2449) for a box domain with a
preserved-solid mask and the plain compliance objective (no overhang
filter, heat, gravity, buckling, blockage or member limits; those branches
are inactive there):
- blueprint xp = P(H x / hs, beta, eta) and analysed (eroded) design
  xe = P(H x / hs, beta, eta + erode), preserved cells 1, void cells 0
                                                   simp.py:1775-1812
- E = E_MIN + xe^3 (1 - E_MIN), E_MIN = 1e-6       simp.py:42, 55, 1861
- c = sum_i w_i c_i, dc_xp = -sum_i w_i 3 xe^2 (1 - E_MIN) ce_i
                                                   simp.py:1897-1904
- dc = H^T (dc_xp dP(xt, eta_e) / hs), dv = H^T (w dP(xt, eta) / hs),
  dP zero on preserved cells; dc <= 0, dv >= 1e-12
                                                   simp.py:1994-2001, 2077, 2190-2204
- the measured step: along the last step dx the Lagrangian's slope was
  a = g_old . dx and is b = g_new . dx; the limit scales by
  clip(a / (a - b), 1/2, 6/5), at most opts.move  simp.py:2228-2245, 1238-1239
- the OC step x clip(sqrt(-dc / dv / lambda)) with the volume multiplier
  found by outward bracketing in log lambda from the last one, then
  Illinois false position to |V - target| <= 1e-13 target
                                                   simp.py:2246-2381, 1235
- projection continuation: beta doubles when change < 2 tol on a free step
  or after beta_every iterations at that beta, the limit halves; at
  beta_max it stops on change < tol (free step) or, the scheduled stop,
  after beta_every iterations with the volume bracketed
                                                   simp.py:2413-2449, 1436-1440
- mirror planes and extrusion columns: the clamped dc and dv and every
  trial design of the volume root are tied (mirror_average, then
  column_average; void cells emptied again after the columns)
                                                   simp.py:2223-2226, 2288-2304
- the analysis' live-cell reduction: alive = xe > ALIVE or an anchor cell
  (touching a supported or loaded node, or preserved); under REDUCE_BELOW
  of the cells alive, only the live cells 26-connected to a live cell on a
  support are solved (the others carry no stiffness); a load on a dof they
  do not hold, or a singular reduced system, solves the whole grid
                                                   simp.py:59-61, 154-155, 264-282, 720-826

Pinned by tests/test_oracle_loop.py:
- the chain rule (dc, dv through filter, projection and erosion) against
  central finite differences of the eroded compliance and the blueprint
  volume on a tiny grid;
- the volume root against a dense scan plus bisection in log lambda (V is
  monotone in lambda), and |V - target| <= 1e-13 target at the root;
- the measured step on a quadratic, where a / (a - b) is the exact line
  minimum;
- the loop's invariants on a small cantilever: the volume on target to
  1e-10 once bracketed, the limit never above opts.move, beta doubling on
  the schedule, compliance falling at the sharpest stage;
- the ties against the reference's index-map formulation, to the bit;
- the reduction: the kept cells are scipy.ndimage.label's components
  (the reference's own call) that touch a support, and the reduced solve
  equals the full solve of the system whose dropped cells have modulus 0
  (an island carries no load: its displacement is zero).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .simp import Evaluator, GridFilter, dproject, element_energy, project

E_MIN = 1e-6          # simp.py:42
ALIVE = 1e-3          # simp.py:59
REDUCE_BELOW = 0.8    # simp.py:61
PENALTY = 3.0         # simp.py:55
VOLUME_RTOL = 1e-13   # simp.py:1235
STEP_SHRINK = 0.5     # simp.py:1238
STEP_GROW = 1.2       # simp.py:1239


@dataclass
class LoopOptions:
    """The SimpOptions fields the plain box loop reads (simp.py:899-1000)."""
    volume_fraction: float
    filter_radius: float
    max_iterations: int = 120
    beta_max: float = 32.0
    beta_every: int = 10
    beta_schedule: tuple | None = None
    move: float = 0.2
    adaptive_move: bool = True
    tol: float = 0.01
    stall_rtol: float = 2e-3
    stall_window: int = 6
    eta: float = 0.5
    erode: float = 0.15
    beta_start: float = 1.0
    pad: bool = True
    mirror: str = ""              # mirror planes through the box centre (domain.mirror)
    extrusion: int | None = None  # extrusion axis (SimpOptions.extrusion)
    reduce: bool = True           # the multigrid path's live-cell reduction (_solve_mg)


def mirror_average(v, dims, axes):
    """simp._symmetrize (simp.py:1476-1482) on a box: for each mirror axis
    in turn, every cell and its image across the centre plane both take
    0.5 (v + v_image) (domain.mirror, domain.py:1981-1988)."""
    out = np.array(v, float).reshape(dims)
    for ax in axes or ():
        a = "xyz".index(ax) if isinstance(ax, str) else int(ax)
        nxt = out.copy()
        for k in range(dims[a]):
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[a], hi[a] = k, dims[a] - 1 - k
            nxt[tuple(lo)] = 0.5 * (out[tuple(lo)] + out[tuple(hi)])
        out = nxt
    return out.reshape(np.shape(v))


def column_average(v, dims, axis, design):
    """simp._extrude (simp.py:1453-1463) with manufacture.tie_groups's
    columns along ``axis`` (manufacture.py:413-427) on a box: each column's
    design cells take their mean, summed cell by cell along the axis (the
    order np.bincount adds a column's cells in)."""
    if axis is None:
        return v
    a = "xyz".index(axis) if isinstance(axis, str) else int(axis)
    g = np.array(v, float).reshape(dims)
    d = np.asarray(design, bool).reshape(dims)
    gm = np.moveaxis(g, a, 0)
    dm = np.moveaxis(d, a, 0)
    total = np.zeros(gm.shape[1:])
    count = np.zeros(gm.shape[1:])
    for k in range(gm.shape[0]):
        total = total + np.where(dm[k], gm[k], 0.0)
        count = count + dm[k]
    mean = total / np.maximum(count, 1)
    for k in range(gm.shape[0]):
        gm[k] = np.where(dm[k], mean, gm[k])
    return np.moveaxis(gm, 0, a).reshape(np.shape(v))


def beta_hold(opts: LoopOptions, beta: float) -> int:
    """Most iterations at projection sharpness beta (simp.py:1436-1440)."""
    s = opts.beta_schedule
    k = int(round(np.log2(max(beta, 1.0))))
    return s[k] if s is not None and k < len(s) else opts.beta_every


def step_fraction(a_s: float, b_s: float) -> float:
    """Where along the last step the parabola through the two slopes has
    its minimum: a / (a - b), infinite when the slope did not rise
    (simp.py:2237-2241). Only meaningful for a < 0."""
    return a_s / (a_s - b_s) if b_s > a_s else np.inf


def volume_root(stepped, target: float, lam_last):
    """The multiplier at which the stepped design's volume meets the target
    (simp.py:2303-2373): bracket outward in log lambda from the last one,
    reach x4 per step, then Illinois false position. stepped(lambda) ->
    (x_new, V - target). Returns (x_new, lambda, f, f1, f2, evaluations)."""
    n_eval = [0]

    def ev(lm):
        n_eval[0] += 1
        return stepped(lm)

    l1, l2 = 1e-12, 1e12
    s_lo, s_hi = np.log(l1), np.log(l2)
    sm = min(max(np.log(lam_last), s_lo), s_hi) if lam_last else 0.0
    x_new, f = ev(float(np.exp(sm)))
    lm = float(np.exp(sm))
    s1 = s2 = sm
    f1 = f2 = None
    if f > 0:
        f1 = f
    else:
        f2 = f
    reach = 0.25
    while (f1 is None or f2 is None) and f != 0.0:
        if f1 is None:
            if s1 <= s_lo:
                break
            s1 = max(s1 - reach, s_lo)
            sm = s1
        else:
            if s2 >= s_hi:
                break
            s2 = min(s2 + reach, s_hi)
            sm = s2
        reach *= 4.0
        lm = float(np.exp(sm))
        x_new, f = ev(lm)
        if f > 0:
            f1, s1 = f, sm
        else:
            f2, s2 = f, sm
    if f1 is not None and f2 is not None:
        side = 0
        for _ in range(60):
            if abs(f) <= VOLUME_RTOL * target or f2 == f1 or s2 - s1 <= 4e-16 * max(abs(s1), abs(s2), 1.0):
                break
            sm = (s1 * f2 - s2 * f1) / (f2 - f1)
            if not s1 < sm < s2:
                sm = 0.5 * (s1 + s2)
            lm = float(np.exp(sm))
            x_new, f = ev(lm)
            if f > 0:
                s1, f1 = sm, f
                if side == 1:
                    f2 *= 0.5
                side = 1
            else:
                s2, f2 = sm, f
                if side == -1:
                    f1 *= 0.5
                side = -1
    return x_new, lm, f, f1, f2, n_eval[0]


class Loop:
    """The loop's per-iteration pieces on one problem (filter, physical and
    eroded designs, sensitivities), shared by the loop and the tests."""

    def __init__(self, p, opts: LoopOptions, preserve=None, weights=None, solve=None, element="im", void=None):
        self.p, self.opts = p, opts
        self.ev = Evaluator(p, element)
        n = self.ev.mesh.n_elem
        self.n = n
        self.pres = np.zeros(n, bool) if preserve is None else np.asarray(preserve, bool).ravel()
        # void cells (opts.void): held at 0, still counted as design cells in
        # the volume target (simp.py:1671-1682)
        self.void = np.zeros(n, bool) if void is None else np.asarray(void, bool).ravel()
        self.design = ~self.pres
        self.wdesign = self.design.astype(float)
        self.target = opts.volume_fraction * float(self.wdesign.sum())
        self.filt = GridFilter(p.dims, opts.filter_radius, pad=opts.pad)
        self.hs = self.filt.hs
        self.weights = np.ones(len(self.ev.fs)) if weights is None else np.asarray(weights, float)
        self.solve = solve if solve is not None else (lambda e: [self.ev.solve(e, f) for f in self.ev.fs])
        self.eta_e = opts.eta + opts.erode
        # cells touching a supported node, and the anchors: also touching a
        # loaded node, or preserved (simp.py:154-155)
        enode = self.ev.mesh.enode
        fixed = self.ev.fixed.reshape(-1, 3).any(axis=1)
        loaded = np.zeros(len(fixed), bool)
        for f in self.ev.fs:
            loaded |= (f.reshape(-1, 3) != 0).any(axis=1)
        self.support_touch = fixed[enode].any(axis=1)
        self.anchor = self.support_touch | loaded[enode].any(axis=1) | self.pres
        self.reduced_solves = 0
        self.full_solves = 0

    def keep_cells(self, xe):
        """The cells the multigrid path keeps (simp._Analysis._solve_mg and
        _connected, simp.py:264-282, 795-807), or None for the whole grid."""
        from scipy import ndimage

        alive = (xe > ALIVE) | self.anchor
        if alive.mean() >= REDUCE_BELOW:
            return None
        g = alive.reshape(self.p.dims)
        lbl, _ = ndimage.label(g, structure=np.ones((3, 3, 3), bool))
        el = lbl.ravel()
        roots = np.unique(el[alive & self.support_touch])
        return alive & np.isin(el, roots[roots > 0])

    def reduced_solve(self, e, keep):
        """Direct solve of the kept cells' system (their dofs off the
        supports; every other dof zero), or None where it cannot stand in for
        the whole grid: a load on a dof the kept cells do not hold, or a
        singular system (a hinge)."""
        import scipy.sparse.linalg as spla
        import warnings

        ev = self.ev
        held = np.zeros(ev.mesh.n_nodes, bool)
        held[ev.mesh.enode[keep].ravel()] = True
        live = np.repeat(held, 3)
        if any(np.any(f[~live]) for f in ev.fs):
            return None
        ek = np.where(keep, e, 0.0)
        data = (ek[:, None, None] * ev.ke[None]).ravel()
        import scipy.sparse as sp

        n = ev.mesh.n_dof
        k = sp.csr_matrix((data, (ev._rows, ev._cols)), shape=(n, n))
        free = np.flatnonzero(live & ~ev.fixed)
        kf = k[free][:, free].tocsc()
        out = []
        with warnings.catch_warnings():
            warnings.simplefilter("error")
            try:
                for f in ev.fs:
                    u = np.zeros(n)
                    u[free] = spla.spsolve(kf, f[free])
                    if not np.all(np.isfinite(u)):
                        return None
                    out.append(u)
            except Exception:
                return None
        return out

    def solve_design(self, e, xe):
        keep = self.keep_cells(xe) if self.opts.reduce else None
        if keep is not None:
            us = self.reduced_solve(e, keep)
            if us is not None:
                self.reduced_solves += 1
                return us, np.where(keep, e, 0.0)
        self.full_solves += 1
        return self.solve(e), e

    def physical(self, x, beta):
        xt = (self.filt @ x) / self.hs
        xp = project(xt, beta, self.opts.eta)
        xp[self.pres] = 1.0
        xp[self.void] = 0.0
        return xt, xp

    def eroded(self, xt, beta):
        xe = project(xt, beta, self.eta_e)
        xe[self.pres] = 1.0
        xe[self.void] = 0.0
        return xe

    def analyse(self, x, beta):
        """compliance, dc, dv, xp, cases' displacements (simp.py:1860-2204)."""
        xt, xp = self.physical(x, beta)
        xe = self.eroded(xt, beta)
        e = E_MIN + xe ** PENALTY * (1 - E_MIN)
        us, _ = self.solve_design(e, xe)
        comp = 0.0
        dc_xp = np.zeros(self.n)
        for w, u, f in zip(self.weights, us, self.ev.fs):
            ce = element_energy(u, self.ev.edof, self.ev.ke)
            comp += w * float(f @ u)
            dc_xp -= w * PENALTY * xe ** (PENALTY - 1) * (1 - E_MIN) * ce
        dproj_e = dproject(xt, beta, self.eta_e)
        dproj_e[self.pres] = 0.0
        dproj_e[self.void] = 0.0
        dproj = dproject(xt, beta, self.opts.eta)
        dproj[self.pres] = 0.0
        dproj[self.void] = 0.0
        dc = self.filt.T @ (dc_xp * dproj_e / self.hs)
        dv = self.filt.T @ (self.wdesign * dproj / self.hs)
        return comp, dc, dv, xp, us


def optimize_loop(p, opts: LoopOptions, preserve=None, weights=None, solve=None, element="im", on_iter=None,
                  void=None):
    """simp._optimize_run's loop (simp.py:1840-2449) on a box. Returns
    dict(x, xphys, history, converged, iterations)."""
    t0 = time.time()
    L = Loop(p, opts, preserve, weights, solve, element, void)
    x = np.full(L.n, opts.volume_fraction)
    x[L.pres] = 1.0
    x[L.void] = 0.0
    beta = float(opts.beta_start)
    since_beta = 0
    measured = opts.adaptive_move
    scheduled = measured
    step_prev = None
    move_k = opts.move
    lam_last = None
    c0 = None
    history = []
    converged = False
    it = 0
    for it in range(1, opts.max_iterations + 1):
        sec = None
        comp, dc, dv, xp, _ = L.analyse(x, beta)
        if c0 is None:
            c0 = comp
        dc = np.minimum(dc, 0.0)
        dv = np.maximum(dv, 1e-12)
        dc = column_average(mirror_average(dc, p.dims, opts.mirror), p.dims, opts.extrusion, L.design)
        dv = column_average(mirror_average(dv, p.dims, opts.mirror), p.dims, opts.extrusion, L.design)
        if measured:
            if step_prev is not None:
                dxp, gp, lam_p = step_prev
                a_s = float(gp @ dxp)
                b_s = float((dc + lam_p * dv) @ dxp)
                if a_s < 0:
                    t_s = step_fraction(a_s, b_s)
                    sec = t_s
                    move_k = min(move_k * min(max(t_s, STEP_SHRINK), STEP_GROW), opts.move)
            move = move_k
        else:
            move = opts.move if beta <= 4 else opts.move / 2
        lo_x = np.maximum(0.0, x - move)
        hi_x = np.minimum(1.0, x + move)
        ratio = np.sqrt(-dc / dv)

        def stepped(lm_):
            xn = np.clip(x * ratio / np.sqrt(lm_), lo_x, hi_x)
            xn[L.pres] = 1.0
            xn[L.void] = 0.0
            if opts.mirror or opts.extrusion is not None:
                xn = column_average(mirror_average(xn, p.dims, opts.mirror), p.dims, opts.extrusion, L.design)
                xn[L.void] = 0.0
            v = float(L.physical(xn, beta)[1] @ L.wdesign)
            return xn, v - L.target

        x_new, lm, _, f1, f2, n_eval = volume_root(stepped, L.target, lam_last)
        lam_last = lm
        vol_held = f1 is not None and f2 is not None
        change = float(np.abs(x_new - x)[L.design].max())
        if measured:
            step_prev = (x_new - x, dc + lm * dv, lm)
        x = x_new
        since_beta += 1
        rec = {"iteration": it, "compliance": comp, "relative": comp / c0 if c0 else 1.0,
               "volume": float(xp[L.design].mean()), "beta": beta, "change": change,
               "seconds": time.time() - t0, "volumeSolves": n_eval, "lambda": lm, "volHeld": vol_held}
        if opts.adaptive_move:
            rec["move"] = move
            if sec is not None:
                rec["stepFit"] = float(min(sec, 99.0))
        history.append(rec)
        if on_iter is not None:
            on_iter(rec)
        free_step = not measured or change < 0.999 * move
        if beta < opts.beta_max and ((change < opts.tol * 2 and free_step) or since_beta >= beta_hold(opts, beta)):
            beta = min(beta * 2, opts.beta_max)
            since_beta = 0
            step_prev = None
            move_k = move_k / 2
        elif beta >= opts.beta_max and it >= 3:
            if change < opts.tol and free_step:
                converged = True
                break
            w = opts.stall_window
            if scheduled and vol_held:
                if since_beta >= beta_hold(opts, beta):
                    converged = True
                    break
            elif since_beta >= w and len(history) > w:
                old = history[-1 - w]["compliance"]
                if abs(comp - old) <= opts.stall_rtol * abs(old):
                    converged = True
                    break
    xt, xp = L.physical(x, beta)
    return dict(x=x.reshape(p.dims), xphys=xp.reshape(p.dims), history=history, converged=converged,
                iterations=it, beta=beta)


# ---- delivery: the design handed over after the loop (simp.py:2452-2548) ----

PRUNE_BELOW = 1e-3    # simp.py:1230
PRUNE_PASSES = 3      # simp.py:1240
_CONN26 = np.ones((3, 3, 3), bool)


def keep_cores(dims, opened, analysed, blueprint, radius, pitch=1.0):
    """simp._keep_cores (simp.py:1317-1358): where any cell was analysed as
    solid, the blueprint within radius / 2.6 + half a cell of the analysed
    design comes back (max with the opening)."""
    from scipy import ndimage

    lost = analysed > 0.5
    if not ((analysed > 0.5) & (opened < 0.5)).any():
        return opened
    reach = radius / 2.6 + 0.5 * pitch
    n = int(np.ceil(reach / pitch))
    g = np.mgrid[-n:n + 1, -n:n + 1, -n:n + 1]
    ball = (g ** 2).sum(axis=0) * pitch ** 2 <= reach ** 2 * (1.0 + 1e-9)
    near = ndimage.binary_dilation(lost.reshape(dims), structure=ball).ravel()
    return np.where(near, np.maximum(opened, blueprint), opened)


def _label26(mask, dims):
    from scipy import ndimage

    return ndimage.label(mask.reshape(dims), structure=_CONN26)[0].ravel()


def anchor_cells(L: Loop):
    """_Analysis.anchor (simp.py:155): cells touching a support or a loaded
    node, and the preserved cells."""
    mesh = L.ev.mesh
    held = np.zeros(mesh.n_nodes, bool)
    held[np.unique(np.flatnonzero(L.ev.fixed) // 3)] = True
    loaded = np.zeros(mesh.n_nodes, bool)
    for f in L.ev.fs:
        loaded |= np.abs(f.reshape(-1, 3)).sum(axis=1) > 0
    return held[mesh.enode].any(axis=1) | loaded[mesh.enode].any(axis=1) | L.pres


def connected(alive, anchor, dims):
    """_Analysis._connected (simp.py:264-282)."""
    lbl = _label26(alive, dims)
    keep_l = np.unique(lbl[alive & anchor])
    return alive & np.isin(lbl, keep_l[keep_l > 0])


def held_pieces(xp, anchor, dims) -> int:
    """simp._held_pieces (simp.py:1361-1370)."""
    lbl = _label26((xp > 0.5) | anchor, dims)
    return len(np.unique(lbl[anchor]))


def prune_dead(L: Loop, xp, us):
    """simp._prune_dead (simp.py:1244-1314) for the plain problem: cells in
    clusters of >= 4 that carry under 1e-3 of the median energy density in
    every case are removed, then what that cut loose from the anchors; a
    pass that parts the design or leaves a mechanism is undone."""
    dims = L.p.dims
    anchor = anchor_cells(L)
    for _ in range(PRUNE_PASSES):
        solid = L.design & (xp > 0.5) & ~L.pres
        if not solid.any():
            break
        live = np.zeros(len(xp), bool)
        for u in us:
            w = element_energy(u, L.ev.edof, L.ev.ke) * (xp ** PENALTY)
            ref = np.median(w[xp > 0.5])
            if ref <= 0:
                live[:] = True
                break
            live |= w >= PRUNE_BELOW * ref
        dead = solid & ~live
        if dead.sum() < 0.002 * solid.sum():
            break
        lbl = _label26(dead, dims)
        n = int(lbl.max())
        if n == 0:
            break
        sizes = np.bincount(lbl[dead], minlength=n + 1)
        dead &= sizes[lbl] >= 4
        if not dead.any():
            break
        before = (xp, us)
        held = held_pieces(xp, anchor, dims)
        xp = xp.copy()
        xp[dead] = 0.0
        loose = (xp > 0.5) & ~connected((xp > 0.5) | anchor, anchor, dims) & ~L.pres
        xp[loose] = 0.0
        if held_pieces(xp, anchor, dims) > held:
            xp, us = before
            break
        us = L.solve(E_MIN + xp ** PENALTY * (1 - E_MIN))
    return xp, us


def deliver(L: Loop, x, beta, prune=True):
    """The delivered design after the loop (simp.py:2452-2548, plain
    problem): the opening of the eroded design with the analysed cores kept,
    preserved cells solid, solved, dead members pruned. Returns dict(xphys,
    compliance, case_compliance, volume_fraction)."""
    o = L.opts
    xt = (L.filt @ x) / L.hs
    xp = project(xt, beta, o.eta)
    xp[L.pres] = 1.0
    if o.erode > 0:
        xe = L.eroded(xt, beta)
        opened = project((L.filt @ xe) / L.hs, beta, o.eta - o.erode)
        xp = keep_cores(L.p.dims, opened, xe, project(xt, beta, o.eta), o.filter_radius)
        xp[L.pres] = 1.0
    xp[L.void] = 0.0
    us = L.solve(E_MIN + xp ** PENALTY * (1 - E_MIN))
    if prune:
        xp, us = prune_dead(L, xp, us)
    cases = [float(f @ u) for f, u in zip(L.ev.fs, us)]
    return dict(xphys=xp, compliance=float(np.dot(L.weights, cases)), case_compliance=cases,
                volume_fraction=float(xp[L.design].mean()))
