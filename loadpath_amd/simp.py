"""The SIMP loop on the GPU: density filter, sensitivities and the OC update
around the batched MG-PCG solve (liblpmi.so lp_simp_*).

Mirrors:
- the top3d loop ``optimize_reference(p, "top3d")`` (benchmarks/reference.py
  :214-257): SIMP p = 3, Emin 1e-9, density filter H/Hs, OC bisection with
  l2 = 1e9 and relative tolerance 1e-3, stop when change <= 0.01 or 200
  iterations, returning ``x, iterations, compliance_reported, seconds``;
- ``density_filter(domain, radius, pad)`` / ``GridFilter(domain, radius)`` (simp.py:1143-1189):
  ``h @ v``, ``h.T``, ``hs``, on the GPU.

- the production loop ``simp.optimize`` (simp.py:1570-1640, its loop
  simp.py:1840-2449) on a box: ``SimpOptions`` / ``optimize`` below
  (Heaviside projection with continuation, the analysed eroded design, the
  measured step and the volume multiplier by bracketing and Illinois false
  position, all on the device).

Several load cases are solved together and their compliances summed.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .fem import hex8_element, hex8_ke


def cone_taps(radius: float, reach: int, pitch: float = 1.0):
    """[((dx, dy, dz), w)], w = radius - |d| pitch > 0, |d_i| <= reach."""
    taps = []
    for dx in range(-reach, reach + 1):
        for dy in range(-reach, reach + 1):
            for dz in range(-reach, reach + 1):
                d = np.sqrt(dx * dx + dy * dy + dz * dz)
                w = radius - (d if pitch == 1.0 else d * pitch)
                if w > 0:
                    taps.append(((dx, dy, dz), float(w)))
    return taps


class GridFilter:
    """simp.GridFilter(domain, radius) (simp.py:1143-1179): the cone
    filter's unnormalised, symmetric matrix H as a stencil on the domain's
    box, ``h @ v`` on the GPU (lp_grid_filter). domain: an object with
    ``mesh.elem_ijk`` (the mesh's cells), ``dims`` (the box) and ``pitch``,
    as loadpath's Domain; or a dims tuple for a full box of unit cells.
    Reach floor(radius / pitch), weights radius - |d| pitch; cells outside
    the mesh count as empty. ``h.T is h``; ``h.sum()`` the row sums."""

    def __init__(self, domain, radius: float, reach: int | None = None):
        if hasattr(domain, "mesh"):
            self.dims = tuple(int(d) for d in domain.dims)
            self.pitch = float(domain.pitch)
            ijk = np.asarray(domain.mesh.elem_ijk, np.int64)
        else:
            self.dims = tuple(int(d) for d in domain)
            self.pitch = 1.0
            ijk = None
        self.radius = float(radius)
        if reach is not None and self.pitch != 1.0:
            raise ValueError("an explicit reach is for unit cells")
        self.reach = int(np.floor(self.radius / self.pitch)) if reach is None else int(reach)
        self.taps = cone_taps(self.radius, self.reach, self.pitch)
        box = int(np.prod(self.dims))
        if ijk is None or (len(ijk) == box and np.array_equal(
                np.ravel_multi_index(tuple(ijk.T), self.dims), np.arange(box))):
            self._flat = None
        else:
            self._flat = np.ravel_multi_index(tuple(ijk.T), self.dims)
        self.n = box if self._flat is None else len(self._flat)
        self.shape = (self.n, self.n)
        self.T = self

    def __matmul__(self, v):
        v = np.asarray(v, np.float64).ravel()
        if self._flat is None:
            g = np.ascontiguousarray(v.reshape(self.dims))
        else:
            g = np.zeros(self.dims)
            g.reshape(-1)[self._flat] = v
        out = np.empty(self.dims)
        if self.pitch == 1.0:
            rc = _lib.lib().lp_filter_apply(*self.dims, self.radius, self.reach, _lib.ptr(g), _lib.ptr(out), 0)
        else:
            rc = _lib.lib().lp_grid_filter(*self.dims, self.pitch, self.radius, _lib.ptr(g), _lib.ptr(out), 0)
        _lib.check(rc, "filter")
        out = out.ravel()
        return out if self._flat is None else out[self._flat]

    def tocsr(self):
        return self

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


def density_filter(domain, radius: float, pad: bool = True):
    """(h, hs) as simp.density_filter(domain, radius, pad) (simp.py:1181-
    1222): h the GridFilter stencil over the mesh's cells, hs the full
    kernel's weight (pad: cells outside the design space count as void) or
    h's row sums. domain as GridFilter takes it (a Domain-like object or a
    box's dims)."""
    h = GridFilter(domain, radius)
    hs = np.full(h.n, sum(w for _, w in h.taps)) if pad else h.sum()
    return h, hs


class GridSimp:
    """One SIMP study on the device: the grid hierarchy, the load cases and
    the design fields stay in HBM between iterations."""

    def __init__(self, problem, *, precision="fp32", smoother="chebyshev", omega=0.6, degree=2, rtol=1e-8,
                 maxiter=500, coarse_cells=2000, max_levels=0, element=None, nu=0.3, penal=3.0, emin=1e-9,
                 move=0.2, filter="top3d", warm_start=True, optimizer="oc", volfrac=None, rmin=None, active=None):
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
        self.volfrac = problem.volfrac if volfrac is None else float(volfrac)
        self.rmin = problem.rmin if rmin is None else float(rmin)
        self.filter = filter
        # active: the design cells (passive ones are held at 0), None: all
        self._active = None if active is None else np.ascontiguousarray(
            np.asarray(active, bool).reshape(problem.dims), np.uint8)
        _lib.check(_lib.lib().lp_simp_create(C.byref(h), self.grid._h, self.nrhs, _lib.ptr(self._loads),
                                             None if self._active is None else _lib.ptr(self._active),
                                             self.volfrac, self.rmin, _lib.FILTER[filter], penal, emin, move,
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
            d = dict(compliance=s.compliance, change=s.change, volume=s.volume, cg_iters=s.cg_iters,
                     cg_iters_total=s.cg_iters_total, oc_evals=s.oc_evals, seconds=s.seconds)
            if self.optimizer == "loop":
                d.update(beta=s.beta, move=s.move, step_fit=None if s.step_fit < 0 else s.step_fit, lam=s.lam,
                         vol_held=bool(s.vol_held), converged=bool(s.converged))
            out.append(d)
        return out

    def set_loop(self, opts: "SimpOptions", preserve=None, case_weights=None, mirror=None, reduce=True):
        """Switch to the production loop (lp_simp_set_loop) and restart it
        from x = volume fraction (preserved cells 1). mirror: mirror planes
        through the box centre as axes ("y", "xz"; domain.mirror), tied with
        opts.extrusion's columns (lp_simp_set_ties); reduce: the multigrid
        path's live-cell reduction of simp._Analysis (lp_simp_set_reduction,
        the reference's behaviour)."""
        sched = None if opts.beta_schedule is None else np.ascontiguousarray(opts.beta_schedule, np.int32)
        o = _lib.LoopOpts(beta_start=opts.beta_start, beta_max=opts.beta_max, beta_every=opts.beta_every,
                          eta=opts.eta, erode=opts.erode, move=opts.move, adaptive_move=int(opts.adaptive_move),
                          tol=opts.tol, stall_rtol=opts.stall_rtol, stall_window=opts.stall_window,
                          emin=E_MIN, penal=PENALTY,
                          schedule=None if sched is None else sched.ctypes.data,
                          schedule_len=0 if sched is None else len(sched))
        self._pres = None if preserve is None else np.ascontiguousarray(np.asarray(preserve, bool).reshape(
            self.problem.dims), np.uint8)
        self._cw = None if case_weights is None else np.ascontiguousarray(case_weights, np.float64)
        if self._cw is not None and len(self._cw) != self.nrhs:
            raise ValueError(f"case_weights: {self.nrhs} load cases, {len(self._cw)} weights")
        _lib.check(_lib.lib().lp_simp_set_loop(self._h, C.byref(o), None if self._pres is None else _lib.ptr(self._pres),
                                               None if self._cw is None else _lib.ptr(self._cw)), "simp_set_loop")
        axes = np.ascontiguousarray([_AXIS[a] if isinstance(a, str) else int(a) for a in (mirror or "")], np.int32)
        ext = -1 if opts.extrusion is None else (_AXIS[opts.extrusion] if isinstance(opts.extrusion, str)
                                                 else int(opts.extrusion))
        _lib.check(_lib.lib().lp_simp_set_ties(self._h, _lib.ptr(axes) if len(axes) else None, len(axes), ext),
                   "simp_set_ties")
        if reduce:
            self._support, self._anchor = self._anchor_cells(self._pres)
            _lib.check(_lib.lib().lp_simp_set_reduction(self._h, _lib.ptr(self._support), _lib.ptr(self._anchor)),
                       "simp_set_reduction")
        else:
            _lib.check(_lib.lib().lp_simp_set_reduction(self._h, None, None), "simp_set_reduction")
        self.optimizer = "loop"
        self.loop_opts = opts
        self.iterations = 0

    def _anchor_cells(self, pres):
        """(support_touch, anchor) per cell, uint8: the cell touches a node
        with a held dof; it does, or touches a loaded node, or is preserved
        (simp._Analysis, simp.py:154-155)."""
        held = self.problem.fixed_grid().any(axis=0)
        loaded = (self._loads != 0).any(axis=(0, 1))

        def touch(nodes):
            c = np.zeros(self.problem.dims, bool)
            for a in (0, 1):
                for b in (0, 1):
                    for d in (0, 1):
                        c |= nodes[a:a + c.shape[0], b:b + c.shape[1], d:d + c.shape[2]]
            return c

        sup = touch(held)
        anc = sup | touch(loaded)
        if pres is not None:
            anc |= pres.astype(bool)
        return np.ascontiguousarray(sup, np.uint8), np.ascontiguousarray(anc, np.uint8)

    def loop_info(self):
        """The reduction's counters: iterations solved reduced, whole-grid
        repeats after a failed reduced solve, iterations sent to the whole
        grid by a load on a dead dof."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _lib.check(_lib.lib().lp_simp_loop_info(self._h, C.byref(a), C.byref(b), C.byref(c)), "simp_loop_info")
        return dict(reduced=a.value, full_retries=b.value, full_for_loads=c.value)

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


E_MIN = 1e-6       # simp.py:42 (the production loop's void modulus)
_AXIS = {"x": 0, "y": 1, "z": 2}
PENALTY = 3.0      # simp.py:55


@dataclass
class SimpOptions:
    """The fields of simp.SimpOptions (simp.py:899-1000) that the loop on a
    box reads, with the reference's defaults. solve_rtol is each
    iteration's PCG tolerance (simp.py:929)."""
    volume_fraction: float
    filter_radius: float
    max_iterations: int = 120
    beta_max: float = 32.0
    beta_every: int = 10
    beta_schedule: tuple | None = None
    move: float = 0.2
    adaptive_move: bool = True
    tol: float = 0.01
    solve_rtol: float = 1e-4
    stall_rtol: float = 2e-3
    stall_window: int = 6
    eta: float = 0.5
    erode: float = 0.15
    beta_start: float = 1.0
    prune: bool = True
    extrusion: int | str | None = None   # the axis of a 2.5D part (SimpOptions.extrusion)


@dataclass
class SimpResult:
    """simp.SimpResult's fields that the loop fills (simp.py:1093-1121):
    x the design variables, xphys the delivered design (deliver(); the
    blueprint P(H x / hs) at the final sharpness when optimize(...,
    delivered=False)), compliance / case_compliance / volume_fraction of it,
    history one record per iteration (the reference's keys: iteration,
    compliance, relative, volume, beta, change, seconds, cgIterations,
    volumeSolves, move, stepFit)."""
    x: np.ndarray
    xphys: np.ndarray
    compliance: float
    volume_fraction: float
    iterations: int
    converged: bool
    seconds: float
    history: list = field(default_factory=list)
    case_compliance: list = field(default_factory=list)


def optimize(problem, opts: SimpOptions, *, preserve=None, void=None, case_weights=None, precision="mixed", pad=True,
             on_iter=None, delivered=True, mirror=None, reduce=True, **kw) -> SimpResult:
    """simp.optimize (simp.py:1570-1640) for a box problem with the plain
    compliance objective, every iteration on the device (loop
    simp.py:1840-2449). preserve: cells held solid (domain.preserve);
    void: cells held empty (opts.void), still counted in the volume
    target; case_weights: the load cases' weights (domain.case_weights).
    delivered: hand over the reference's delivered design (the opening of
    the eroded design with its cores kept, pruned when opts.prune:
    deliver()); False: the blueprint at the loop's end. mirror: mirror
    planes through the box centre as axes (domain.mirror), tied with
    opts.extrusion's columns; reduce: the multigrid path's live-cell
    reduction (simp.py:795-826). kw go to GridSimp (precision,
    coarse_cells, element, ...)."""
    kw.setdefault("rtol", opts.solve_rtol)
    s = GridSimp(problem, precision=precision, filter="padded" if pad else "top3d", volfrac=opts.volume_fraction,
                 rmin=opts.filter_radius, active=None if void is None else ~np.asarray(void, bool), **kw)
    s.set_loop(opts, preserve=preserve, case_weights=case_weights, mirror=mirror, reduce=reduce)
    t0 = time.time()
    history = []
    converged = False
    c0 = None
    for it in range(1, opts.max_iterations + 1):
        st = s.iterate(1)[0]
        c0 = st["compliance"] if c0 is None else c0
        rec = {"iteration": it, "compliance": st["compliance"], "relative": st["compliance"] / c0 if c0 else 1.0,
               "volume": st["volume"], "beta": st["beta"], "change": st["change"], "seconds": time.time() - t0,
               "cgIterations": st["cg_iters"], "volumeSolves": st["oc_evals"], "lambda": st["lam"],
               "volHeld": st["vol_held"]}
        if opts.adaptive_move:
            rec["move"] = st["move"]
            if st["step_fit"] is not None:
                rec["stepFit"] = st["step_fit"]
        history.append(rec)
        if on_iter is not None:
            on_iter(rec)
        if st["converged"]:
            converged = True
            break
    x = s.x
    beta = _final_beta(history, opts)
    if delivered:
        d = deliver(s, x, beta, opts, preserve=preserve, void=void, pad=pad, prune=opts.prune)
        return SimpResult(x=x, xphys=d["xphys"], compliance=d["compliance"], volume_fraction=d["volume_fraction"],
                          iterations=len(history), converged=converged, seconds=time.time() - t0, history=history,
                          case_compliance=d["case_compliance"])
    xphys = blueprint(x, opts, pad=pad, beta=beta, preserve=preserve, void=void)
    return SimpResult(x=x, xphys=xphys, compliance=history[-1]["compliance"] if history else 0.0,
                      volume_fraction=float(xphys[~_pres_mask(preserve, x.shape)].mean()),
                      iterations=len(history), converged=converged, seconds=time.time() - t0, history=history)


def _pres_mask(preserve, shape):
    return np.zeros(shape, bool) if preserve is None else np.asarray(preserve, bool).reshape(shape)


def _final_beta(history, opts):
    """The sharpness after the last iteration's continuation test
    (simp.py:2413-2421): the records hold each iteration's own beta."""
    if not history:
        return opts.beta_start
    beta = history[-1]["beta"]
    since = 0
    for h in reversed(history):
        if h["beta"] != beta:
            break
        since += 1
    last = history[-1]
    free = not opts.adaptive_move or last["change"] < 0.999 * last.get("move", opts.move)
    if beta < opts.beta_max and ((last["change"] < opts.tol * 2 and free) or since >= _beta_hold(opts, beta)):
        beta = min(beta * 2, opts.beta_max)
    return beta


def _beta_hold(opts, beta):
    s = opts.beta_schedule
    k = int(round(np.log2(max(beta, 1.0))))
    return s[k] if s is not None and k < len(s) else opts.beta_every


def blueprint(x, opts: SimpOptions, *, pad=True, beta=None, preserve=None, void=None):
    """The blueprint xp = P(H x / hs, beta, eta), preserved cells 1, void
    cells 0 (simp.physical, simp.py:1775-1781), on the device."""
    x = np.asarray(x, float)
    f = GridFilter(x.shape, opts.filter_radius)
    hs = np.full(x.size, sum(w for _, w in f.taps)) if pad else f.sum()
    xt = (f @ x.ravel()) / hs
    xp = project(xt, opts.beta_max if beta is None else beta, opts.eta).reshape(x.shape)
    xp[_pres_mask(preserve, x.shape)] = 1.0
    if void is not None:
        xp[np.asarray(void, bool).reshape(x.shape)] = 0.0
    return xp


def label26(mask):
    """scipy.ndimage.label(mask, structure=np.ones((3, 3, 3))) on the device
    (lp_label26): (labels int32 of mask's shape, count), components
    numbered by their first cell in C order."""
    _lib.require_gpu()
    m = np.ascontiguousarray(np.asarray(mask, bool), np.uint8)
    if m.ndim != 3:
        raise ValueError("label26 takes a 3-D cell mask")
    labels = np.empty(m.shape, np.int32)
    n = C.c_int()
    _lib.check(_lib.lib().lp_label26(*m.shape, _lib.ptr(m), _lib.ptr(labels), C.byref(n)), "label26")
    return labels, n.value


def connected(alive, anchor):
    """The cells of `alive` in a 26-connected piece that holds an `anchor`
    cell (simp._Analysis._connected, simp.py:264-282: the supports' cells)."""
    alive = np.asarray(alive, bool)
    lbl, _ = label26(alive)
    keep = np.unique(lbl[alive & np.asarray(anchor, bool)])
    return alive & np.isin(lbl, keep[keep > 0])


PRUNE_BELOW = 1e-3   # simp.py:1230
PRUNE_PASSES = 3     # simp.py:1240


def _corner_touch(node_mask, dims):
    """Cells with any of their 8 corners on a marked node."""
    nx, ny, nz = dims
    out = np.zeros(dims, bool)
    for b in range(8):
        bx, by, bz = b & 1, (b >> 1) & 1, (b >> 2) & 1
        out |= node_mask[bx:bx + nx, by:by + ny, bz:bz + nz]
    return out


def anchor_cells(problem, preserve=None):
    """_Analysis.anchor (simp.py:155) on a box: cells touching a support or a
    loaded node, and the preserved cells."""
    dims = problem.dims
    nodes = problem.fixed_grid().any(axis=0) | (np.abs(problem.load_grids()).sum(axis=(0, 1)) > 0)
    return (_corner_touch(nodes, dims) | _pres_mask(preserve, dims)).ravel()


def keep_cores(dims, opened, analysed, blueprint, radius, pitch=1.0):
    """simp._keep_cores (simp.py:1317-1358) on the device: where any cell
    was analysed as solid, the blueprint within radius / 2.6 + half a cell
    of the analysed design comes back (max with the opening). The ball's
    dilation is the cone filter with every weight positive."""
    if not ((analysed > 0.5) & (opened < 0.5)).any():
        return opened
    reach = radius / 2.6 + 0.5 * pitch
    n = int(np.ceil(reach / pitch))
    # centre distance <= reach (the reference's ball, 1e-9 relative slack)
    ball = GridFilter(dims, reach / pitch * np.sqrt(1.0 + 1e-9), reach=n)
    near = (ball @ (analysed > 0.5).astype(float)) > 0
    return np.where(near, np.maximum(opened, blueprint), opened)


def _held_pieces(xp, anchor, dims):
    lbl, _ = label26(((xp > 0.5) | anchor).reshape(dims))
    return len(np.unique(lbl.ravel()[anchor]))


def deliver(s: "GridSimp", x, beta, opts: SimpOptions, *, preserve=None, void=None, pad=True, prune=True):
    """The design handed over after the loop (simp.py:2452-2548, plain
    problem), on the device: the opening of the eroded design with the
    analysed cores kept (simp._keep_cores), preserved cells solid, solved at
    rtol 1e-6, dead members pruned (simp._prune_dead). s: the study's
    GridSimp (its grid solves the delivery). Returns dict(xphys,
    compliance, case_compliance, volume_fraction)."""
    dims = s.problem.dims
    x = np.asarray(x, float).ravel()
    pres = _pres_mask(preserve, dims).ravel()
    vmask = np.zeros(x.size, bool) if void is None else np.asarray(void, bool).ravel()
    design = ~pres
    f = GridFilter(dims, opts.filter_radius)
    hs = np.full(x.size, sum(w for _, w in f.taps)) if pad else f.sum()
    xt = (f @ x) / hs
    bp = project(xt, beta, opts.eta)
    xp = bp.copy()
    xp[pres] = 1.0
    if opts.erode > 0:
        xe = project(xt, beta, opts.eta + opts.erode)
        xe[pres] = 1.0
        xe[vmask] = 0.0
        opened = project((f @ xe) / hs, beta, opts.eta - opts.erode)
        xp = keep_cores(dims, opened, xe, bp, opts.filter_radius)
        xp[pres] = 1.0
    xp[vmask] = 0.0
    g = s.grid
    g.rtol = 1e-6
    _, loads = s.problem.dof_vectors(g.mesh)
    weights = np.ones(len(loads)) if getattr(s, "_cw", None) is None else s._cw

    def solve(xv):
        g.set_moduli(E_MIN + xv ** PENALTY * (1 - E_MIN))
        return g.solve_many(list(loads))

    us = solve(xp)
    if prune:
        anchor = anchor_cells(s.problem, preserve)
        for _ in range(PRUNE_PASSES):
            solid = design & (xp > 0.5) & ~pres
            if not solid.any():
                break
            live = np.zeros(x.size, bool)
            for u in us:
                w = g.strain_energy(u) * (xp ** PENALTY)
                ref = np.median(w[xp > 0.5])
                if ref <= 0:
                    live[:] = True
                    break
                live |= w >= PRUNE_BELOW * ref
            dead = solid & ~live
            if dead.sum() < 0.002 * solid.sum():
                break
            lbl, nlab = label26(dead.reshape(dims))
            lbl = lbl.ravel()
            if nlab == 0:
                break
            sizes = np.bincount(lbl[dead], minlength=nlab + 1)
            dead &= sizes[lbl] >= 4
            if not dead.any():
                break
            before = (xp, us)
            held = _held_pieces(xp, anchor, dims)
            xp = xp.copy()
            xp[dead] = 0.0
            loose = (xp > 0.5) & ~connected(((xp > 0.5) | anchor).reshape(dims), anchor.reshape(dims)).ravel() & ~pres
            xp[loose] = 0.0
            if _held_pieces(xp, anchor, dims) > held:
                xp, us = before
                break
            try:
                us = solve(xp)
            except RuntimeError:
                xp, us = before
                break
    cases = [float(fv @ u) for fv, u in zip(loads, us)]
    return dict(xphys=xp.reshape(dims), compliance=float(np.dot(weights, cases)), case_compliance=cases,
                volume_fraction=float(xp[design].mean()))


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
