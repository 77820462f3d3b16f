"""Drop-in ``GridElasticity`` on MI355X: matrix-free hex8 multigrid-preconditioned
CG for K(E) u = f, with the hierarchy, V-cycle and CG on the GPU (liblpmi.so).

Mirrors loadpath/mg.py's public surface (mg.py:411-988):
``GridElasticity(mesh, nu, fixed, springs=None, rtol=1e-6, maxiter=500,
coarse_cells=2000, degree=2, dtype=np.float32, lu_max_dofs=60_000,
device=None, lowrank=None)`` with ``set_moduli(e)``, ``solve_many(fs, x0s)``,
``strain_energy(u)``, ``live_dofs()``, ``apply(u_grid)``,
``precondition(r_grid)``, ``to_grid`` / ``from_grid`` / ``cell_grid`` and the
attributes ``last_iters``, ``strict``, ``levels``, ``rtol``, ``maxiter``; and
the module functions ``prolong`` / ``restrict`` (mg.py:84-132).

Differences, by design: all right-hand sides of ``solve_many`` are solved in
one batched PCG (the reference loops over them, mg.py:932-949); the
coarsest level is factored on the GPU (FP64 Cholesky -> inverse) instead of
SuperLU on the host, so ``lu_max_dofs`` only bounds its size; ``dtype``
float64 selects the all-FP64 V-cycle. Extra keywords: ``smoother``
("chebyshev" | "jacobi"), ``omega``, ``max_rhs``, ``max_levels``.
``springs`` / ``lowrank`` (extra sparse terms) are outside this package.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .fem import hex8_element, hex8_ke

__all__ = ["GridElasticity", "prolong", "restrict"]


def _grid_op(fn, src: np.ndarray, out_shape, cx, cy, cz):
    src = np.ascontiguousarray(src, np.float64)
    lead = src.shape[:-3]
    flat = src.reshape((-1,) + src.shape[-3:])
    pad = (-flat.shape[0]) % 3
    if pad:
        flat = np.concatenate([flat, np.zeros((pad,) + flat.shape[1:])])
    outs = []
    for k in range(0, flat.shape[0], 3):
        o = np.empty((3,) + out_shape)
        _lib.check(fn(_lib.ptr(np.ascontiguousarray(flat[k:k + 3])), _lib.ptr(o), cx, cy, cz, 0), fn.__name__)
        outs.append(o)
    res = np.concatenate(outs)[: flat.shape[0] - pad]
    return res.reshape(lead + out_shape)


def prolong(c: np.ndarray) -> np.ndarray:
    """Trilinear prolongation (..., n0, n1, n2) -> (..., 2n0-1, 2n1-1, 2n2-1)
    on the GPU (mg.py:84-108)."""
    cx, cy, cz = c.shape[-3:]
    return _grid_op(_lib.lib().lp_prolong, c, (2 * cx - 1, 2 * cy - 1, 2 * cz - 1), cx, cy, cz)


def restrict(f: np.ndarray) -> np.ndarray:
    """Transpose of prolong (mg.py:111-132)."""
    fx, fy, fz = f.shape[-3:]
    cx, cy, cz = (fx + 1) // 2, (fy + 1) // 2, (fz + 1) // 2
    if (2 * cx - 1, 2 * cy - 1, 2 * cz - 1) != (fx, fy, fz):
        raise ValueError("restrict needs odd node counts 2n-1 per axis")
    return _grid_op(_lib.lib().lp_restrict, f, (cx, cy, cz), cx, cy, cz)


class _LevelInfo:
    def __init__(self, dims, lmax):
        self.dims = tuple(int(d) for d in dims)
        self.lmax = float(lmax)

    @property
    def node_shape(self):
        nx, ny, nz = self.dims
        return (3, nx + 1, ny + 1, nz + 1)


class GridElasticity:
    """See the module docstring; mg.GridElasticity (mg.py:411)."""

    STALL = 100

    def __init__(self, mesh, nu, fixed, springs=None, rtol=1e-6, maxiter=500, coarse_cells=2000, degree=2,
                 dtype=np.float32, lu_max_dofs=60_000, device=None, lowrank=None, *, smoother="chebyshev",
                 omega=0.6, max_rhs=8, max_levels=0, element=None, ke=None, precision=None):
        if springs is not None or lowrank is not None:
            raise NotImplementedError("springs / lowrank extra terms are not part of the MI355X build")
        _lib.require_gpu()
        self.mesh = mesh
        self.nu = nu
        self._rtol = float(rtol)
        self._maxiter = int(maxiter)
        self.coarse_cells = int(coarse_cells)
        self.degree = int(degree)
        self.dtype = np.dtype(dtype)
        self.lu_max_dofs = int(lu_max_dofs)
        self.smoother = smoother
        self.omega = float(omega)
        self.max_rhs = int(max_rhs)
        if precision is None:
            precision = "fp64" if self.dtype == np.float64 else "fp32"
        self.precision = precision
        cells = mesh.elem_ijk.max(axis=0) + 1
        cells = cells + cells % 2
        self.cells = tuple(int(c) for c in cells)
        self.node_shape = tuple(int(c) + 1 for c in cells)
        self.nn = int(np.prod(self.node_shape))
        self.node_flat = np.ravel_multi_index(tuple(mesh.node_ijk.T), self.node_shape)
        self.elem_ijk = mesh.elem_ijk
        self.element = element or hex8_element()
        self.ke = (hex8_ke(nu, self.element) if ke is None else np.asarray(ke, float)) * mesh.pitch
        fixed = np.asarray(fixed, bool)
        fg = np.zeros((3, self.nn), bool)
        for i in range(3):
            fg[i, self.node_flat] = fixed[i::3]
        self.fixed_grid = fg.reshape((3,) + self.node_shape)
        self.levels = []
        self._strict = False
        self.last_iters = 0
        self._lmax_started = set()
        self._mask = None
        opt = _lib.Options(rtol=self._rtol, maxiter=self._maxiter, coarse_cells=self.coarse_cells, degree=self.degree,
                           smoother=_lib.SMOOTH[smoother], omega=self.omega, precision=_lib.PREC[precision],
                           max_rhs=self.max_rhs, lmax_iters=12, max_levels=int(max_levels))
        self._opt = opt
        h = C.c_void_p()
        fgu = np.ascontiguousarray(self.fixed_grid, np.uint8)
        _lib.check(_lib.lib().lp_grid_create(C.byref(h), *self.cells, _lib.ptr(fgu),
                                             _lib.ptr(np.ascontiguousarray(self.ke)), C.byref(opt)), "lp_grid_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().lp_grid_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---- layout conversion (mg.py:495-511) --------------------------------
    def to_grid(self, v, dtype=np.float64):
        out = np.zeros((3, self.nn), dtype)
        for i in range(3):
            out[i, self.node_flat] = v[i::3]
        return out.reshape((3,) + self.node_shape)

    def from_grid(self, g):
        flat = np.asarray(g).reshape(3, -1)
        out = np.empty(3 * len(self.node_flat))
        for i in range(3):
            out[i::3] = flat[i, self.node_flat]
        return out

    def cell_grid(self, e, dtype=np.float64):
        g = np.zeros(self.cells, dtype)
        g[tuple(self.elem_ijk.T)] = e
        return g

    # ---- hierarchy --------------------------------------------------------
    def set_moduli(self, e):
        """mg.py:515-591, on the device."""
        e = np.asarray(e, float)
        self._e = e
        eg = np.ascontiguousarray(self.cell_grid(e))
        lib = _lib.lib()
        cheb = self.smoother == "chebyshev"
        _lib.check(lib.lp_set_moduli(self._h, _lib.ptr(eg), 2 if cheb else 0), "set_moduli")
        self._refresh(lmax=cheb)

    @property
    def cell_order_is_grid(self) -> bool:
        """The mesh's cells are the padded cell grid in C order (a full box
        with even dims): element vectors and cell-grid fields coincide, so
        device tensors pass through unchanged."""
        if not hasattr(self, "_grid_order"):
            n = int(np.prod(self.cells))
            ijk = self.elem_ijk
            self._grid_order = len(ijk) == n and bool(np.array_equal(
                np.ravel_multi_index(tuple(ijk.T), self.cells), np.arange(n)))
        return self._grid_order

    def set_moduli_device(self, e):
        """set_moduli from a CUDA float64 tensor of the padded cell grid's
        moduli (C order); the hierarchy is rebuilt from device memory."""
        import torch

        e = e.detach().to(torch.float64).contiguous()
        if not e.is_cuda or e.numel() != int(np.prod(self.cells)):
            raise ValueError("set_moduli_device: a CUDA tensor over the padded cell grid")
        self._e = None
        torch.cuda.synchronize()
        cheb = self.smoother == "chebyshev"
        _lib.check(_lib.lib().lp_set_moduli(self._h, C.c_void_p(e.data_ptr()), 1 | (2 if cheb else 0)),
                   "set_moduli")
        self._refresh(lmax=cheb)

    def solve_device(self, f_grid, x0=None):
        """Batched solve of CUDA node-grid tensors (nrhs, 3, nodes...):
        masked on the device (lp_mask_rhs), solved in place of x0 (or from
        zero). Returns (x, iterations, converged) with x a CUDA tensor."""
        import torch

        f = f_grid.detach().to(torch.float64).contiguous()
        nrhs = f.shape[0]
        b = torch.empty_like(f)
        x = torch.zeros_like(f) if x0 is None else x0.detach().to(torch.float64).contiguous().clone()
        torch.cuda.synchronize()
        lib = _lib.lib()
        _lib.check(lib.lp_mask_rhs(self._h, nrhs, C.c_void_p(f.data_ptr()), C.c_void_p(b.data_ptr()), 1), "mask_rhs")
        if x0 is not None:
            _lib.check(lib.lp_mask_rhs(self._h, nrhs, C.c_void_p(x.data_ptr()), C.c_void_p(x.data_ptr()), 1),
                       "mask_rhs")
        its = np.zeros(nrhs, np.int32)
        ok = np.zeros(nrhs, np.int32)
        rc = lib.lp_solve(self._h, nrhs, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), int(x0 is not None), 1,
                          _lib.ptr(its), _lib.ptr(ok))
        if rc not in (_lib.LP_OK, _lib.LP_NOT_CONVERGED):
            _lib.check(rc, "solve")
        # the host path's x * mask, kept on the device
        _lib.check(lib.lp_mask_rhs(self._h, nrhs, C.c_void_p(x.data_ptr()), C.c_void_p(x.data_ptr()), 1), "mask_rhs")
        self._refresh(lmax=False)
        self.last_iters = int(its.max())
        return x, its, ok, b

    def _refresh(self, lmax: bool):
        lib = _lib.lib()
        nl = C.c_int()
        dims = np.zeros(3 * 16, np.int32)
        strict = C.c_int()
        lm = np.zeros(16)
        _lib.check(lib.lp_grid_info(self._h, C.byref(nl), _lib.ptr(dims), C.byref(strict), _lib.ptr(lm)), "info")
        n = nl.value
        if lmax:
            # the reference's start vectors: default_rng(i).random - 0.5 on the
            # level's node grid, then the last vector (mg.py:771-794)
            for i in range(n):
                shape = (3,) + tuple(int(d) + 1 for d in dims[3 * i:3 * i + 3])
                if i not in self._lmax_started:
                    start = np.ascontiguousarray(np.random.default_rng(i).random(shape) - 0.5)
                    _lib.check(lib.lp_estimate_lmax(self._h, i, _lib.ptr(start)), "lmax")
                    self._lmax_started.add(i)
                else:
                    _lib.check(lib.lp_estimate_lmax(self._h, i, None), "lmax")
            _lib.check(lib.lp_grid_info(self._h, C.byref(nl), _lib.ptr(dims), C.byref(strict), _lib.ptr(lm)), "info")
        self._strict = bool(strict.value)
        self.levels = [_LevelInfo(dims[3 * i:3 * i + 3], lm[i]) for i in range(n)]
        self._mask = None

    def level_mask(self, level=0) -> np.ndarray:
        """(3, nodes...) float mask of a level (1 = free and carrying stiffness)."""
        lv = self.levels[level]
        bits = np.empty(int(np.prod(lv.node_shape[1:])), np.uint8)
        _lib.check(_lib.lib().lp_level_data(self._h, level, 0, _lib.ptr(bits)), "level_data")
        m = np.stack([(bits >> c) & 1 for c in range(3)]).astype(np.float64)
        return m.reshape(lv.node_shape)

    def level_dinv(self, level=0) -> np.ndarray:
        lv = self.levels[level]
        out = np.empty(lv.node_shape)
        _lib.check(_lib.lib().lp_level_data(self._h, level, 1, _lib.ptr(out)), "level_data")
        return out

    def cell_dofs(self, level=0) -> np.ndarray:
        """(cells, 24) int64 grid dof index per cell corner dof, as the
        device kernels compute it."""
        lv = self.levels[level]
        out = np.empty((int(np.prod(lv.dims)), 24), np.int64)
        _lib.check(_lib.lib().lp_level_data(self._h, level, 2, _lib.ptr(out)), "level_data")
        return out

    def _mask0(self):
        if self._mask is None:
            self._mask = self.level_mask(0)
        return self._mask

    # ---- operators ------------------------------------------------------------
    def apply(self, u, level=0, dtype=np.float64):
        """The level operator m*K(m*u) + (1-m)*u (mg.py:317-346); level 0 FP64
        by default; dtype float32 runs the V-cycle copy."""
        dt = np.dtype(dtype)
        lv = self.levels[level]
        u = np.ascontiguousarray(u, dt)
        nrhs = u.size // int(np.prod(lv.node_shape))
        out = np.empty_like(u)
        _lib.check(_lib.lib().lp_apply(self._h, level, dt.itemsize, _lib.ptr(u), _lib.ptr(out), nrhs, 0), "apply")
        return out

    def precondition(self, r):
        """One V-cycle, FP64 in and out (mg.py:862-864)."""
        r = np.ascontiguousarray(r, np.float64)
        nrhs = r.size // (3 * self.nn)
        z = np.empty_like(r)
        _lib.check(_lib.lib().lp_precondition(self._h, _lib.ptr(r), _lib.ptr(z), nrhs, 0), "precondition")
        return z

    # ---- solve (mg.py:932-965) -----------------------------------------------
    def solve_grids(self, b, x0=None):
        """Batched PCG on node grids (nrhs, 3, nodes...), masked. Returns
        (x, iterations per rhs, converged per rhs)."""
        b = np.ascontiguousarray(b, np.float64)
        nrhs = b.shape[0]
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        its = np.zeros(nrhs, np.int32)
        ok = np.zeros(nrhs, np.int32)
        lib = _lib.lib()
        opt_max = self._opt.maxiter
        rc = lib.lp_solve(self._h, nrhs, _lib.ptr(b), _lib.ptr(x), int(x0 is not None), 0, _lib.ptr(its), _lib.ptr(ok))
        del opt_max
        if rc not in (_lib.LP_OK, _lib.LP_NOT_CONVERGED):
            _lib.check(rc, "solve")
        self._refresh(lmax=False)
        return x, its, ok

    # strict: the hierarchy the next set_moduli builds (mg.py:561-586); a
    # solve that stalls in the plain one sets it, and it stays set
    @property
    def strict(self):
        return self._strict

    @strict.setter
    def strict(self, v):
        self._strict = bool(v)
        if getattr(self, "_h", None) is not None:
            _lib.lib().lp_tune(self._h, 10, int(self._strict))

    # rtol / maxiter: the reference's callers set them before a solve
    # (simp.py:734-746); they reach the device handle at once
    @property
    def rtol(self):
        return self._rtol

    @rtol.setter
    def rtol(self, v):
        self._rtol = float(v)
        if getattr(self, "_h", None) is not None:
            _lib.check(_lib.lib().lp_set_solve_tol(self._h, self._rtol, self._maxiter), "set_solve_tol")

    @property
    def maxiter(self):
        return self._maxiter

    @maxiter.setter
    def maxiter(self, v):
        self._maxiter = int(v)
        if getattr(self, "_h", None) is not None:
            _lib.check(_lib.lib().lp_set_solve_tol(self._h, self._rtol, self._maxiter), "set_solve_tol")

    def solve_many(self, fs, x0s=None):
        fs = list(fs)
        if not fs:
            return []
        m = self._mask0()
        out = []
        iters = 0
        for k in range(0, len(fs), self.max_rhs):
            chunk = fs[k:k + self.max_rhs]
            b = np.stack([self.to_grid(f) * m for f in chunk])
            x0 = None
            if x0s is not None and any(v is not None for v in x0s[k:k + self.max_rhs]):
                x0 = np.stack([self.to_grid(v) * m if v is not None else np.zeros_like(m)
                               for v in x0s[k:k + self.max_rhs]])
            x, its, ok = self.solve_grids(b, x0)
            m = self._mask0()
            for j in range(len(chunk)):
                iters = max(iters, int(its[j]))
                if not ok[j]:
                    raise RuntimeError(f"linear solve did not converge ({int(its[j])} iterations); "
                                       "the structure may be a mechanism (check supports)")
                out.append(self.from_grid(x[j] * m))
        self.last_iters = iters
        return out

    def live_dofs(self):
        bits = np.empty(self.nn, np.uint8)
        held = np.empty(self.nn, np.uint8)
        _lib.check(_lib.lib().lp_level_data(self._h, 0, 0, _lib.ptr(bits)), "level_data")
        _lib.check(_lib.lib().lp_level_data(self._h, 0, 3, _lib.ptr(held)), "level_data")
        alive = bits | held
        out = np.empty(3 * len(self.node_flat), bool)
        for i in range(3):
            out[i::3] = ((alive >> i) & 1)[self.node_flat].astype(bool)
        return out

    def strain_energy(self, u):
        """Per element u_e^T Ke u_e at unit modulus, mesh order (mg.py:975-988)."""
        g = np.ascontiguousarray(self.to_grid(u)[None])
        se = np.empty((1,) + self.cells)
        _lib.check(_lib.lib().lp_strain_energy(self._h, 1, _lib.ptr(g), _lib.ptr(se), 0), "strain_energy")
        return se[0][tuple(self.elem_ijk.T)]

    def element_product(self, a, b):
        """Per element a_e^T Ke b_e, mesh order (the adjoint term of
        constrained.py:659; strain_energy when a is b)."""
        ga = np.ascontiguousarray(self.to_grid(a)[None])
        gb = ga if b is a else np.ascontiguousarray(self.to_grid(b)[None])
        out = np.empty((1,) + self.cells)
        _lib.check(_lib.lib().lp_element_product(self._h, 1, _lib.ptr(ga), _lib.ptr(gb), _lib.ptr(out), 0),
                   "element_product")
        return out[0][tuple(self.elem_ijk.T)]

    _TUNED = {}

    def autotune(self, nrhs=None, reps=5) -> int:
        """Pick the fine-level x-chunk of the V-cycle operator by timing
        (lp_autotune); cached per (cells, precision, batch) for the process."""
        nrhs = int(nrhs or self.max_rhs)
        key = (self.cells, self.precision, nrhs)
        lib = _lib.lib()
        if key in GridElasticity._TUNED:
            fine, weights = GridElasticity._TUNED[key]
            lib.lp_tune(self._h, 1, fine)
            lib.lp_tune(self._h, 8, weights)
            return fine
        xc = C.c_int()
        ms = C.c_double()
        _lib.check(lib.lp_autotune(self._h, nrhs, int(reps), C.byref(xc), C.byref(ms)), "autotune")
        GridElasticity._TUNED[key] = (xc.value, int(lib.lp_tune(self._h, 8, 0)))
        return xc.value

    def time_apply(self, level=0, dtype=np.float64, nrhs=1, reps=20) -> float:
        """Mean device time (ms) of one operator apply (HIP events)."""
        ms = C.c_double()
        _lib.check(_lib.lib().lp_time_apply(self._h, level, np.dtype(dtype).itemsize, nrhs, reps, C.byref(ms)),
                   "time_apply")
        return ms.value
