"""The backend plugin slot of loadpath's GridElasticity, on MI355X.

loadpath.mg.GridElasticity calls ``resolve_device(device)`` and, when that
returns a device, constructs ``self._dev = Backend(self, dev)``
(mg.py:476-484), calls ``self._dev.load()`` after each hierarchy build
(mg.py:587-589) and ``self._dev.solve(b_grid, x0_grid) -> (x, iterations,
converged)`` in place of its own PCG (mg.py:953-964). ``HipBackend`` has
those three members: ``load()`` takes the fine moduli, supports and element
matrix from the host grid object and rebuilds the whole hierarchy on the
GPU (liblpmi.so); ``solve`` runs the batched MG-PCG there. The reference's
TorchMG (gpu.py:280-521) uploads the host hierarchy instead and solves the
coarsest level on the host.

``resolve_device`` mirrors gpu.resolve_device (gpu.py:56-79):
LOADPATH_DEVICE = auto | cuda | mi355x | hip selects the GPU (torch-ROCm
names the HIP device "cuda"), cpu / numpy / off returns None.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

__all__ = ["resolve_device", "HipBackend"]

_GPU_NAMES = ("auto", "cuda", "mi355x", "hip", "gpu")


def resolve_device(name: str | None = None):
    """'mi355x' when the HIP path is selected and a GPU is present, else None."""
    name = (name if name is not None else os.environ.get("LOADPATH_DEVICE", "auto")).strip().lower()
    if name in ("", "cpu", "none", "off", "numpy"):
        return None
    if name not in _GPU_NAMES:
        raise RuntimeError(f"LOADPATH_DEVICE={name}: the MI355X backend takes one of {_GPU_NAMES} or cpu")
    import torch

    if not torch.cuda.is_available():
        if name == "auto":
            return None
        raise RuntimeError(f"LOADPATH_DEVICE={name} but no GPU is visible")
    _lib.load_library()
    return "mi355x"


class HipBackend:
    """Device copy of a GridElasticity's operator and its MG-PCG solve.

    ``grid`` is the host GridElasticity (loadpath's or loadpath_amd's): the
    backend reads ``cells``, ``fixed_grid``, ``ke``, ``rtol``, ``maxiter``,
    ``coarse_cells``, ``degree`` and the fine moduli (``levels[0].e`` or
    ``_e`` through ``cell_grid``)."""

    def __init__(self, grid, device="mi355x", precision="fp32", max_rhs=1, smoother="chebyshev", omega=0.6):
        _lib.require_gpu()
        self.g = grid
        self.dev = device
        self.precision = precision
        self.max_rhs = int(max_rhs)
        opt = _lib.Options(rtol=float(grid.rtol), maxiter=int(grid.maxiter), coarse_cells=int(grid.coarse_cells),
                           degree=int(getattr(grid, "degree", 2)), smoother=_lib.SMOOTH[smoother], omega=float(omega),
                           precision=_lib.PREC[precision], max_rhs=self.max_rhs, lmax_iters=12, max_levels=0)
        self._opt = opt
        h = C.c_void_p()
        fg = np.ascontiguousarray(grid.fixed_grid, np.uint8)
        _lib.check(_lib.lib().lp_grid_create(C.byref(h), *grid.cells, _lib.ptr(fg),
                                             _lib.ptr(np.ascontiguousarray(grid.ke, np.float64)), C.byref(opt)),
                   "lp_grid_create")
        self._h = h
        # the host grid runs the STALL -> strict fallback around solve()
        # (mg.py:950-965): the device solves to the cap it is given
        _lib.lib().lp_tune(h, 9, 0)
        self.last_iters = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().lp_grid_destroy(h)
            except Exception:
                pass
            self._h = None

    def _moduli(self):
        g = self.g
        lv0 = g.levels[0] if getattr(g, "levels", None) else None
        e = getattr(lv0, "e", None)
        if e is None or not isinstance(e, np.ndarray) or e.shape != tuple(g.cells):
            e = g.cell_grid(g._e)
        return np.ascontiguousarray(e, np.float64)

    def load(self):
        """After each set_moduli of the host grid: rebuild on the device, in
        the plain or the strict hierarchy as the host grid's ``strict`` says
        (the host sets it before the rebuild that follows a stall,
        mg.py:960-962)."""
        self.opt_sync()
        _lib.lib().lp_tune(self._h, 10, int(bool(getattr(self.g, "strict", False))))
        _lib.check(_lib.lib().lp_set_moduli(self._h, _lib.ptr(self._moduli()), 0), "set_moduli")

    def opt_sync(self):
        """The host grid's rtol and maxiter onto the device handle: callers set
        ``GridElasticity.rtol`` between solves (simp.py:800) and ``_solve``
        caps maxiter at STALL around the first attempt (mg.py:953-956)."""
        self._opt.maxiter = int(self.g.maxiter)
        self._opt.rtol = float(self.g.rtol)
        _lib.check(_lib.lib().lp_set_solve_tol(self._h, self._opt.rtol, self._opt.maxiter), "set_solve_tol")

    def solve(self, b, x0):
        """b, x0: float64 node grids (masked). Returns (x, iterations, converged)."""
        self.opt_sync()
        b = np.ascontiguousarray(b, np.float64)[None]
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64)[None].copy()
        its = np.zeros(1, np.int32)
        ok = np.zeros(1, np.int32)
        rc = _lib.lib().lp_solve(self._h, 1, _lib.ptr(b), _lib.ptr(x), int(x0 is not None), 0, _lib.ptr(its),
                                 _lib.ptr(ok))
        if rc not in (_lib.LP_OK, _lib.LP_NOT_CONVERGED):
            _lib.check(rc, "solve")
        self.last_iters = int(its[0])
        return x[0], int(its[0]), bool(ok[0])
