"""The reference's multigrid kernel switches (loadpath/mgops.py) on the GPU.

Same names, arguments and results as mgops (mgops.py:53-54, 220-261,
557-585, 753-773):

- ``available()``: True when liblpmi.so loads and a GPU is visible (the
  reference: numba present and LOADPATH_MGOPS != "numpy");
- ``Cells(active)``: a level's cells that carry stiffness (rows, n);
- ``apply_masked(kind, u, a, b, cells, m, out, finish)``: out = K (u m),
  and with finish out = out m + (1 - m) u, for kind "scaled" (a = Ke,
  b = e), "weights" (a = M_q, b = w) or "explicit" (a = K_c);
- ``coarse_residual(b, ax, m, cm, coarse_shape)``: restrict(((b - ax) m)
  padded) cm;
- ``add_correction(xc, cm, m, x) -> bool``: x += prolong(xc cm) m;
- ``dot(a, b) -> float``, ``cast(a, dtype)``.

Every call runs on the device (lp_cells_apply, lp_coarse_residual,
lp_add_correction, lp_dot, lp_cast); host arrays in and out.
"""
from __future__ import annotations

import numpy as np

from . import _lib

__all__ = ["available", "Cells", "apply_masked", "coarse_residual", "add_correction", "dot", "cast"]

_KIND = {"scaled": 0, "weights": 1, "explicit": 2}


def available() -> bool:
    try:
        _lib.require_gpu()
    except Exception:
        return False
    return True


class Cells:
    """mgops.Cells (mgops.py:220-244): the cells of a level that carry
    stiffness, x-major; rows[x, y, z] a cell's index in that list or -1."""

    def __init__(self, active: np.ndarray):
        ijk = np.argwhere(active)
        self.cells = np.ascontiguousarray(ijk, np.int64)
        self.n = len(ijk)
        self.rows = np.full(active.shape, -1, np.int64)
        self.rows[tuple(ijk.T)] = np.arange(self.n)


def _f(a, dt):
    return np.ascontiguousarray(a, dt)


def apply_masked(kind, u, a, b, cells, m, out, finish):
    """K (u m) into ``out`` (u, m, out: (3, nodes...) of one dtype); with
    finish also out = out m + (1 - m) u (mg._Level.apply)."""
    dt = np.dtype(u.dtype)
    if dt not in (np.float32, np.float64):
        raise TypeError("apply_masked: float32 or float64 fields")
    nx, ny, nz = cells.rows.shape
    uu = _f(u, dt)
    mm = None if (m is None or np.size(m) <= 1) else _f(m, dt)
    aa = _f(a, dt)
    bb = None if kind == "explicit" else _f(b, dt)
    res = np.empty(uu.shape, dt)
    rows = _f(cells.rows, np.int64)
    _lib.check(_lib.lib().lp_cells_apply(_KIND[kind], nx, ny, nz, _lib.ptr(rows), cells.n, _lib.ptr(aa),
                                         None if bb is None else _lib.ptr(bb), _lib.ptr(uu),
                                         None if mm is None else _lib.ptr(mm), _lib.ptr(res), int(bool(finish)),
                                         dt.itemsize), "apply_masked")
    out[...] = res.reshape(out.shape)
    return out


def coarse_residual(b, ax, m, cm, coarse_shape):
    """restrict(((b - ax) m) zero-padded to the grid twice coarse_shape's
    cells) cm."""
    dt = np.dtype(b.dtype)
    if not (ax.dtype == m.dtype == cm.dtype == dt) or dt not in (np.float32, np.float64):
        return None
    fx, fy, fz = b.shape[1:]
    cx, cy, cz = coarse_shape[1:]
    c = np.empty(tuple(coarse_shape), dt)
    _lib.check(_lib.lib().lp_coarse_residual(_lib.ptr(_f(b, dt)), _lib.ptr(_f(ax, dt)), _lib.ptr(_f(m, dt)),
                                             _lib.ptr(_f(cm, dt)), _lib.ptr(c), fx, fy, fz, cx, cy, cz,
                                             dt.itemsize), "coarse_residual")
    return c


def add_correction(xc, cm, m, x) -> bool:
    """x += prolong(xc cm)[:, :x's nodes] m in one pass; False when it
    cannot (mixed dtypes, a non-contiguous x)."""
    dt = np.dtype(x.dtype)
    if not (xc.dtype == cm.dtype == m.dtype == dt) or dt not in (np.float32, np.float64) or \
            not x.flags.c_contiguous:
        return False
    cx, cy, cz = xc.shape[1:]
    sx, sy, sz = x.shape[1:]
    _lib.check(_lib.lib().lp_add_correction(_lib.ptr(_f(xc, dt)), _lib.ptr(_f(cm, dt)), _lib.ptr(_f(m, dt)),
                                            _lib.ptr(x), cx, cy, cz, sx, sy, sz, dt.itemsize), "add_correction")
    return True


def dot(a, b) -> float:
    """vdot of two same-shape fields, summed in double on the device."""
    dt = np.result_type(a, b)
    dt = np.float64 if dt == np.float64 else np.float32
    out = np.zeros(1)
    aa, bb = _f(a, dt).ravel(), _f(b, dt).ravel()
    _lib.check(_lib.lib().lp_dot(_lib.ptr(aa), _lib.ptr(bb), aa.size, np.dtype(dt).itemsize, _lib.ptr(out)), "dot")
    return float(out[0])


def cast(a, dtype):
    """a.astype(dtype) (a copy when the dtype is the same)."""
    dtype = np.dtype(dtype)
    if a.dtype == dtype:
        return a.copy()
    if a.dtype not in (np.float32, np.float64) or dtype not in (np.float32, np.float64):
        return a.astype(dtype)
    src = np.ascontiguousarray(a)
    out = np.empty(a.shape, dtype)
    _lib.check(_lib.lib().lp_cast(_lib.ptr(src), src.dtype.itemsize, _lib.ptr(out), dtype.itemsize, src.size), "cast")
    return out
