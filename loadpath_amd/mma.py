"""The method of moving asymptotes on the GPU: mma.MMA (loadpath/mma.py:31-214)
with the same constructor, ``update`` and ``tighten``; the n-length work
runs in liblpmi.so (lp_mma_*), the m x m pieces on the host.

    opt = MMA(n, m, xmin=0.0, xmax=1.0, move=0.2)
    xnew = opt.update(x, f0, df0, f, df)     # df: (m, n)

Arrays may be numpy (host) or torch CUDA tensors (float64, contiguous;
then the new design comes back as a CUDA tensor).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

__all__ = ["MMA"]


def _is_cuda(a) -> bool:
    return hasattr(a, "is_cuda") and a.is_cuda


class MMA:
    def __init__(self, n, m, xmin=0.0, xmax=1.0, move=0.2, c=1000.0, asyinit=0.5, asyincr=1.2, asydecr=0.7):
        _lib.require_gpu()
        self.n, self.m = int(n), int(m)
        lo = np.ascontiguousarray(np.broadcast_to(np.asarray(xmin, float), (self.n,)))
        hi = np.ascontiguousarray(np.broadcast_to(np.asarray(xmax, float), (self.n,)))
        self.xmin, self.xmax = lo.copy(), hi.copy()
        h = C.c_void_p()
        _lib.check(_lib.lib().lp_mma_create(C.byref(h), self.n, self.m, _lib.ptr(lo), _lib.ptr(hi), float(move),
                                            float(c), float(asyinit), float(asyincr), float(asydecr)), "mma_create")
        self._h = h
        self._move = float(move)

    @property
    def move(self):
        """The move limit of the next update (settable, as the reference's
        loops set it)."""
        return self._move

    @move.setter
    def move(self, value):
        _lib.check(_lib.lib().lp_mma_set_move(self._h, float(value)), "mma_set_move")
        self._move = float(value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().lp_mma_destroy(h)
            except Exception:
                pass
            self._h = None

    def tighten(self, factor):
        _lib.check(_lib.lib().lp_mma_tighten(self._h, float(factor)), "mma_tighten")

    def update(self, x, f0, df0, f, df):
        fh = np.ascontiguousarray(np.asarray(f, float).reshape(self.m))
        if _is_cuda(x):
            import torch

            xd = x.detach().to(torch.float64).contiguous()
            d0 = df0.detach().to(torch.float64).contiguous()
            d1 = df.detach().to(torch.float64).reshape(self.m, self.n).contiguous()
            out = torch.empty_like(xd)
            _lib.check(_lib.lib().lp_mma_update(self._h, C.c_void_p(xd.data_ptr()), float(f0),
                                                C.c_void_p(d0.data_ptr()), _lib.ptr(fh), C.c_void_p(d1.data_ptr()),
                                                C.c_void_p(out.data_ptr()), 1), "mma_update")
            return out
        xh = np.ascontiguousarray(np.asarray(x, float).reshape(self.n))
        d0 = np.ascontiguousarray(np.asarray(df0, float).reshape(self.n))
        d1 = np.ascontiguousarray(np.asarray(df, float).reshape(self.m, self.n))
        out = np.empty(self.n)
        _lib.check(_lib.lib().lp_mma_update(self._h, _lib.ptr(xh), float(f0), _lib.ptr(d0), _lib.ptr(fh),
                                            _lib.ptr(d1), _lib.ptr(out), 0), "mma_update")
        return out

    def _get(self, what, size):
        out = np.empty(size)
        _lib.check(_lib.lib().lp_mma_get(self._h, what, _lib.ptr(out)), "mma_get")
        return out

    @property
    def low(self):
        return self._get(0, self.n)

    @property
    def upp(self):
        return self._get(1, self.n)

    @property
    def lam(self):
        return self._get(2, self.m)

    @property
    def iter(self):
        return int(self._get(3, 1)[0])
