"""The reference's element kernel switch (loadpath/femops.py) on the GPU:
``available()`` and ``element_energy(u, v, edof, ke)`` =
einsum("ij,jk,ik->i", u[edof], ke, v[edof]) with the reference's summation
order (femops.py:4-7, 43-44, 236-242), on the device (lp_element_energy).
The assembly kernels of femops serve fem.Stiffness, whose device route is
loadpath_amd.fem.Stiffness."""
from __future__ import annotations

import numpy as np

from . import _lib

__all__ = ["available", "element_energy"]


def available() -> bool:
    try:
        _lib.require_gpu()
    except Exception:
        return False
    return True


def element_energy(u, v, edof, ke):
    """Per element u_e^T Ke v_e (v None: u)."""
    u = np.ascontiguousarray(u, np.float64)
    vv = None if v is None else np.ascontiguousarray(v, np.float64)
    ed = np.ascontiguousarray(edof, np.int64)
    k = np.ascontiguousarray(ke, np.float64)
    out = np.empty(len(ed))
    _lib.check(_lib.lib().lp_element_energy(_lib.ptr(u), None if vv is None else _lib.ptr(vv), _lib.ptr(ed), len(ed),
                                            u.size, _lib.ptr(k), _lib.ptr(out)), "element_energy")
    return out
