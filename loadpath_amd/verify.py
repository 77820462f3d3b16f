"""Stress recovery of loadpath's verification on the device:
verify.nodal_stress / nodal_von_mises (verify.py:197-206, 307-337).

Each cell's stress is evaluated at its own 8 corners (the element's strain
there, incompatible-mode bubbles included for the "im" element), scaled by
the cell's modulus, and averaged over the cells sharing a node — on the GPU
(liblpmi.so lp_nodal_stress), one thread per node over its cells in a fixed
order, FP64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .fem import hex8_element


def _grid_of(mesh):
    """The cell box of a HexMesh, its present cells and the grid index of
    each mesh node."""
    lo = mesh.elem_ijk.min(axis=0)
    dims = tuple(int(d) for d in mesh.elem_ijk.max(axis=0) - lo + 1)
    present = np.zeros(dims, np.uint8)
    e = mesh.elem_ijk - lo
    present[e[:, 0], e[:, 1], e[:, 2]] = 1
    nd = mesh.node_ijk - lo
    nodes = np.ravel_multi_index((nd[:, 0], nd[:, 1], nd[:, 2]), tuple(d + 1 for d in dims))
    cells = np.ravel_multi_index((e[:, 0], e[:, 1], e[:, 2]), dims)
    return dims, present, nodes, cells


def nodal_stress(stiff, u, modulus, thermal=None, *, vm_out=False):
    """verify.nodal_stress(stiff, u, modulus, thermal) (verify.py:307-337):
    (n_nodes, 6) nodal-averaged stress tensors. stiff: anything with .mesh
    (a HexMesh), .nu and optionally .element and .stress_mask (per element),
    e.g. a GridElasticity or a fem.Stiffness-like record. modulus: one
    number or one per element; thermal: a free expansion strain per
    element, or None."""
    _lib.require_gpu()
    mesh = stiff.mesh
    element = getattr(stiff, "element", None) or hex8_element()
    dims, present, nodes, cells = _grid_of(mesh)
    ncell = int(np.prod(dims))
    nn = int(np.prod([d + 1 for d in dims]))
    ug = np.zeros((nn, 3))
    ug[nodes] = np.asarray(u, float).reshape(-1, 3)
    mod_arr = None
    mod_scalar = 0.0
    m = np.asarray(modulus, float)
    if m.ndim == 0:
        mod_scalar = float(m)
    else:
        mod_arr = np.zeros(ncell)
        mod_arr[cells] = np.broadcast_to(m, (mesh.n_elem,))
    th = None
    if thermal is not None:
        th = np.zeros(ncell)
        th[cells] = np.broadcast_to(np.asarray(thermal, float), (mesh.n_elem,))
    cnt = None
    mask = getattr(stiff, "stress_mask", None)
    if mask is not None:
        cnt = np.zeros(ncell, np.uint8)
        cnt[cells] = np.asarray(mask, bool)
    out = np.empty((nn, 6))
    vm = np.empty(nn) if vm_out else None
    _lib.check(_lib.lib().lp_nodal_stress(*dims, float(mesh.pitch), float(stiff.nu), 0 if element == "im" else 1,
                                          _lib.ptr(ug), None if mod_arr is None else _lib.ptr(mod_arr), mod_scalar,
                                          None if th is None else _lib.ptr(th), _lib.ptr(present.ravel()),
                                          None if cnt is None else _lib.ptr(cnt),
                                          _lib.ptr(out), None if vm is None else _lib.ptr(vm), 0), "nodal_stress")
    return (out[nodes], vm[nodes]) if vm_out else out[nodes]


def nodal_von_mises(stiff, u, modulus, thermal=None):
    """verify.nodal_von_mises (verify.py:197-206): fem.von_mises of the
    nodal-averaged stress."""
    return nodal_stress(stiff, u, modulus, thermal, vm_out=True)[1]



class MaterialAverage:
    """verify._MaterialAverage(node_xyz, enode, fill, pitch, radius)
    (verify.py:214-305) on the device, for a HexMesh: () applied to nodal
    tensors sig (n_nodes, 6) gives each node's stress averaged over the
    material within radius (lp_material_average), with the reference's
    samples (quarter points, centres, blocks of cells past 3.3 cells).
    fill: the filled share of each element (None: full)."""

    def __init__(self, mesh, fill=None, radius=1.0):
        self.mesh = mesh
        self.radius = float(radius)
        self._dims, _, self._nodes, cells = _grid_of(mesh)
        self._fill = np.zeros(int(np.prod(self._dims)))
        self._fill[cells] = 1.0 if fill is None else np.broadcast_to(np.asarray(fill, float), (mesh.n_elem,))

    def __call__(self, sig):
        _lib.require_gpu()
        nn = int(np.prod([d + 1 for d in self._dims]))
        sg = np.zeros((nn, 6))
        sg[self._nodes] = np.asarray(sig, float).reshape(-1, 6)
        out = np.empty((nn, 6))
        _lib.check(_lib.lib().lp_material_average(*self._dims, float(self.mesh.pitch), self.radius,
                                                  _lib.ptr(self._fill), _lib.ptr(sg), _lib.ptr(out)),
                   "material_average")
        return out[self._nodes]


def material_average(mesh, sig, fill=None, radius=1.0):
    """MaterialAverage(mesh, fill, radius)(sig)."""
    return MaterialAverage(mesh, fill, radius)(sig)
