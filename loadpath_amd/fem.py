"""Voxel hex8 element and mesh numbering (host side of the drop-in).

- ``hex8_ke(nu, element)``: fem.hex8_ke (loadpath/fem.py:220-224), computed
  by liblpmi's FP64 host routine (lp_hex8_ke).
- ``hex8_element()``: the LOADPATH_HEX8 switch (fem.py:151-160).
- ``HexMesh``: fem.HexMesh (fem.py:317-374): node ids ascend in C order over
  the node grid, ``edof = 3 * node + component``.
- ``Stiffness``: fem.Stiffness (fem.py:377-526), the assembled route's CSR
  stiffness: its layout built once by liblpmi's host code, ``assemble(scale)``
  on the GPU (lp_stiffness_*).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

CORNERS = np.array([[(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)])
_ELEMENTS = ("im", "plain")


def hex8_element() -> str:
    e = os.environ.get("LOADPATH_HEX8", "im").strip().lower() or "im"
    if e not in _ELEMENTS:
        raise ValueError(f"LOADPATH_HEX8 must be one of {_ELEMENTS}, not {e!r}")
    return e


def hex8_ke(nu: float, element: str | None = None) -> np.ndarray:
    element = element or hex8_element()
    if element not in _ELEMENTS:
        raise ValueError(f"element must be one of {_ELEMENTS}, not {element!r}")
    out = np.empty((24, 24))
    _lib.check(_lib.lib().lp_hex8_ke(float(nu), 0 if element == "im" else 1, _lib.ptr(out)), "lp_hex8_ke")
    return out


class HexMesh:
    """Active voxels of a grid and the nodes they use (fem.py:317-374).

    The contract is loadpath's numbering: node ids ascend with the node's
    C-order index on the (max + 2)-node grid, and ``edof = 3 * node +
    component``. Two constructions give it: a full box of cells starting at
    (0, 0, 0) numbers every grid node, so ids and corner ids are closed
    forms; any other set takes the sorted distinct corner keys, where a
    node's id is its rank among them (``np.unique``'s inverse) and
    ``node_id`` is a binary search."""

    def __init__(self, elem_ijk: np.ndarray, pitch: float, origin=(0.0, 0.0, 0.0)):
        self.elem_ijk = np.asarray(elem_ijk, np.int64).reshape(-1, 3)
        self.pitch = float(pitch)
        self.origin = np.asarray(origin, float)
        if len(self.elem_ijk) == 0:
            raise ValueError("no active elements")
        if self.elem_ijk.min() < 0:
            raise ValueError("cell indices must be non-negative")
        cells = self.elem_ijk.max(axis=0) + 1
        self.node_dims = cells + 1
        nd = tuple(int(d) for d in self.node_dims)
        stride = np.array([nd[1] * nd[2], nd[2], 1], np.int64)
        corner_step = CORNERS @ stride                        # (8,) key offsets of the corners
        cell_key = self.elem_ijk @ stride                     # key of each cell's corner 0
        self._box = len(self.elem_ijk) == int(np.prod(cells)) and self._is_c_order_box(cells)
        if self._box:
            # every node of the grid is used: id == key
            self._keys = None
            self.enode = cell_key[:, None] + corner_step[None, :]
            self.n_nodes = int(np.prod(nd))
            ids = np.arange(self.n_nodes, dtype=np.int64)
        else:
            keys, inv = np.unique((cell_key[:, None] + corner_step[None, :]).ravel(), return_inverse=True)
            self._keys = keys
            self.enode = inv.reshape(-1, 8).astype(np.int64)
            self.n_nodes = len(keys)
            ids = keys
        self.node_ijk = np.stack([ids // stride[0], (ids // stride[1]) % nd[1], ids % nd[2]], axis=1)
        self.node_xyz = self.origin + self.node_ijk * self.pitch
        self.n_dof = 3 * self.n_nodes
        self._stride = stride
        self._edof = None

    def _is_c_order_box(self, cells) -> bool:
        """The cells are exactly the box [0, cells) listed in C order."""
        ijk = self.elem_ijk
        want = np.arange(len(ijk), dtype=np.int64)
        got = (ijk[:, 0] * cells[1] + ijk[:, 1]) * cells[2] + ijk[:, 2]
        return bool(np.array_equal(got, want))

    @property
    def n_elem(self):
        return len(self.elem_ijk)

    @property
    def edof(self):
        if self._edof is None:
            self._edof = (3 * self.enode[:, :, None] + np.arange(3)).reshape(self.n_elem, 24)
        return self._edof

    def elem_centers(self):
        return self.origin + (self.elem_ijk + 0.5) * self.pitch

    def node_id(self, ijk) -> np.ndarray:
        """Node id at integer node coordinates, -1 when absent or outside."""
        ijk = np.asarray(ijk, np.int64)
        inside = np.all((ijk >= 0) & (ijk < self.node_dims), axis=-1)
        key = np.where(inside, ijk @ self._stride, 0)
        if self._box:
            return np.where(inside, key, -1)
        pos = np.searchsorted(self._keys, key)
        pos_c = np.minimum(pos, self.n_nodes - 1)
        hit = inside & (self._keys[pos_c] == key)
        return np.where(hit, pos_c, -1).astype(np.int64)


def box_mesh(nx: int, ny: int, nz: int, pitch: float = 1.0) -> HexMesh:
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return HexMesh(np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1), pitch)


class Stiffness:
    """fem.Stiffness(mesh, nu, fixed, element) (fem.py:377-526): the global
    stiffness of a voxel mesh in a fixed CSR layout, Dirichlet conditions as
    identity rows. ``assemble(scale)`` (per-element moduli, mesh order)
    returns a scipy CSR matrix whose values are assembled on the GPU, equal
    to the reference's to the bit; ``indptr``, ``indices``, ``nnz``,
    ``ke`` (pitch included), ``fixed``, ``free`` as the reference's."""

    def __init__(self, mesh, nu, fixed, element: str | None = None):
        _lib.require_gpu()
        self.mesh = mesh
        self.nu = nu
        self.element = element or hex8_element()
        self.ke = hex8_ke(nu, self.element) * mesh.pitch
        fixed = np.asarray(fixed, bool)
        if fixed.shape != (mesh.n_dof,):
            raise ValueError("fixed must be a bool per dof")
        self.fixed = fixed
        self.free = ~fixed
        lo = mesh.elem_ijk.min(axis=0)
        dims = tuple(int(d) for d in mesh.elem_ijk.max(axis=0) - lo + 1)
        self._dims = dims
        e = mesh.elem_ijk - lo
        self._cells = np.ravel_multi_index((e[:, 0], e[:, 1], e[:, 2]), dims)
        present = np.zeros(int(np.prod(dims)), np.uint8)
        present[self._cells] = 1
        nd = mesh.node_ijk - lo
        nodes = np.ravel_multi_index((nd[:, 0], nd[:, 1], nd[:, 2]), tuple(d + 1 for d in dims))
        if np.any(np.diff(nodes) <= 0):
            raise ValueError("mesh node ids must ascend in C order over the node grid")
        fg = np.zeros((int(np.prod([d + 1 for d in dims])), 3), np.uint8)
        fg[nodes] = fixed.reshape(-1, 3)
        h = C.c_void_p()
        n_dof, nnz = C.c_longlong(), C.c_longlong()
        _lib.check(_lib.lib().lp_stiffness_create(C.byref(h), *dims, _lib.ptr(present), _lib.ptr(fg),
                                                  _lib.ptr(np.ascontiguousarray(self.ke)), C.byref(n_dof),
                                                  C.byref(nnz)), "stiffness_create")
        self._h = h
        if n_dof.value != mesh.n_dof:
            raise ValueError("the mesh's nodes are not the nodes its cells use")
        self.nnz = int(nnz.value)
        self.indptr = np.empty(mesh.n_dof + 1, np.int64)
        self.indices = np.empty(self.nnz, np.int64)
        _lib.check(_lib.lib().lp_stiffness_pattern(h, _lib.ptr(self.indptr), _lib.ptr(self.indices)),
                   "stiffness_pattern")
        self._scale = np.zeros(int(np.prod(dims)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().lp_stiffness_destroy(h)
            except Exception:
                pass
            self._h = None

    def assemble_data(self, scale) -> np.ndarray:
        """The CSR values for per-element moduli scale (mesh order)."""
        self._scale[self._cells] = np.asarray(scale, float)
        data = np.empty(self.nnz)
        _lib.check(_lib.lib().lp_stiffness_assemble(self._h, _lib.ptr(self._scale), _lib.ptr(data), 0),
                   "stiffness_assemble")
        return data

    def assemble(self, scale):
        import scipy.sparse as sp

        k = sp.csr_matrix((self.assemble_data(scale), self.indices, self.indptr),
                          shape=(self.mesh.n_dof, self.mesh.n_dof))
        k.has_sorted_indices = True
        return k


def band_order(mesh) -> np.ndarray:
    """The dof order of the band factor: mesh nodes sorted with the longest
    axis of their box slowest (the band is then 3 x the node count of a
    cross-section of the two shorter axes), three dofs per node; new
    position -> mesh dof."""
    ijk = np.asarray(mesh.node_ijk)
    ext = ijk.max(axis=0) - ijk.min(axis=0)
    slow, mid, fast = np.argsort(-ext, kind="stable")
    nodes = np.lexsort((ijk[:, fast], ijk[:, mid], ijk[:, slow]))
    return (3 * nodes[:, None] + np.arange(3)[None, :]).ravel().astype(np.int64)


class DirectSolver:
    """direct.DirectSolver (direct.py:158-246) on the device: one Cholesky
    factor of the assembled stiffness, shared by every right-hand side of
    that matrix, as a block-tridiagonal factor of its band in HBM
    (lp_band_*). ``stiffness``: a Stiffness of the mesh (its pattern is the
    reference's fem.Stiffness layout). ``solve_many(k, fs)`` factors k (a CSR
    matrix with that pattern, e.g. Stiffness.assemble's) unless it is the
    matrix already factored, then solves with one step of iterative
    refinement; ``factor(scale)`` assembles and factors on the device
    directly. Raises RuntimeError on a non-positive pivot (a mechanism)."""

    def __init__(self, stiffness: Stiffness, refine: int = 1):
        _lib.require_gpu()
        self.stiffness = stiffness
        self.refine = int(refine)
        self.perm = band_order(stiffness.mesh)
        h = C.c_void_p()
        band, nbytes = C.c_longlong(), C.c_longlong()
        _lib.check(_lib.lib().lp_band_create(C.byref(h), stiffness._h, _lib.ptr(self.perm), C.byref(band),
                                             C.byref(nbytes)), "band_create")
        self._h = h
        self.band = int(band.value)
        self.factor_bytes = int(nbytes.value)
        self._k = None
        self.last_relres = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().lp_band_destroy(h)
            except Exception:
                pass
            self._h = None

    def _check(self, rc, what):
        if rc == _lib.LP_SINGULAR:
            raise RuntimeError(f"{what}: the factorization met a non-positive pivot; "
                               "the structure may be a mechanism (check supports)")
        _lib.check(rc, what)

    def factor(self, scale):
        """Assemble K for per-element moduli scale (mesh order) and factor it."""
        st = self.stiffness
        st._scale[st._cells] = np.asarray(scale, float)
        self._k = None
        self._check(_lib.lib().lp_band_factor(self._h, _lib.ptr(st._scale), len(st._scale), 0), "band_factor")
        self._k = "scale"

    def _factor_csr(self, k):
        st = self.stiffness
        if k.shape != (st.mesh.n_dof, st.mesh.n_dof) or k.nnz != st.nnz or not np.array_equal(k.indptr, st.indptr):
            raise ValueError("the matrix does not have this stiffness's pattern")
        self._k = None
        self._check(_lib.lib().lp_band_factor_csr(self._h, _lib.ptr(np.ascontiguousarray(k.data, np.float64)), 0),
                    "band_factor")
        self._k = k

    def solve_many(self, k, fs, x0s=None):
        """x_i = K^-1 f_i for each right-hand side (x0s ignored: a direct
        solve needs no start)."""
        del x0s
        if k is not None and k is not self._k:
            self._factor_csr(k)
        if self._k is None:
            raise RuntimeError("DirectSolver: nothing factored")
        fs = [np.asarray(f, float) for f in fs]
        out = []
        for j in range(0, len(fs), 64):
            f = np.ascontiguousarray(np.stack(fs[j:j + 64]))
            x = np.empty_like(f)
            rr = C.c_double()
            self._check(_lib.lib().lp_band_solve(self._h, f.shape[0], _lib.ptr(f), _lib.ptr(x), 0, self.refine,
                                                 C.byref(rr)), "band_solve")
            self.last_relres = float(rr.value)
            out.extend(x)
        return out


class Solver:
    """fem.Solver(mesh, rtol, maxiter, ...) (fem.py:586-792) for the
    assembled route, on the device: ``solve(k, f, x0)`` / ``solve_many(k, fs,
    x0s)`` with k a CSR stiffness in the reference's fem.Stiffness layout.
    The reference runs CG with a smoothed-aggregation AMG preconditioner and
    falls back on a direct factor (direct.py) where CG stalls; here every
    matrix is factored once on the device (DirectSolver: its band Cholesky
    in HBM) and its right-hand sides are solved directly with iterative
    refinement against K itself, the steps doubled (1, 2, 4, 8, at most
    maxiter) until |f - K x| <= rtol |f|; one suffices in practice.
    last_iters: the refinement steps of the last solve."""

    def __init__(self, mesh, rtol=1e-6, maxiter=800, rebuild_after_iters=25, refresh=True):
        del rebuild_after_iters, refresh   # the AMG hierarchy's knobs: nothing to rebuild here
        self.mesh = mesh
        self.rtol = float(rtol)
        self.maxiter = int(maxiter)
        self.last_iters = 0
        self._direct = None
        self._pattern = None

    def invalidate(self):
        if self._direct is not None:
            self._direct._k = None

    def _solver(self, k):
        if self._direct is None:
            # the pattern is the mesh's: any ke and supports give it
            self._pattern = Stiffness(self.mesh, 0.3, np.zeros(self.mesh.n_dof, bool))
            self._direct = DirectSolver(self._pattern)
        return self._direct

    def solve_many(self, k, fs, x0s=None):
        d = self._solver(k)
        fs = [np.asarray(f, float) for f in fs]
        steps = 1
        while True:
            d.refine = steps
            xs = d.solve_many(k, fs)
            if d.last_relres <= self.rtol or steps >= min(self.maxiter, 8):
                break
            steps *= 2
        self.last_iters = steps
        if d.last_relres > self.rtol:
            raise RuntimeError(f"linear solve did not converge (relative residual {d.last_relres:.2e} after "
                               f"{steps} refinement steps); the structure may be a mechanism (check supports)")
        return xs

    def solve(self, k, f, x0=None):
        return self.solve_many(k, [f])[0]
