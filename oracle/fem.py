"""Oracle: the voxel hex8 element and the mesh numbering (FP64, numpy).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows loadpath/fem.py:
- corner order bit0=x, bit1=y, bit2=z                      fem.py:30
- B matrix of the unit cube, Voigt (xx,yy,zz,xy,yz,zx)     fem.py:127-148
- incompatible bubble modes (1 - x_m^2), condensed          fem.py:172-217
- hex8_ke(nu, element), times pitch in the solver           fem.py:220-224, mg.py:446
- HexMesh node ids in ascending C-order grid index; edof    fem.py:317-374

Pinned by (tests/test_oracle_fem.py):
- "plain" Ke equals the closed-form brick stiffness of top3d (Liu & Tovar
  2014, their lk_H8), written out independently from the published table;
- both elements: symmetric, PSD, exactly 6 zero eigenvalues, rigid modes in
  the kernel (tests/test_hex8_im.py:86-94 of the reference);
- constant-strain patch test on a 3x3x3 block (test_hex8_im.py:27-43);
- "im" pure bending one cell deep exact to 1e-6 (test_hex8_im.py:76-83);
- numbering: brute force with np.unique over corner ids on random voxel sets.
"""
from __future__ import annotations

import numpy as np

CORNERS = np.array([[(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], dtype=np.int64)
GAUSS = 1.0 / np.sqrt(3.0)


def isotropic_c(e: float, nu: float) -> np.ndarray:
    """6x6 isotropic elasticity, engineering shear strains (fem.py:110-121)."""
    lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = e / (2.0 * (1.0 + nu))
    c = np.zeros((6, 6))
    c[:3, :3] = lam
    for i in range(3):
        c[i, i] = lam + 2.0 * mu
        c[3 + i, 3 + i] = mu
    return c


def _strain_rows(grad: np.ndarray) -> np.ndarray:
    """(6, 3*m) strain operator for m vector fields whose scalar gradients
    (x, y, z derivative on the unit cube) are the rows of grad (m, 3)."""
    m = grad.shape[0]
    b = np.zeros((6, 3 * m))
    for n in range(m):
        gx, gy, gz = grad[n]
        ux, uy, uz = 3 * n, 3 * n + 1, 3 * n + 2
        b[0, ux] = gx
        b[1, uy] = gy
        b[2, uz] = gz
        b[3, ux], b[3, uy] = gy, gx      # gamma_xy
        b[4, uy], b[4, uz] = gz, gy      # gamma_yz
        b[5, ux], b[5, uz] = gz, gx      # gamma_zx
    return b


def b_matrix(xi: float, eta: float, zeta: float) -> np.ndarray:
    """Trilinear strain-displacement (6, 24) of the unit cube at a natural
    point in [-1, 1]^3 (fem.py:127-148)."""
    s = 2.0 * CORNERS - 1.0
    nat = np.array([xi, eta, zeta])
    grad = np.empty((8, 3))
    for c in range(8):
        f = 1.0 + nat * s[c]           # the three 1-D factors
        # dN/dx = 2 dN/dxi on the unit cube; N = f0 f1 f2 / 8
        grad[c, 0] = 0.25 * s[c, 0] * f[1] * f[2]
        grad[c, 1] = 0.25 * s[c, 1] * f[0] * f[2]
        grad[c, 2] = 0.25 * s[c, 2] * f[0] * f[1]
    return _strain_rows(grad)


def b_bubble(xi: float, eta: float, zeta: float) -> np.ndarray:
    """Strain (6, 9) of the bubble modes (1 - x_m^2) per component
    (fem.py:172-189): d/dx_m of (1 - xi_m^2) on the unit cube is -4 xi_m."""
    nat = (xi, eta, zeta)
    grad = np.zeros((3, 3))
    for m in range(3):
        grad[m, m] = -4.0 * nat[m]
    return _strain_rows(grad)


def _gauss_points():
    return [(a, b, c) for a in (-GAUSS, GAUSS) for b in (-GAUSS, GAUSS) for c in (-GAUSS, GAUSS)]


def element_parts(c: np.ndarray, element: str):
    """(ke (24,24), R (9,24) or None) of the unit cube for elasticity c
    (fem.py:191-217). Gauss weight 1, det J = 1/8."""
    kuu = np.zeros((24, 24))
    kua = np.zeros((24, 9))
    kaa = np.zeros((9, 9))
    for p in _gauss_points():
        b = b_matrix(*p)
        ba = b_bubble(*p)
        kuu += 0.125 * (b.T @ c @ b)
        kua += 0.125 * (b.T @ c @ ba)
        kaa += 0.125 * (ba.T @ c @ ba)
    if element == "plain":
        return 0.5 * (kuu + kuu.T), None
    if element != "im":
        raise ValueError(f"element must be 'im' or 'plain', not {element!r}")
    r = -np.linalg.solve(kaa, kua.T)       # bubble amplitudes per nodal dof
    ke = kuu + kua @ r
    return 0.5 * (ke + ke.T), r


def hex8_ke(nu: float, element: str = "im") -> np.ndarray:
    """Unit-cube Ke at E = 1 (fem.py:220-224)."""
    return element_parts(isotropic_c(1.0, nu), element)[0]


def hex8_b(xi, eta, zeta, nu, element: str = "im") -> np.ndarray:
    """Strain-displacement incl. the condensed bubbles (fem.py:227-237)."""
    b = b_matrix(xi, eta, zeta)
    r = element_parts(isotropic_c(1.0, nu), element)[1]
    if r is not None:
        b = b + b_bubble(xi, eta, zeta) @ r
    return b


class Mesh:
    """Active voxels and their nodes (fem.HexMesh, fem.py:317-374): node ids
    ascend in C order over the node grid (max cell index + 2 per axis)."""

    def __init__(self, elem_ijk, pitch: float = 1.0, origin=(0.0, 0.0, 0.0)):
        self.elem_ijk = np.asarray(elem_ijk, np.int64).reshape(-1, 3)
        if len(self.elem_ijk) == 0:
            raise ValueError("no active elements")
        self.pitch = float(pitch)
        self.origin = np.asarray(origin, float)
        self.node_dims = self.elem_ijk.max(axis=0) + 2
        nd = self.node_dims
        corner_nodes = self.elem_ijk[:, None, :] + CORNERS[None, :, :]          # (ne, 8, 3)
        flat = (corner_nodes[..., 0] * nd[1] + corner_nodes[..., 1]) * nd[2] + corner_nodes[..., 2]
        used, inv = np.unique(flat.ravel(), return_inverse=True)
        self.enode = inv.reshape(flat.shape).astype(np.int64)
        self.node_ijk = np.stack(np.unravel_index(used, tuple(nd)), axis=1).astype(np.int64)
        self.n_nodes = len(used)
        self.n_dof = 3 * self.n_nodes
        self._flat_used = used

    @property
    def n_elem(self):
        return len(self.elem_ijk)

    @property
    def edof(self):
        return (3 * self.enode[:, :, None] + np.arange(3)).reshape(self.n_elem, 24)

    def node_id(self, ijk):
        ijk = np.asarray(ijk, np.int64)
        nd = self.node_dims
        ok = np.all((ijk >= 0) & (ijk < nd), axis=-1)
        flat = (ijk[..., 0] * nd[1] + ijk[..., 1]) * nd[2] + ijk[..., 2]
        pos = np.searchsorted(self._flat_used, np.where(ok, flat, -1))
        pos = np.clip(pos, 0, self.n_nodes - 1)
        hit = ok & (self._flat_used[pos] == flat)
        return np.where(hit, pos, -1)


def box_mesh(nx, ny, nz, pitch=1.0):
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return Mesh(np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1), pitch)


def assemble(mesh: Mesh, ke: np.ndarray, e: np.ndarray):
    """Brute-force global K = sum_c e_c Ke scattered by edof (COO, FP64), the
    evaluator's assembly (benchmarks/evaluate.py:103-113)."""
    import scipy.sparse as sp

    edof = mesh.edof
    rows = np.repeat(edof, 24, axis=1).ravel()
    cols = np.tile(edof, (1, 24)).ravel()
    data = (np.asarray(e, float)[:, None, None] * ke[None]).ravel()
    return sp.csr_matrix((data, (rows, cols)), shape=(mesh.n_dof, mesh.n_dof))


def rigid_modes(xyz: np.ndarray) -> np.ndarray:
    """(3n, 6) translations and infinitesimal rotations of node positions."""
    n = len(xyz)
    x, y, z = (xyz - xyz.mean(axis=0)).T
    b = np.zeros((3 * n, 6))
    for i in range(3):
        b[i::3, i] = 1.0
    b[0::3, 3], b[1::3, 3] = -y, x
    b[1::3, 4], b[2::3, 4] = -z, y
    b[0::3, 5], b[2::3, 5] = z, -x
    return b


def von_mises_matrix() -> np.ndarray:
    """vm^2 = s^T V s for Voigt stress (constrained.py:107-114)."""
    v = np.zeros((6, 6))
    v[:3, :3] = -0.5
    for i in range(3):
        v[i, i] = 1.0
        v[3 + i, 3 + i] = 3.0
    return v


def stiffness_csr(mesh: Mesh, ke: np.ndarray, scale, fixed):
    """fem.Stiffness(mesh, nu, fixed).assemble(scale) (fem.py:377-526), as its
    numpy path builds it: (data, indices, indptr). Rows 3n+i; for each of
    the node's neighbours (27 offsets, dx outermost, kept where a cell joins
    the two) the columns 3m+j; values scale_e Ke[a, b] added pair by pair
    (a, then b) over the elements; Dirichlet rows and columns zeroed and a
    fixed diagonal set to 1."""
    enode = mesh.enode
    nn = mesh.n_nodes
    fixed = np.asarray(fixed, bool)
    corners = CORNERS

    def off_index(d):
        return (d[..., 0] + 1) * 9 + (d[..., 1] + 1) * 3 + (d[..., 2] + 1)

    adj = np.zeros((nn, 27), bool)
    pair_off = np.empty((8, 8), np.int64)
    for a in range(8):
        for b in range(8):
            pair_off[a, b] = off_index(corners[b] - corners[a])
            adj[enode[:, a], pair_off[a, b]] = True
    cnt = adj.sum(axis=1)
    rank = np.cumsum(adj, axis=1) - 1
    row_len = np.repeat(3 * cnt, 3)
    indptr = np.concatenate([[0], np.cumsum(row_len)]).astype(np.int64)
    nnz = int(indptr[-1])
    indices = np.empty(nnz, np.int64)
    offsets = np.array([[dx, dy, dz] for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)])
    lookup = {tuple(ijk): i for i, ijk in enumerate(mesh.node_ijk)}
    n_idx, o_idx = np.nonzero(adj)
    nbr = np.array([lookup[tuple(v)] for v in mesh.node_ijk[n_idx] + offsets[o_idx]], np.int64)
    for i in range(3):
        base = indptr[3 * n_idx + i] + 3 * rank[n_idx, o_idx]
        for j in range(3):
            indices[base + j] = 3 * nbr + j
    data = np.zeros(nnz)
    s = np.asarray(scale, float)[:, None]
    for a in range(8):
        na = enode[:, a]
        row = 3 * cnt[na]
        for b in range(8):
            base = indptr[3 * na] + 3 * rank[na, pair_off[a, b]]
            slots = base[:, None] + np.repeat(np.arange(3), 3)[None, :] * row[:, None] + np.tile(np.arange(3), 3)[None]
            data[slots] += s * ke[3 * a:3 * a + 3, 3 * b:3 * b + 3].ravel()[None]
    row_of = np.repeat(np.arange(mesh.n_dof), row_len)
    data *= ~(fixed[row_of] | fixed[indices])
    data[np.nonzero((row_of == indices) & fixed[row_of])[0]] = 1.0
    return data, indices, indptr
