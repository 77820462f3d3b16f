"""Oracle: matrix-free geometric multigrid + PCG for K(E) u = f on a voxel
grid (FP64 outer CG; V-cycle in a chosen dtype), numpy only.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates loadpath/mg.py and the arithmetic of loadpath/mgops.py:
- trilinear prolongation / restriction (its transpose)       mg.py:84-132
- child prolongations P_k (8, 24, 24)                        mg.py:57-73
- padded even cell grid, node grid (3, nx+1, ny+1, nz+1),
  fixed mask, to_grid / from_grid / cell_grid                mg.py:428-511
- level operator m*K(m*u) + (1-m)*u                          mg.py:317-346
- level diagonal, alive / mask / dinv                        mg.py:348-371, 532-541
- Galerkin coarsening: first level with the fine mask
  (the reference's weights+patch form), deeper levels
  unmasked unless strict; live = d > 1e-12 max d             mg.py:552-677
- strict hierarchy (held injection, kept >= 1)               mg.py:561-574
- coarsest level assembled D K D + (I - D), exact solve      mg.py:708-766
- lambda_max by 12 power iterations, x1.3                    mg.py:77-81, 771-794
- Chebyshev smoother on [lmax/30, lmax]                      mg.py:796-835
- damped Jacobi V(nu1, nu2) (the configs' smoother; not in the reference)
- V-cycle                                                    mg.py:837-860
- PCG with energy-optimal warm start                         mg.py:872-926
- solve_many with the STALL -> strict fallback               mg.py:928-965
- strain_energy u_e^T Ke u_e at unit modulus                 mg.py:975-988

Pinned by (tests/test_oracle_mg.py):
- transfers adjoint to 1e-10, linear fields prolong exactly (test_mg.py:56-63);
- level operators equal brute-force assembled K (oracle.fem.assemble) and
  the explicit Galerkin product P^T D K D P built from sparse prolongation
  rows (closed form) on tiny grids;
- converged PCG solutions equal scipy spsolve on the free dofs to <1e-6
  relative (test_mg.py:66-78), both smoothers, warm starts;
- strain energy equals einsum over the assembled element gather;
- lambda_max bounded by / close to scipy eigsh of D^-1 A.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .fem import CORNERS, hex8_ke

LMAX_SAFETY = 1.3
STALL = 100
CORNER_OFF = [tuple(int(v) for v in c) for c in CORNERS]


# --------------------------------------------------------------------------
# grid transfers
# --------------------------------------------------------------------------

def _prolong_axis(a: np.ndarray, ax: int) -> np.ndarray:
    n = a.shape[ax]
    shape = list(a.shape)
    shape[ax] = 2 * n - 1
    f = np.empty(shape, a.dtype)
    ev = [slice(None)] * a.ndim
    od = [slice(None)] * a.ndim
    ev[ax] = slice(0, None, 2)
    od[ax] = slice(1, None, 2)
    lo = [slice(None)] * a.ndim
    hi = [slice(None)] * a.ndim
    lo[ax] = slice(0, n - 1)
    hi[ax] = slice(1, n)
    f[tuple(ev)] = a
    f[tuple(od)] = (a[tuple(lo)] + a[tuple(hi)]) * a.dtype.type(0.5)
    return f


def _restrict_axis(f: np.ndarray, ax: int) -> np.ndarray:
    n = (f.shape[ax] + 1) // 2
    ev = [slice(None)] * f.ndim
    od = [slice(None)] * f.ndim
    ev[ax] = slice(0, None, 2)
    od[ax] = slice(1, None, 2)
    c = f[tuple(ev)].copy()
    half = f[tuple(od)] * f.dtype.type(0.5)
    lo = [slice(None)] * f.ndim
    hi = [slice(None)] * f.ndim
    lo[ax] = slice(0, n - 1)
    hi[ax] = slice(1, n)
    c[tuple(lo)] += half
    c[tuple(hi)] += half
    return c


def prolong(c: np.ndarray) -> np.ndarray:
    """(..., n0, n1, n2) -> (..., 2n0-1, 2n1-1, 2n2-1), trilinear (mg.py:84-108)."""
    out = c
    for ax in (-3, -2, -1):
        out = _prolong_axis(out, out.ndim + ax)
    return out


def restrict(f: np.ndarray) -> np.ndarray:
    """Transpose of prolong (mg.py:111-132)."""
    out = f
    for ax in (-3, -2, -1):
        out = _restrict_axis(out, out.ndim + ax)
    return out


def child_prolongations() -> np.ndarray:
    """P_k (8, 24, 24): coarse cell dofs -> dofs of child k (mg.py:57-73).
    Fine corner q of child k sits at CORNERS[k] + CORNERS[q] in fine units
    (0..2); its trilinear weight on coarse corner r is prod_d (t or 1 - t)."""
    p = np.zeros((8, 24, 24))
    for k in range(8):
        for q in range(8):
            pos = (CORNERS[k] + CORNERS[q]) / 2.0
            for r in range(8):
                w = np.prod(np.where(CORNERS[r] == 1, pos, 1.0 - pos))
                if w != 0.0:
                    for i in range(3):
                        p[k, 3 * q + i, 3 * r + i] = w
    return p


PK = child_prolongations()


def prolong_rows(fine_ijk: np.ndarray, coarse_node_shape) -> sp.csr_matrix:
    """Scalar trilinear prolongation rows for fine nodes (mg.py:390-408):
    closed form used to pin the Galerkin levels."""
    n = len(fine_ijk)
    rows, cols, vals = [], [], []
    for corner in range(8):
        w = np.ones(n)
        ci = np.empty((n, 3), np.int64)
        for d in range(3):
            i = fine_ijk[:, d]
            odd = (i % 2) == 1
            hi = (corner >> d) & 1
            ci[:, d] = i // 2 + (hi & odd)
            w *= np.where(odd, 0.5, 0.0 if hi else 1.0)
        ok = w > 0
        rows.append(np.nonzero(ok)[0])
        cols.append(np.ravel_multi_index(tuple(ci[ok].T), coarse_node_shape))
        vals.append(w[ok])
    nc = int(np.prod(coarse_node_shape))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, nc))


# --------------------------------------------------------------------------
# levels
# --------------------------------------------------------------------------

def _gather(u, x0, x1, ny, nz):
    """(24, x1-x0, ny, nz) corner values of the cells in x-slab [x0, x1)."""
    out = np.empty((24, x1 - x0, ny, nz), u.dtype)
    for c, (dx, dy, dz) in enumerate(CORNER_OFF):
        out[3 * c:3 * c + 3] = u[:, x0 + dx:x1 + dx, dy:dy + ny, dz:dz + nz]
    return out


def _scatter_add(f, x0, x1, ny, nz, out):
    for c, (dx, dy, dz) in enumerate(CORNER_OFF):
        out[:, x0 + dx:x1 + dx, dy:dy + ny, dz:dz + nz] += f[3 * c:3 * c + 3]


class Level:
    """One grid level with cells (nx, ny, nz) and node grid
    (3, nx+1, ny+1, nz+1). kind "fine": cell matrix e[c]*ke; kind
    "explicit": k (24, 24, nx, ny, nz). Operator m*K(m*u) + (1-m)*u."""

    SLAB_BYTES = 64e6

    def __init__(self, kind, dims, dtype):
        self.kind = kind
        self.dims = tuple(int(d) for d in dims)
        self.dtype = np.dtype(dtype)
        self.ke = None
        self.e = None
        self.k = None
        self.mask = None
        self.dinv = None
        self.lmax = None

    @property
    def node_shape(self):
        nx, ny, nz = self.dims
        return (3, nx + 1, ny + 1, nz + 1)

    def astype(self, dtype):
        lv = Level(self.kind, self.dims, dtype)
        if self.kind == "fine":
            lv.ke = self.ke.astype(dtype)
            lv.e = self.e.astype(dtype)
        else:
            lv.k = self.k.astype(dtype)
        lv.mask = self.mask.astype(dtype)
        lv.dinv = self.dinv.astype(dtype)
        lv.lmax = self.lmax
        return lv

    def _step(self):
        nx, ny, nz = self.dims
        per = 24 * 24 * ny * nz * self.dtype.itemsize
        return max(1, int(self.SLAB_BYTES // max(per, 1)))

    def apply_k(self, u):
        """Unmasked K u."""
        nx, ny, nz = self.dims
        out = np.zeros(self.node_shape, u.dtype)
        step = self._step()
        for a in range(0, nx, step):
            b = min(a + step, nx)
            g = _gather(u, a, b, ny, nz).reshape(24, -1)
            if self.kind == "fine":
                f = (self.ke.astype(u.dtype) @ g) * self.e[a:b].reshape(1, -1).astype(u.dtype)
            else:
                kk = self.k[:, :, a:b].reshape(24, 24, -1).astype(u.dtype)
                f = np.einsum("ijn,jn->in", kk, g)
            _scatter_add(f.reshape(24, b - a, ny, nz), a, b, ny, nz, out)
        return out

    def apply(self, u):
        m = self.mask.astype(u.dtype)
        out = self.apply_k(u * m)
        return out * m + (1 - m) * u

    def diagonal(self):
        nx, ny, nz = self.dims
        d = np.zeros(self.node_shape)
        if self.kind == "fine":
            dc = np.diag(self.ke).astype(np.float64)[:, None, None, None] * self.e[None].astype(np.float64)
        else:
            dc = np.einsum("iixyz->ixyz", self.k).astype(np.float64)
        for c, (dx, dy, dz) in enumerate(CORNER_OFF):
            d[:, dx:dx + nx, dy:dy + ny, dz:dz + nz] += dc[3 * c:3 * c + 3]
        return d

    def cell_matrices(self, a, b):
        """(24, 24, b-a, ny, nz) float64 of x-slab [a, b)."""
        if self.kind == "fine":
            return self.ke.astype(np.float64)[:, :, None, None, None] * self.e[a:b][None, None].astype(np.float64)
        return self.k[:, :, a:b].astype(np.float64)

    def active_cells(self) -> int:
        if self.kind == "fine":
            return int((self.e != 0).sum())
        return int((np.abs(self.k).max(axis=(0, 1)) > 0).sum())


def corner_mask(mask, a, b, ny, nz):
    """(24, b-a, ny, nz) mask values at the cells' corner dofs."""
    dm = np.empty((24, b - a, ny, nz))
    for q, (qx, qy, qz) in enumerate(CORNER_OFF):
        dm[3 * q:3 * q + 3] = mask[:, a + qx:b + qx, qy:qy + ny, qz:qz + nz]
    return dm


def galerkin(fine: Level, masked: bool) -> Level:
    """Coarse cells (half, padded even) with K_c = sum_k P_k^T K_child P_k,
    K_child masked by the fine mask when `masked` (mg.py:601-677). Processed
    by coarse x-slabs to bound memory."""
    nx, ny, nz = fine.dims
    half = (nx // 2, ny // 2, nz // 2)
    cdims = tuple(h + h % 2 if h > 1 else h for h in half)
    k = np.zeros((24, 24) + cdims)
    hx, hy, hz = half
    per = 24 * 24 * 8 * ny * nz * 8
    step = max(1, int(256e6 // max(per, 1)))
    for c0 in range(0, hx, step):
        c1 = min(c0 + step, hx)
        kf = fine.cell_matrices(2 * c0, 2 * c1)
        if masked:
            dm = corner_mask(fine.mask, 2 * c0, 2 * c1, ny, nz)
            kf = kf * dm[:, None] * dm[None, :]
        for ch, (ox, oy, oz) in enumerate(CORNER_OFF):
            sub = kf[:, :, ox::2, oy:2 * hy:2, oz:2 * hz:2]
            t = np.einsum("ai,abxyz,bj->ijxyz", PK[ch], sub, PK[ch], optimize=True)
            k[:, :, c0:c1, :hy, :hz] += t
    lv = Level("explicit", cdims, np.float64)
    lv.k = k
    return lv


class GridOracle:
    """FP64 restatement of mg.GridElasticity (mg.py:411-988).

    smoother: "chebyshev" (the reference's, degree `degree`) or "jacobi"
    (damped Jacobi, `omega`, `degree` sweeps pre and post). vdtype: the
    V-cycle's precision (the reference uses float32)."""

    def __init__(self, mesh, nu, fixed, rtol=1e-6, maxiter=500, coarse_cells=2000, degree=2,
                 vdtype=np.float32, element="im", smoother="chebyshev", omega=0.6, ke=None):
        self.mesh = mesh
        self.rtol = rtol
        self.maxiter = maxiter
        self.coarse_cells = int(coarse_cells)
        self.degree = int(degree)
        self.vdtype = np.dtype(vdtype)
        self.smoother = smoother
        self.omega = float(omega)
        cells = mesh.elem_ijk.max(axis=0) + 1
        cells = cells + cells % 2
        self.cells = tuple(int(c) for c in cells)
        self.node_shape = tuple(int(c) + 1 for c in cells)
        self.nn = int(np.prod(self.node_shape))
        self.node_flat = np.ravel_multi_index(tuple(mesh.node_ijk.T), self.node_shape)
        self.ke = (hex8_ke(nu, element) if ke is None else np.asarray(ke, float)) * mesh.pitch
        fixed = np.asarray(fixed, bool)
        fg = np.zeros((3, self.nn), bool)
        for i in range(3):
            fg[i, self.node_flat] = fixed[i::3]
        self.fixed_grid = fg.reshape((3,) + self.node_shape)
        self.levels = []
        self.vlevels = []
        self.strict = False
        self.last_iters = 0
        self.coarse = None
        self._lmax_vec = {}

    def to_grid(self, v, dtype=np.float64):
        out = np.zeros((3, self.nn), dtype)
        for i in range(3):
            out[i, self.node_flat] = v[i::3]
        return out.reshape((3,) + self.node_shape)

    def from_grid(self, g):
        flat = g.reshape(3, -1)
        out = np.empty(3 * len(self.node_flat))
        for i in range(3):
            out[i::3] = flat[i, self.node_flat]
        return out

    def cell_grid(self, e, dtype=np.float64):
        g = np.zeros(self.cells, dtype)
        g[tuple(self.mesh.elem_ijk.T)] = e
        return g

    # ---- hierarchy (mg.py:515-591) -----------------------------------------

    def set_moduli(self, e):
        e = np.asarray(e, float)
        self._e = e
        l0 = Level("fine", self.cells, np.float64)
        l0.ke = self.ke.copy()
        l0.e = self.cell_grid(e)
        d0 = l0.diagonal()
        alive = d0 > 1e-300
        l0.mask = (alive & ~self.fixed_grid).astype(np.float64)
        l0.dinv = np.where(l0.mask > 0, 1.0 / np.where(alive, d0, 1.0), 1.0)
        self.alive = alive
        self.levels = [l0]
        fine = l0
        n_active = int((e > 0).sum())
        held = alive & self.fixed_grid
        first = True
        while n_active > self.coarse_cells and min(fine.dims) >= 2:
            coarse = galerkin(fine, masked=first or self.strict)
            first = False
            d = coarse.diagonal()
            dmax = float(d.max()) if d.size else 0.0
            live = d > 1e-12 * max(dmax, 1e-300)
            if self.strict:
                hc = np.zeros((3,) + coarse.node_shape[1:], bool)
                inj = held[:, ::2, ::2, ::2]
                hc[:, :inj.shape[1], :inj.shape[2], :inj.shape[3]] = inj[:, :hc.shape[1], :hc.shape[2], :hc.shape[3]]
                held = hc
                kept = np.zeros((3,) + coarse.node_shape[1:])
                r = restrict(np.asarray(fine.mask, np.float64))
                kept[:, :r.shape[1], :r.shape[2], :r.shape[3]] = r[:, :kept.shape[1], :kept.shape[2], :kept.shape[3]]
                live &= ~held & (kept >= 1.0)
            coarse.mask = live.astype(np.float64)
            coarse.dinv = np.where(live, 1.0 / np.where(live, d, 1.0), 1.0)
            n_active = coarse.active_cells()
            self.levels.append(coarse)
            fine = coarse
        try:
            self._build_coarse()
        except np.linalg.LinAlgError:
            if self.strict:
                raise RuntimeError("coarsest factor is singular")
            self.strict = True
            return self.set_moduli(e)
        self.vlevels = [lv.astype(self.vdtype) for lv in self.levels]
        if self.smoother == "chebyshev":
            for i in range(len(self.vlevels)):
                self.vlevels[i].lmax = self._estimate_lmax(i)

    def _build_coarse(self):
        if len(self.levels) < 2:
            self.coarse = None
            return
        lv = self.levels[-1]
        a = coarse_matrix(lv)
        self.coarse = np.linalg.cholesky(a)            # raises LinAlgError when not SPD

    def _coarse_solve(self, b):
        import scipy.linalg as sla

        rhs = b.reshape(-1).astype(np.float64)
        y = sla.solve_triangular(self.coarse, rhs, lower=True)
        x = sla.solve_triangular(self.coarse.T, y, lower=False)
        return x.reshape(b.shape).astype(b.dtype)

    def _estimate_lmax(self, i, its=12):
        lv = self.vlevels[i]
        prev = self._lmax_vec.get(i)
        if prev is not None and prev.shape == lv.node_shape:
            v = prev.astype(lv.dtype) * lv.mask
        else:
            rng = np.random.default_rng(i)
            v = (rng.random(lv.node_shape) - 0.5).astype(lv.dtype) * lv.mask
        lam = 1.0
        for _ in range(its):
            nv = float(np.vdot(v, v)) ** 0.5
            if nv == 0:
                return 2.0
            v = v / lv.dtype.type(nv)
            av = lv.dinv * lv.apply(v) * lv.mask
            lam = float(np.vdot(v, av))
            v = av
        self._lmax_vec[i] = v
        return LMAX_SAFETY * max(lam, 1e-12)

    # ---- smoothers and V-cycle (mg.py:796-870) -----------------------------

    def _chebyshev(self, lv, b, x):
        hi = lv.lmax
        lo = hi / 30.0
        theta = 0.5 * (hi + lo)
        delta = 0.5 * (hi - lo)
        sigma = theta / delta
        rho = 1.0 / sigma
        t = lv.dtype.type
        if x is None:
            x = np.zeros_like(b)
            r = b.copy()
            d = (lv.dinv * b) * t(1.0 / theta)
        else:
            r = b - lv.apply(x)
            d = t(1.0 / theta) * lv.dinv * r
        for it in range(self.degree):
            last = it + 1 == self.degree
            ad = None if last else lv.apply(d)
            rho_new = 1.0 / (2.0 * sigma - rho)
            x += d
            if not last:
                r -= ad
                d = t(rho_new * rho) * d + t(2.0 * rho_new / delta) * lv.dinv * r
            rho = rho_new
        return x

    def _jacobi(self, lv, b, x):
        w = lv.dtype.type(self.omega)
        sweeps = self.degree
        if x is None:
            x = w * lv.dinv * b
            sweeps -= 1
        for _ in range(sweeps):
            x = x + w * lv.dinv * (b - lv.apply(x))
        return x

    def _smooth(self, lv, b, x):
        if self.smoother == "jacobi":
            return self._jacobi(lv, b, x)
        return self._chebyshev(lv, b, x)

    def _vcycle(self, i, b):
        lv = self.vlevels[i]
        if i == len(self.vlevels) - 1:
            if self.coarse is not None:
                return self._coarse_solve(b)
            x = self._smooth(lv, b, None)
            return self._smooth(lv, b, x)
        x = self._smooth(lv, b, None)
        ax = lv.apply(x)
        nxt = self.vlevels[i + 1]
        cs = nxt.node_shape
        sx, sy, sz = lv.node_shape[1:]
        rf = np.zeros((3,) + tuple(2 * n - 1 for n in cs[1:]), b.dtype)
        rf[:, :sx, :sy, :sz] = (b - ax) * lv.mask
        rc = restrict(rf) * nxt.mask
        xc = self._vcycle(i + 1, rc)
        x += prolong(xc * nxt.mask)[:, :sx, :sy, :sz] * lv.mask
        return self._smooth(lv, b, x)

    def precondition(self, r):
        return self._vcycle(0, r.astype(self.vdtype)).astype(np.float64)

    def apply(self, u):
        return self.levels[0].apply(u)

    # ---- PCG (mg.py:872-926) -----------------------------------------------

    def _pcg(self, b, x0):
        a = self.levels[0]
        x = np.zeros(a.node_shape) if x0 is None else x0.copy()
        bnorm = float(np.vdot(b, b)) ** 0.5
        if bnorm == 0:
            return np.zeros_like(b), 0, True
        r = b.copy()
        if x0 is not None:
            ax = a.apply(x)
            xax = float(np.vdot(x, ax))
            s = float(np.vdot(b, x)) / xax if xax > 0 else 0.0
            if s > 0:
                rs = b - s * ax
                if float(np.vdot(rs, rs)) ** 0.5 < bnorm:
                    x *= s
                    r = rs
                else:
                    x[:] = 0
            else:
                x[:] = 0
        z = self.precondition(r)
        p = z.copy()
        rz = float(np.vdot(r, z))
        tol = self.rtol * bnorm
        for it in range(1, self.maxiter + 1):
            ap = a.apply(p)
            pap = float(np.vdot(p, ap))
            if pap <= 0:
                return x, it, False
            alpha = rz / pap
            x += alpha * p
            r -= alpha * ap
            if float(np.vdot(r, r)) ** 0.5 <= tol:
                return x, it, True
            z = self.precondition(r)
            rz_new = float(np.vdot(r, z))
            p = z + (rz_new / rz) * p
            rz = rz_new
        return x, self.maxiter, False

    def _solve(self, b, g0):
        if self.strict or len(self.levels) < 2:
            return self._pcg(b, g0)
        full = self.maxiter
        self.maxiter = min(full, STALL)
        try:
            x, it, ok = self._pcg(b, g0)
        finally:
            self.maxiter = full
        if ok:
            return x, it, ok
        self.strict = True
        self.set_moduli(self._e)
        x, it2, ok = self._pcg(b, g0)
        return x, it + it2, ok

    def solve_many(self, fs, x0s=None):
        x0s = x0s or [None] * len(fs)
        out = []
        iters = 0
        for f, x0 in zip(fs, x0s):
            b = self.to_grid(f) * self.levels[0].mask
            g0 = self.to_grid(x0) * self.levels[0].mask if x0 is not None else None
            x, it, ok = self._solve(b, g0)
            iters = max(iters, it)
            if not ok:
                raise RuntimeError(f"linear solve did not converge ({it} iterations); "
                                   "the structure may be a mechanism (check supports)")
            out.append(self.from_grid(x * self.levels[0].mask))
        self.last_iters = iters
        return out

    def live_dofs(self):
        a = self.alive.reshape(3, -1)
        out = np.empty(3 * len(self.node_flat), bool)
        for i in range(3):
            out[i::3] = a[i, self.node_flat]
        return out

    def strain_energy(self, u):
        g = self.to_grid(u)
        return grid_strain_energy(g, self.ke, self.cells)[tuple(self.mesh.elem_ijk.T)]


def grid_strain_energy(g, ke, cells):
    """(nx, ny, nz) u_c^T Ke u_c of node-grid field g (mg.py:975-988)."""
    nx, ny, nz = cells
    se = np.zeros(cells)
    step = max(1, int(32e6 // (24 * ny * nz * 8)))
    for a in range(0, nx, step):
        b = min(a + step, nx)
        flat = _gather(g, a, b, ny, nz).reshape(24, -1)
        se[a:b] = np.einsum("in,in->n", flat, ke @ flat).reshape(b - a, ny, nz)
    return se


def coarse_matrix(lv: Level) -> np.ndarray:
    """Dense D K D + (I - D) of a level over all its node-grid dofs in
    component-major order (the reference assembles the active cells' dofs
    only, mg.py:708-741; dofs outside them are identity rows here, masked
    out of every right-hand side, so the solution agrees)."""
    nx, ny, nz = lv.dims
    ns = lv.node_shape[1:]
    nn = int(np.prod(ns))
    kall = lv.cell_matrices(0, nx)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = np.ravel_multi_index((ii.ravel(), jj.ravel(), kk.ravel()), ns)
    idx = np.empty((24, base.size), np.int64)
    for c, (dx, dy, dz) in enumerate(CORNER_OFF):
        node = base + (dx * ns[1] + dy) * ns[2] + dz
        for i in range(3):
            idx[3 * c + i] = i * nn + node
    kc = kall.reshape(24, 24, -1)
    rows = np.broadcast_to(idx[:, None, :], kc.shape).ravel()
    cols = np.broadcast_to(idx[None, :, :], kc.shape).ravel()
    n = 3 * nn
    k = sp.csr_matrix((kc.ravel(), (rows, cols)), shape=(n, n)).toarray()
    m = lv.mask.reshape(-1).astype(np.float64)
    a = m[:, None] * k * m[None, :] + np.diag(1.0 - m)
    return 0.5 * (a + a.T)
