"""Oracle: the p-mean von Mises stress aggregate and its adjoint sensitivity
(FP64, numpy).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows loadpath/constrained.py, the non-averaged stress branch:
- sample points: the centre (fem.hex8_b_center, fem.py:307-308) or the 8
  corners, x outermost, with the condensed bubbles (constrained.py:384-387;
  fem.hex8_b, fem.py:227-237); sigma = C B u_e / pitch
- vm = sqrt(max(s^T V s, 1e-30)), V of constrained.py:107-114
- s_rel = x^q vm / lim; pm = sc (sum (s_rel / sc)^P / nc)^(1/P), sc = max
  (s_max, 1e-30); g = kappa pm                          constrained.py:622-641
- dg/ds = kappa (s_rel / pm)^(P-1) / nc; explicit part dg/dx at fixed u;
  adjoint right-hand side sum dg/ds x^q / (vm lim) (V sigma)^T C B, zero on
  fixed dofs                                            constrained.py:643-653
- dg/dx -= dE/dx lambda_e^T Ke u_e, K lambda = rhs      constrained.py:655-659

Pinned by (tests/test_oracle_stress.py):
- centre stress of an affine displacement field equals C eps exactly (the
  B matrix reproduces any constant strain), and vm of a uniaxial stress is
  |s|, of pure shear sqrt(3)|t| (closed forms);
- pm -> s_max as P grows, pm = s_max when every cell reads the same stress;
- the whole gradient (explicit + adjoint through a direct sparse solve)
  against FP64 central finite differences of the forward functional, on
  seeded elements of a small clamped bar, centre and corner sampling;
- the compliance gradient -p x^(p-1) (1-Emin) u_e^T Ke u_e against the same
  finite differences (the self-adjoint case).

Also verify.nodal_stress / nodal_von_mises (verify.py:197-206, 307-337):
each cell's stress at its own 8 corners (fem.CORNERS order, bit0 = x),
modulus * C (B u_e / pitch, thermal taken off the normal strains), summed
at the corner's node and divided by the number of cells there (the cells
of a stress mask only, when one is given); fem.von_mises (fem.py:311-314).
Pinned (tests/test_oracle_stress.py) by the reference's own closed forms
(tests/test_hex8_im.py:27-84): the constant-strain patch test returns
C eps at every node, and a one-cell-deep strip in pure bending reads the
top-fibre stress M c / I with the "im" element; a free thermal expansion
reads zero stress; and a per-cell loop over corners (no np.add.at).

Also the material-averaged stress reading of the constrained loop
(constrained._Averaged, constrained.py:146-252, used at 517-605; no thin
walls): samples at each cell's quarter points (natural +-0.5, fem.CORNERS
order) while radius <= 1.5 cells, else its centre, sigma_s = C B(s) u_e /
pitch on the cells that moved (any nonzero displacement at a corner) and 0
elsewhere; weights v_s = pitch^3 / ns * mat; per readable node the ball of
samples within radius: D = sum v, sig = sum v sigma / D, rho = sum v mat / D;
vm (and the layer bond) from sig, s_rel = rho^q f / lim, pm as above over
the readable nodes (twice as many terms with a bond), kappa normalized;
the adjoint right-hand side through the samples of moved cells and the
explicit dg/dx through the weights. Pinned by the rhs and dg/dx against
central differences of pm (directional, random directions), with and
without a bond, and by the closed form of a uniform stress (every node
reads C eps, pm = vm / lim).
"""
from __future__ import annotations

import numpy as np

from .fem import b_matrix, hex8_b, isotropic_c, von_mises_matrix

VM = von_mises_matrix()
P_DEFAULT = 8.0     # constrained._P (constrained.py:46)
Q_DEFAULT = 1.0     # constrained._Q (constrained.py:45)


def sample_cb(nu: float, points: str = "center", element: str = "im", pitch: float = 1.0) -> np.ndarray:
    """(npts, 6, 24) C B / pitch at the stress sample points
    (constrained.py:383-387)."""
    c0 = isotropic_c(1.0, nu)
    if points == "center":
        return (c0 @ b_matrix(0.0, 0.0, 0.0) / pitch)[None]
    if points == "corners":
        return np.array([c0 @ hex8_b(a, b, c, nu, element) / pitch
                         for a in (-1.0, 1.0) for b in (-1.0, 1.0) for c in (-1.0, 1.0)])
    raise ValueError(points)


def pnorm_forward(ue, x, cbs, p=P_DEFAULT, q=Q_DEFAULT, lim=1.0, kappa=1.0, nc=None, pmask=None):
    """ue (ne, 24) element displacements, x (ne,) material share, pmask
    (ne, npts) the sample points read (None: all; a point not read has zero
    stress, constrained.py:622-625). Returns (g, pm, smax, parts) with parts
    per sample point (cb, sv, vm, s_rel) (constrained.py:618-641)."""
    ue = np.asarray(ue, float)
    x = np.asarray(x, float)
    if nc is None:
        nc = len(x) * len(cbs)
    parts = []
    for k, cb in enumerate(cbs):
        sig = ue @ cb.T
        if pmask is not None:
            sig = sig * np.asarray(pmask)[:, k][:, None]
        sv = sig @ VM
        vm = np.sqrt(np.maximum(np.einsum("ij,ij->i", sv, sig), 1e-30))
        parts.append((cb, sv, vm, x ** q * vm / lim))
    smax = max((float(sr.max()) for *_, sr in parts if len(sr)), default=0.0)
    sc = max(smax, 1e-30)
    pm = sc * (sum(float(np.sum((sr / sc) ** p)) for *_, sr in parts) / nc) ** (1.0 / p)
    return kappa * pm, pm, smax, parts


def pnorm_partials(ue, x, cbs, p=P_DEFAULT, q=Q_DEFAULT, lim=1.0, kappa=1.0, nc=None, pmask=None):
    """(g, pm, smax, g_explicit (ne,), rhs_e (ne, 24)): the explicit
    derivative in x at fixed u and the element adjoint right-hand sides
    dg/du_e (constrained.py:643-652, dmat = 1)."""
    x = np.asarray(x, float)
    if nc is None:
        nc = len(x) * len(cbs)
    g, pm, smax, parts = pnorm_forward(ue, x, cbs, p, q, lim, kappa, nc, pmask)
    gx = np.zeros(len(x))
    rhs_e = np.zeros((len(x), 24))
    for cb, sv, vm, sr in parts:
        dg_ds = kappa * (sr / max(pm, 1e-30)) ** (p - 1) / nc
        gx += dg_ds * (q * np.maximum(x, 1e-12) ** (q - 1) * vm / lim)
        coef = dg_ds * (x ** q / (vm * lim))
        rhs_e += coef[:, None] * (sv @ cb)
    return g, pm, smax, gx, rhs_e


def assemble_rhs(edof, rhs_e, n_dof, fixed=None):
    rhs = np.bincount(np.asarray(edof).ravel(), weights=np.asarray(rhs_e).ravel(), minlength=n_dof)
    if fixed is not None:
        rhs[np.asarray(fixed, bool)] = 0.0
    return rhs


def element_product(a, b, edof, ke):
    """a_e^T Ke b_e per element (constrained.py:659; the bilinear form of
    femops.element_energy)."""
    return np.einsum("ij,jk,ik->i", a[edof], ke, b[edof])


def pnorm_gradient(mesh, ke, fixed, f, xphys, cbs, penal=3.0, emin=1e-9, p=P_DEFAULT, q=Q_DEFAULT, lim=1.0,
                   kappa=1.0):
    """The whole objective and its gradient in xphys through a direct FP64
    solve: (g, dg/dx, u, lambda). E = emin + x^penal (1 - emin); the stress
    reads the material share x itself (constrained.py:423-428, share None)."""
    import scipy.sparse.linalg as spla

    from .fem import assemble

    x = np.asarray(xphys, float)
    e = emin + x ** penal * (1 - emin)
    de = penal * x ** (penal - 1) * (1 - emin)
    k = assemble(mesh, ke, e)
    free = ~np.asarray(fixed, bool)
    kf = k[free][:, free].tocsc()
    lu = spla.splu(kf)
    u = np.zeros(mesh.n_dof)
    u[free] = lu.solve(f[free])
    edof = mesh.edof
    g, pm, smax, gx, rhs_e = pnorm_partials(u[edof], x, cbs, p, q, lim, kappa)
    rhs = assemble_rhs(edof, rhs_e, mesh.n_dof, fixed)
    lam = np.zeros(mesh.n_dof)
    lam[free] = lu.solve(rhs[free])
    grad = gx - de * element_product(lam, u, edof, ke)
    return g, grad, u, lam


def pnorm_value(mesh, ke, fixed, f, xphys, cbs, penal=3.0, emin=1e-9, p=P_DEFAULT, q=Q_DEFAULT, lim=1.0,
                kappa=1.0):
    import scipy.sparse.linalg as spla

    from .fem import assemble

    x = np.asarray(xphys, float)
    k = assemble(mesh, ke, emin + x ** penal * (1 - emin))
    free = ~np.asarray(fixed, bool)
    u = np.zeros(mesh.n_dof)
    u[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    return pnorm_forward(u[mesh.edof], x, cbs, p, q, lim, kappa)[0]


def nodal_stress(mesh, u, modulus, nu: float, element: str = "im", thermal=None, mask=None):
    """verify.nodal_stress (verify.py:307-337) on a HexMesh: (n_nodes, 6)."""
    c = isotropic_c(1.0, nu)
    mod = np.broadcast_to(np.asarray(modulus, float), (mesh.n_elem,))[:, None]
    if mask is not None:
        mod = mod * np.asarray(mask, float)[:, None]
    ue = np.asarray(u, float)[mesh.edof]
    acc = np.zeros((mesh.n_nodes, 6))
    for k in range(8):
        a, b, cc = 2.0 * (k & 1) - 1.0, 2.0 * ((k >> 1) & 1) - 1.0, 2.0 * ((k >> 2) & 1) - 1.0
        strain = ue @ (hex8_b(a, b, cc, nu, element) / mesh.pitch).T
        if thermal is not None:
            strain[:, :3] -= np.asarray(thermal, float)[:, None]
        np.add.at(acc, mesh.enode[:, k], (strain @ c.T) * mod)
    if mask is not None:
        cnt = np.bincount(mesh.enode[np.asarray(mask, bool)].ravel(), minlength=mesh.n_nodes)
        return acc / np.maximum(cnt, 1)[:, None]
    cnt = np.bincount(mesh.enode.ravel(), minlength=mesh.n_nodes)
    return acc / cnt[:, None]


def von_mises(stress):
    """fem.von_mises (fem.py:311-314)."""
    sx, sy, sz, txy, tyz, tzx = np.asarray(stress, float).T
    return np.sqrt(0.5 * ((sx - sy) ** 2 + (sy - sz) ** 2 + (sz - sx) ** 2) + 3.0 * (txy ** 2 + tyz ** 2 + tzx ** 2))


def material_average(node_xyz, enode, fill, pitch: float, radius: float, sig):
    """verify._MaterialAverage (verify.py:214-305) applied to nodal tensors
    sig (n_nodes, 6): the nodal tensors interpolated trilinearly to each
    live cell's samples (its 8 quarter points while radius <= 1.5 cells,
    else its centre), each sample weighted by the cell's fill * volume /
    samples; past 3.3 cells the samples are merged by blocks of blk^3 cells
    (blk = ceil(radius / 3.3 cells), blocks counted from the smallest sample
    coordinate) into one sample at the block's weighted centroid carrying
    the weighted mean tensor and the summed weight (verify.py:264-274); then
    every node's weighted mean over the samples within radius (its own
    value when there are none)."""
    from scipy.spatial import cKDTree

    node_xyz = np.asarray(node_xyz, float)
    enode = np.asarray(enode)
    corners = np.array([[k & 1, (k >> 1) & 1, (k >> 2) & 1] for k in range(8)], float)
    n, ne = len(node_xyz), len(enode)
    f = np.ones(ne) if fill is None else np.asarray(fill, float)
    live = np.nonzero(f > 0)[0]
    rho = radius / pitch
    q = 0.25 + 0.5 * corners if rho <= 1.5 else np.full((1, 3), 0.5)
    ns = len(q)
    wts = np.prod(np.where(corners[None, :, :] > 0.5, q[:, None, :], 1.0 - q[:, None, :]), axis=2)
    xs = np.einsum("sk,ekd->esd", wts, node_xyz[enode[live]]).reshape(-1, 3)
    w = np.repeat(f[live] * pitch ** 3 / ns, ns)
    sig = np.asarray(sig, float)
    ssig = np.einsum("sk,ekc->esc", wts, sig[enode[live]]).reshape(-1, 6)
    blk = int(np.ceil(rho / 3.3))
    if blk > 1 and len(xs):
        key = np.floor((xs - xs.min(axis=0)) / (blk * pitch) + 1e-9).astype(np.int64)
        _, gid = np.unique(key, axis=0, return_inverse=True)
        gid = np.asarray(gid).ravel()
        nb = int(gid.max()) + 1
        wb = np.bincount(gid, w, nb)
        xs = np.stack([np.bincount(gid, w * xs[:, d], nb) for d in range(3)], axis=1) / wb[:, None]
        ssig = np.stack([np.bincount(gid, w * ssig[:, c], nb) for c in range(6)], axis=1) / wb[:, None]
        w = wb
    pairs = cKDTree(node_xyz).sparse_distance_matrix(cKDTree(xs), radius, output_type="coo_matrix")
    out = np.zeros((n, 6))
    tot = np.zeros(n)
    np.add.at(out, pairs.row, w[pairs.col, None] * ssig[pairs.col])
    np.add.at(tot, pairs.row, w[pairs.col])
    lone = tot <= 0
    out = out / np.where(lone, 1.0, tot)[:, None]
    out[lone] = sig[lone]
    return out


def bond_form(normal):
    """constrained._bond_form (constrained.py:262-273): (m, Q) with the
    stress normal to layers stacked along `normal` m . sigma and the squared
    shear on them sigma . Q . sigma (Voigt: sx, sy, sz, txy, tyz, tzx)."""
    n = np.asarray(normal, float)
    n = n / np.linalg.norm(n)
    t = np.array([[n[0], 0, 0, n[1], 0, n[2]], [0, n[1], 0, n[0], n[2], 0], [0, 0, n[2], 0, n[1], n[0]]])
    m = n @ t
    return m, t.T @ t - np.outer(m, m)


def normalize_kappa(prev, smax, pm, it):
    """constrained._normalize (constrained.py:255-259)."""
    k = smax / max(pm, 1e-30)
    return 0.5 * (prev + k) if it > 1 else k


def averaged_samples(nu, rho, element="im"):
    """(natural coordinates (ns, 3), C B at them (ns, 6, 24) at unit pitch)
    (constrained.py:176-183)."""
    corners = np.array([[k & 1, (k >> 1) & 1, (k >> 2) & 1] for k in range(8)], float)
    nat = corners - 0.5 if rho <= 1.5 else np.zeros((1, 3))
    c0 = isotropic_c(1.0, nu)
    return nat, np.array([c0 @ hex8_b(a, b, c, nu, element) for a, b, c in nat])


def averaged_pnorm(mesh, nu, radius, readable, us, xe, *, p=P_DEFAULT, q=Q_DEFAULT, limit=1.0, bond=None,
                   kappa_prev=None, it=1, element="im", fixed=None):
    """The averaged stress branch of constrained._constrained_run
    (constrained.py:517-605) over mesh (a Mesh), for displacements us (one
    per load case, n_dof each), eroded densities xe, readable nodes (bool
    per node) and bond = (ratio, normal, k) or None. Per case: dict(g, pm,
    smax, kappa, rhs (n_dof; zero on `fixed`), gx (explicit dg/dxe))."""
    import scipy.sparse as sp
    from scipy.spatial import cKDTree

    pitch = mesh.pitch
    nat, cb = averaged_samples(nu, radius / pitch, element)
    cb = cb / pitch
    ns = len(nat)
    ne = mesh.n_elem
    xs = ((mesh.elem_ijk[:, None, :] + 0.5 + 0.5 * nat[None]) * pitch).reshape(-1, 3)
    nodes = np.flatnonzero(readable)
    pr = cKDTree(mesh.node_ijk[nodes] * pitch).sparse_distance_matrix(cKDTree(xs), radius, output_type="coo_matrix")
    ball = sp.csr_matrix((np.ones(len(pr.row)), (pr.row, pr.col)), shape=(len(nodes), len(xs)))
    ballT = ball.T.tocsr()
    w0 = pitch ** 3 / ns
    mat = np.maximum(np.asarray(xe, float), 0.0)
    v = np.repeat(w0 * mat, ns)
    xd = np.repeat(mat, ns)
    dn = ball @ v
    dinv = np.where(dn > 0, 1.0 / np.where(dn > 0, dn, 1.0), 0.0)
    rho_n = (ball @ (v * xd)) * dinv
    nc = max(len(nodes), 1) * (2 if bond is not None else 1)
    if bond is not None:
        bm, bq = bond_form(bond[1])
        k2 = float(bond[2]) ** 2
    edof = mesh.edof
    out = []
    for ci, u in enumerate(us):
        u = np.asarray(u, float)
        moved = np.flatnonzero(u.reshape(-1, 3).any(axis=1)[mesh.enode].any(axis=1))
        ue = u[edof[moved]]
        ss = np.zeros((ne, ns, 6))
        for k in range(ns):
            ss[moved, k] = ue @ cb[k].T
        ss = ss.reshape(-1, 6)
        sig = (ball @ (v[:, None] * ss)) * dinv[:, None]
        sv = sig @ VM
        parts = [(sv, np.sqrt(np.maximum(np.einsum("ij,ij->i", sv, sig), 1e-30)), limit)]
        if bond is not None:
            sn = np.maximum(sig @ bm, 0.0)
            sb = sn[:, None] * bm + k2 * (sig @ bq)
            parts.append((sb, np.sqrt(np.maximum(np.einsum("ij,ij->i", sb, sig), 1e-30)), limit * float(bond[0])))
        srs = [rho_n ** q * f / lim for _, f, lim in parts]
        smax = max((float(sr.max()) for sr in srs if len(sr)), default=0.0)
        sc = max(smax, 1e-30)
        pm = sc * (sum(float(np.sum((sr / sc) ** p)) for sr in srs) / nc) ** (1.0 / p)
        prev = 1.0 if kappa_prev is None else float(np.atleast_1d(kappa_prev)[ci])
        kappa = normalize_kappa(prev, smax, pm, it)
        gsig = np.zeros_like(sig)
        grho = np.zeros(len(sig))
        for (sh, f, lim), sr in zip(parts, srs):
            a = kappa * (sr / max(pm, 1e-30)) ** (p - 1) / nc
            gsig += (a * rho_n ** q / (f * lim))[:, None] * sh
            grho += a * q * np.maximum(rho_n, 1e-12) ** (q - 1) * f / lim
        hs = ballT @ (gsig * dinv[:, None])
        hw = v[:, None] * hs
        t_sig = np.einsum("ij,ij->i", ss, hs) - ballT @ (np.einsum("ij,ij->i", gsig, sig) * dinv)
        t_rho = 2.0 * xd * (ballT @ (grho * dinv)) - ballT @ (grho * rho_n * dinv)
        gx = w0 * (t_sig + t_rho).reshape(ne, ns).sum(axis=1)
        hw = hw.reshape(ne, ns, 6)
        rhs = np.zeros(mesh.n_dof)
        for k in range(ns):
            np.add.at(rhs, edof[moved].ravel(), (hw[moved, k] @ cb[k]).ravel())
        if fixed is not None:
            rhs[np.asarray(fixed)] = 0.0
        out.append(dict(g=kappa * pm, pm=pm, smax=smax, kappa=kappa, rhs=rhs, gx=gx))
    return out
