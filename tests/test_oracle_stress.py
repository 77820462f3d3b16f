"""Pins the oracle's p-mean stress aggregate and its adjoint gradient
(oracle/stress.py) by closed forms and FP64 central finite differences."""
import numpy as np
import pytest

from oracle import stress as S
from oracle.fem import box_mesh, hex8_ke, isotropic_c


def _affine_cell(grad):
    """Corner displacements of the unit cube under u(x) = grad @ x."""
    from oracle.fem import CORNERS

    return np.concatenate([grad @ c.astype(float) for c in CORNERS])


def _voigt_strain(grad):
    e = 0.5 * (grad + grad.T)
    return np.array([e[0, 0], e[1, 1], e[2, 2], 2 * e[0, 1], 2 * e[1, 2], 2 * e[0, 2]])


@pytest.mark.parametrize("points", ["center", "corners"])
def test_affine_field_gives_c_eps(points):
    rng = np.random.default_rng(0)
    grad = rng.standard_normal((3, 3))
    ue = _affine_cell(grad)
    want = isotropic_c(1.0, 0.3) @ _voigt_strain(grad)
    for cb in S.sample_cb(0.3, points):
        assert np.allclose(cb @ ue, want, rtol=0, atol=1e-12)


def test_pitch_scales_stress():
    rng = np.random.default_rng(1)
    ue = rng.standard_normal(24)
    a = S.sample_cb(0.3, "corners", pitch=1.0)
    b = S.sample_cb(0.3, "corners", pitch=0.5)
    assert np.allclose(b @ ue, 2.0 * (a @ ue))


def test_von_mises_closed_forms():
    s = np.array([3.0, 0, 0, 0, 0, 0])
    assert np.sqrt(s @ S.VM @ s) == pytest.approx(3.0)
    t = np.array([0, 0, 0, 0, 2.0, 0])
    assert np.sqrt(t @ S.VM @ t) == pytest.approx(2.0 * np.sqrt(3.0))
    # hydrostatic stress carries no von Mises stress
    h = np.array([5.0, 5.0, 5.0, 0, 0, 0])
    assert abs(h @ S.VM @ h) < 1e-12


def test_pmean_limits():
    rng = np.random.default_rng(2)
    ue = rng.standard_normal((50, 24))
    x = rng.uniform(0.2, 1.0, 50)
    cbs = S.sample_cb(0.3, "center")
    g8, pm8, smax, _ = S.pnorm_forward(ue, x, cbs, p=8)
    g200, pm200, _, _ = S.pnorm_forward(ue, x, cbs, p=200)
    assert pm8 < smax and pm200 < smax
    assert abs(pm200 - smax) < abs(pm8 - smax)
    # identical cells: the P-mean equals the common value for every P
    same = np.repeat(ue[:1], 10, axis=0)
    _, pm, smax1, _ = S.pnorm_forward(same, np.full(10, 0.7), cbs, p=8)
    assert pm == pytest.approx(smax1, rel=1e-12)


def _bar(nx=6, ny=3, nz=3, seed=3):
    mesh = box_mesh(nx, ny, nz)
    fixed = np.zeros(mesh.n_dof, bool)
    root = np.nonzero(mesh.node_ijk[:, 0] == 0)[0]
    for i in range(3):
        fixed[3 * root + i] = True
    f = np.zeros(mesh.n_dof)
    tip = np.nonzero(mesh.node_ijk[:, 0] == nx)[0]
    f[3 * tip + 2] = -1.0 / len(tip)
    x = np.random.default_rng(seed).uniform(0.1, 1.0, mesh.n_elem)
    return mesh, fixed, f, x


@pytest.mark.parametrize("points", ["center", "corners"])
def test_pnorm_gradient_matches_central_differences(points):
    mesh, fixed, f, x = _bar()
    ke = hex8_ke(0.3)
    cbs = S.sample_cb(0.3, points)
    g, grad, _, _ = S.pnorm_gradient(mesh, ke, fixed, f, x, cbs)
    rng = np.random.default_rng(4)
    h = 1e-6
    for e in rng.choice(mesh.n_elem, 8, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (S.pnorm_value(mesh, ke, fixed, f, xp, cbs) - S.pnorm_value(mesh, ke, fixed, f, xm, cbs)) / (2 * h)
        assert grad[e] == pytest.approx(fd, rel=2e-5, abs=1e-9 * abs(g)), (e, grad[e], fd)


def test_compliance_gradient_matches_central_differences():
    """The self-adjoint case: dc/dx = -p x^(p-1) (1 - Emin) u_e^T Ke u_e."""
    import scipy.sparse.linalg as spla

    from oracle.fem import assemble
    from oracle.simp import compliance_sensitivity, element_energy

    mesh, fixed, f, x = _bar(seed=5)
    ke = hex8_ke(0.3)
    free = ~fixed

    def comp(xv):
        k = assemble(mesh, ke, 1e-9 + xv ** 3 * (1 - 1e-9))
        u = np.zeros(mesh.n_dof)
        u[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
        return f @ u, u

    c, u = comp(x)
    dc = compliance_sensitivity(x, element_energy(u, mesh.edof, ke))
    h = 1e-6
    for e in np.random.default_rng(6).choice(mesh.n_elem, 6, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (comp(xp)[0] - comp(xm)[0]) / (2 * h)
        assert dc[e] == pytest.approx(fd, rel=1e-5)


# ---- verify.nodal_stress (oracle/stress.py nodal_stress) -------------------

def _mesh(ijk, pitch=1.0, origin=(0.0, 0.0, 0.0)):
    from oracle.fem import Mesh

    return Mesh(np.asarray(ijk), pitch, origin)


@pytest.mark.parametrize("element", ["im", "plain"])
def test_nodal_stress_constant_strain_patch(element):
    """A linear field on a 3x3x3 block returns C eps at every node
    (the reference's tests/test_hex8_im.py:27-43)."""
    e, nu = 1000.0, 0.3
    mesh = _mesh(np.argwhere(np.ones((3, 3, 3), bool)), 0.7, (0.1, -0.2, 0.3))
    a = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4], [-1e-4, 3e-4, 7e-4]])
    x = mesh.node_ijk * mesh.pitch
    u = (x @ a.T).ravel()
    eps = np.array([a[0, 0], a[1, 1], a[2, 2], a[0, 1] + a[1, 0], a[1, 2] + a[2, 1], a[2, 0] + a[0, 2]])
    sig = isotropic_c(e, nu) @ eps
    s = S.nodal_stress(mesh, u, e, nu, element)
    assert np.allclose(s, sig[None], rtol=0, atol=1e-9 * np.abs(sig).max())
    # a free thermal expansion of the same size reads no stress
    th = 4e-4
    s0 = S.nodal_stress(mesh, (x * th).ravel(), e, nu, element, thermal=np.full(mesh.n_elem, th))
    assert np.abs(s0).max() < 1e-9 * e * th


def _pure_bending(t, element, w=2, length=12):
    """A strip t cells deep under a couple on its end face, the reference's
    set-up (tests/test_hex8_im.py:46-73): top-fibre stress / (M c / I)."""
    import scipy.sparse.linalg as spla
    from oracle.fem import assemble

    e, nu = 1000.0, 0.3
    mesh = _mesh([(i, j, k) for i in range(length) for j in range(t) for k in range(w)])
    x = mesh.node_ijk.astype(float)
    yc = t / 2.0
    root, end = np.isclose(x[:, 0], 0), np.isclose(x[:, 0], length)
    fixed = np.zeros(mesh.n_dof, bool)
    fixed[3 * np.nonzero(root)[0]] = True
    n0 = np.nonzero(root & np.isclose(x[:, 1], 0) & np.isclose(x[:, 2], 0))[0][0]
    n1 = np.nonzero(root & np.isclose(x[:, 1], 0) & np.isclose(x[:, 2], w))[0][0]
    fixed[[3 * n0 + 1, 3 * n0 + 2, 3 * n1 + 1]] = True
    f = np.zeros(mesh.n_dof)
    for n in np.nonzero(end)[0]:
        y, z = x[n, 1], x[n, 2]
        fy = sum((y - yc) / 3.0 + (y + d - yc) / 6.0 for d in (-1, 1) if 0 <= y + d <= t)
        fz = sum(0.5 for d in (-1, 1) if 0 <= z + d <= w)
        f[3 * n] = fy * fz
    k = assemble(mesh, hex8_ke(nu, element), np.full(mesh.n_elem, e))
    free = ~fixed
    u = np.zeros(mesh.n_dof)
    u[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    s = S.nodal_stress(mesh, u, e, nu, element)
    top = np.isclose(x[:, 1], t) & (x[:, 0] > 2) & (x[:, 0] < length - 2)
    return s[top, 0].mean() / yc, (mesh, u, e, nu)


def test_nodal_stress_pure_bending_closed_form():
    """The incompatible-mode element reads M c / I one cell deep; the plain
    trilinear one shear-locks (tests/test_hex8_im.py:76-83)."""
    for t in (1, 2):
        assert _pure_bending(t, "im")[0] == pytest.approx(1.0, abs=1e-6)
    assert _pure_bending(1, "plain")[0] < 0.8


def test_nodal_stress_matches_a_per_cell_loop():
    """Against a plain loop over cells and corners (no np.add.at), with a
    stress mask, per-cell moduli and a thermal strain."""
    from oracle.fem import hex8_b

    rng = np.random.default_rng(2)
    g = rng.random((4, 3, 3)) < 0.8
    mesh = _mesh(np.argwhere(g), 0.5)
    nu = 0.3
    u = rng.standard_normal(mesh.n_dof)
    mod = rng.uniform(0.5, 2.0, mesh.n_elem)
    th = rng.uniform(-1e-2, 1e-2, mesh.n_elem)
    mask = rng.random(mesh.n_elem) < 0.7
    s = S.nodal_stress(mesh, u, mod, nu, "im", thermal=th, mask=mask)
    c = isotropic_c(1.0, nu)
    acc = np.zeros((mesh.n_nodes, 6))
    cnt = np.zeros(mesh.n_nodes)
    for ci in range(mesh.n_elem):
        if not mask[ci]:
            continue
        ue = u[mesh.edof[ci]]
        for k in range(8):
            b = hex8_b(2.0 * (k & 1) - 1.0, 2.0 * ((k >> 1) & 1) - 1.0, 2.0 * ((k >> 2) & 1) - 1.0, nu, "im") / 0.5
            strain = b @ ue
            strain[:3] -= th[ci]
            acc[mesh.enode[ci, k]] += mod[ci] * (c @ strain)
            cnt[mesh.enode[ci, k]] += 1
    want = acc / np.maximum(cnt, 1)[:, None]
    assert np.allclose(s, want, rtol=1e-12, atol=1e-12)
    vm = S.von_mises(s)
    assert np.allclose(vm, np.sqrt(np.einsum("ni,ij,nj->n", s, S.VM, s)), rtol=1e-12)


# ---- verify._MaterialAverage (oracle/stress.py material_average) -----------

@pytest.mark.parametrize("radius", [0.9, 1.4, 2.6])
def test_material_average_against_a_brute_force_ball(radius):
    from oracle.fem import Mesh

    rng = np.random.default_rng(5)
    mesh = Mesh(np.argwhere(rng.random((6, 5, 4)) < 0.8), 0.7, (0.3, 0.0, -0.4))
    xyz = mesh.origin + mesh.node_ijk * mesh.pitch
    fill = rng.uniform(0.0, 1.0, mesh.n_elem)
    fill[rng.random(mesh.n_elem) < 0.2] = 0.0
    sig = rng.standard_normal((mesh.n_nodes, 6))
    r = radius * mesh.pitch
    got = S.material_average(xyz, mesh.enode, fill, mesh.pitch, r, sig)
    rho = radius
    qs = [(0.25 + 0.5 * (k & 1), 0.25 + 0.5 * ((k >> 1) & 1), 0.25 + 0.5 * ((k >> 2) & 1)) for k in range(8)] \
        if rho <= 1.5 else [(0.5, 0.5, 0.5)]
    want = np.zeros_like(sig)
    for i in range(mesh.n_nodes):
        acc, tot = np.zeros(6), 0.0
        for c in range(mesh.n_elem):
            if fill[c] <= 0:
                continue
            for q in qs:
                wk = np.array([(q[0] if k & 1 else 1 - q[0]) * (q[1] if (k >> 1) & 1 else 1 - q[1])
                               * (q[2] if (k >> 2) & 1 else 1 - q[2]) for k in range(8)])
                xs = wk @ xyz[mesh.enode[c]]
                if np.linalg.norm(xs - xyz[i]) <= r:
                    w = fill[c] * mesh.pitch ** 3 / len(qs)
                    acc += w * (wk @ sig[mesh.enode[c]])
                    tot += w
        want[i] = acc / tot if tot > 0 else sig[i]
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_material_average_keeps_a_uniform_field():
    from oracle.fem import Mesh

    mesh = Mesh(np.argwhere(np.ones((5, 4, 3), bool)), 1.0)
    sig = np.tile(np.array([1.0, -2.0, 0.5, 0.3, 0.0, -0.7]), (mesh.n_nodes, 1))
    fill = np.linspace(0.1, 1.0, mesh.n_elem)
    got = S.material_average(mesh.node_ijk * 1.0, mesh.enode, fill, 1.0, 1.2, sig)
    assert np.allclose(got, sig, rtol=0, atol=1e-14)


@pytest.mark.parametrize("radius", [3.7, 7.1])
def test_material_average_blocks_on_a_linear_field(radius):
    """Past 3.3 cells the samples merge by blocks: for a linear nodal field
    the trilinear samples are exact, a block's merged tensor is the field at
    its filled centroid, and a node's average is the field at the weighted
    centroid of the blocks in its ball. Blocks built here by a dictionary
    over integer cell keys, balls by brute force."""
    from oracle.fem import Mesh

    rng = np.random.default_rng(9)
    mesh = Mesh(np.argwhere(rng.random((13, 11, 9)) < 0.7) + np.array([1, 2, 0]), 0.5, (0.2, 0.1, -0.3))
    xyz = mesh.origin + mesh.node_ijk * mesh.pitch
    fill = rng.uniform(0.05, 1.0, mesh.n_elem)
    fill[rng.random(mesh.n_elem) < 0.15] = 0.0
    a = rng.standard_normal((6, 3))
    b = rng.standard_normal(6)
    sig = xyz @ a.T + b
    r = radius * mesh.pitch
    got = S.material_average(xyz, mesh.enode, fill, mesh.pitch, r, sig)
    blk = int(np.ceil(radius / 3.3))
    live = fill > 0
    lo = mesh.elem_ijk[live].min(axis=0)
    blocks = {}
    for e in np.nonzero(live)[0]:
        k = tuple((mesh.elem_ijk[e] - lo) // blk)
        w = fill[e] * mesh.pitch ** 3
        c = mesh.origin + (mesh.elem_ijk[e] + 0.5) * mesh.pitch
        acc = blocks.setdefault(k, [0.0, np.zeros(3)])
        acc[0] += w
        acc[1] += w * c
    wb = np.array([v[0] for v in blocks.values()])
    cb = np.array([v[1] / v[0] for v in blocks.values()])
    for i in range(mesh.n_nodes):
        inside = np.linalg.norm(cb - xyz[i], axis=1) <= r
        if inside.any():
            want = a @ (wb[inside] @ cb[inside] / wb[inside].sum()) + b
        else:
            want = sig[i]
        assert np.allclose(got[i], want, rtol=1e-10, atol=1e-10)
    assert len(blocks) < live.sum()


# ---- the constrained loop's averaged stress reading (constrained.py:146-252, 517-605)

def _averaged_case(seed=3, shape=(5, 4, 3)):
    from oracle.fem import Mesh

    rng = np.random.default_rng(seed)
    g = rng.random(shape) < 0.85
    g[0, 0, 0] = True
    mesh = Mesh(np.argwhere(g), 0.7)
    u = rng.standard_normal(mesh.n_dof)
    xe = rng.uniform(0.2, 1.0, mesh.n_elem)
    readable = rng.random(mesh.n_nodes) < 0.8
    return rng, mesh, u, xe, readable


@pytest.mark.parametrize("radius,bond", [(1.2, None), (2.0, None), (1.2, (0.6, (0.3, -0.2, 0.9), 0.8))])
def test_averaged_gradients_against_central_differences(radius, bond):
    """The adjoint right-hand side is kappa dpm/du and the explicit part
    kappa dpm/dxe at fixed u: directional central differences of pm along
    random directions (every cell moved)."""
    rng, mesh, u, xe, readable = _averaged_case()
    r = radius * mesh.pitch
    kw = dict(p=8.0, q=1.0, limit=1.7, bond=bond)
    ev = S.averaged_pnorm(mesh, 0.3, r, readable, [u], xe, **kw)[0]

    def pm(uu, xx):
        return S.averaged_pnorm(mesh, 0.3, r, readable, [uu], xx, **kw)[0]["pm"]

    for _ in range(3):
        du = rng.standard_normal(mesh.n_dof)
        h = 1e-5
        fd = (pm(u + h * du, xe) - pm(u - h * du, xe)) / (2 * h)
        assert ev["rhs"] @ du / ev["kappa"] == pytest.approx(fd, rel=1e-6)
        dx = rng.standard_normal(mesh.n_elem)
        fd = (pm(u, xe + h * dx) - pm(u, xe - h * dx)) / (2 * h)
        assert ev["gx"] @ dx / ev["kappa"] == pytest.approx(fd, rel=1e-6)


def test_averaged_uniform_stress_reads_c_eps():
    """An affine displacement field on a full, fully dense block: every
    sample reads C eps, so every node does, rho = 1 and pm = s_max =
    vm(C eps) / lim; kappa = 1 at the first iteration, the mean of the two
    readings after it."""
    from oracle.fem import Mesh, isotropic_c

    mesh = Mesh(np.argwhere(np.ones((4, 3, 3), bool)), 0.5)
    a = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4], [-1e-4, 3e-4, 7e-4]])
    u = (mesh.node_ijk * mesh.pitch @ a.T).ravel()
    eps = np.array([a[0, 0], a[1, 1], a[2, 2], a[0, 1] + a[1, 0], a[1, 2] + a[2, 1], a[2, 0] + a[0, 2]])
    sig = isotropic_c(1.0, 0.3) @ eps
    vm = np.sqrt(sig @ S.VM @ sig)
    for radius in (0.8, 1.9):
        ev = S.averaged_pnorm(mesh, 0.3, radius * mesh.pitch, np.ones(mesh.n_nodes, bool), [u],
                              np.ones(mesh.n_elem), limit=2.0)[0]
        assert ev["pm"] == pytest.approx(vm / 2.0, rel=1e-10)
        assert ev["smax"] == pytest.approx(vm / 2.0, rel=1e-10)
        assert ev["kappa"] == pytest.approx(1.0, rel=1e-10)
        ev2 = S.averaged_pnorm(mesh, 0.3, radius * mesh.pitch, np.ones(mesh.n_nodes, bool), [u],
                               np.ones(mesh.n_elem), limit=2.0, kappa_prev=[3.0], it=2)[0]
        assert ev2["kappa"] == pytest.approx(2.0, rel=1e-10)
