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
