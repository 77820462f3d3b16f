"""Pins oracle.fem (the element and the numbering) with closed forms,
invariants and brute force, independent of the oracle's own formulas."""
import itertools

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle.fem import CORNERS, Mesh, assemble, box_mesh, hex8_ke, isotropic_c, rigid_modes


def _plain_ke_physical(nu, order=3):
    """Trilinear hex8 on the unit cube in physical coordinates, N_c =
    prod_d (x_d or 1 - x_d), integrated with an order^3 Gauss-Legendre rule
    (the integrand is quadratic per axis, so this is exact)."""
    pts, wts = np.polynomial.legendre.leggauss(order)
    pts = 0.5 * (pts + 1.0)
    wts = 0.5 * wts
    c = isotropic_c(1.0, nu)
    k = np.zeros((24, 24))
    for (x, wx), (y, wy), (z, wz) in itertools.product(zip(pts, wts), repeat=3):
        p = np.array([x, y, z])
        b = np.zeros((6, 24))
        for n in range(8):
            f = np.where(CORNERS[n] == 1, p, 1.0 - p)
            s = np.where(CORNERS[n] == 1, 1.0, -1.0)
            g = np.array([s[0] * f[1] * f[2], s[1] * f[0] * f[2], s[2] * f[0] * f[1]])
            b[0, 3 * n], b[1, 3 * n + 1], b[2, 3 * n + 2] = g
            b[3, 3 * n], b[3, 3 * n + 1] = g[1], g[0]
            b[4, 3 * n + 1], b[4, 3 * n + 2] = g[2], g[1]
            b[5, 3 * n], b[5, 3 * n + 2] = g[2], g[0]
        k += wx * wy * wz * b.T @ c @ b
    return k


@pytest.mark.parametrize("nu", [0.3, 0.2, 0.45])
def test_plain_ke_is_the_exact_trilinear_integral(nu):
    a, b = hex8_ke(nu, "plain"), _plain_ke_physical(nu)
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


@pytest.mark.parametrize("element", ["im", "plain"])
def test_ke_symmetric_psd_six_rigid_modes(element):
    k = hex8_ke(0.3, element)
    assert np.array_equal(k, k.T)
    ev = np.linalg.eigvalsh(k)
    assert np.sum(np.abs(ev) < 1e-10) == 6 and ev.min() > -1e-10
    rb = rigid_modes(CORNERS.astype(float))
    assert np.abs(k @ rb).max() < 1e-12


def _bending_nodal(kappa, nu):
    """Nodal values of the exact pure-bending field of a unit cube centred at
    the origin: sigma_xx = E kappa y."""
    x = CORNERS - 0.5
    u = np.zeros((8, 3))
    u[:, 0] = kappa * x[:, 0] * x[:, 1]
    u[:, 1] = -0.5 * kappa * x[:, 0] ** 2 - 0.5 * nu * kappa * (x[:, 1] ** 2 - x[:, 2] ** 2)
    u[:, 2] = -nu * kappa * x[:, 1] * x[:, 2]
    return u.ravel()


def test_im_element_exact_in_pure_bending():
    """Energy of the exact bending field = E kappa^2 I: 1/24 per unit cube
    for "im"; the plain trilinear element locks (stiffer)."""
    for nu in (0.3, 0.0, 0.4):
        u = _bending_nodal(1.0, nu)
        e_im = 0.5 * u @ hex8_ke(nu, "im") @ u
        e_plain = 0.5 * u @ hex8_ke(nu, "plain") @ u
        assert e_im == pytest.approx(1.0 / 24.0, rel=1e-12)
        assert e_plain > 1.1 / 24.0


@pytest.mark.parametrize("element", ["im", "plain"])
def test_constant_strain_patch(element):
    """A 3x3x3 block held on its outer nodes at a linear field: the inner
    nodes come back exact (test_hex8_im.py:27-43 of the reference)."""
    mesh = box_mesh(3, 3, 3, pitch=0.7)
    a = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4], [-1e-4, 3e-4, 7e-4]])
    xyz = mesh.node_ijk * 0.7
    exact = (xyz @ a.T).ravel()
    edge = np.any((mesh.node_ijk == 0) | (mesh.node_ijk == 3), axis=1)
    fixed = np.repeat(edge, 3)
    k = assemble(mesh, hex8_ke(0.3, element) * 0.7, np.full(mesh.n_elem, 1000.0))
    free = ~fixed
    u = exact.copy()
    rhs = -(k[free][:, fixed] @ exact[fixed])
    u[free] = spla.spsolve(k[free][:, free].tocsc(), rhs)
    assert np.allclose(u, exact, atol=1e-12)


def test_mesh_numbering_brute_force():
    rng = np.random.default_rng(0)
    for _ in range(4):
        g = rng.random((5, 4, 6)) < 0.5
        ijk = np.argwhere(g)
        m = Mesh(ijk)
        corners = ijk[:, None, :] + CORNERS[None]
        assert np.array_equal(m.node_ijk[m.enode], corners)
        flat = np.ravel_multi_index(tuple(m.node_ijk.T), tuple(m.node_dims))
        assert np.all(np.diff(flat) > 0)                       # ascending C order
        assert set(map(tuple, m.node_ijk)) == set(map(tuple, corners.reshape(-1, 3)))
        assert np.array_equal(m.edof, 3 * np.repeat(m.enode, 3, axis=1) + np.tile(np.arange(3), 8))
