"""verify.nodal_stress / nodal_von_mises on the device (lp_nodal_stress,
loadpath_amd.verify) against the oracle's restatement and the reference's
closed forms (tests/test_hex8_im.py:27-84)."""
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stiff(mesh, nu=0.3, element="im", mask=None):
    return SimpleNamespace(mesh=mesh, nu=nu, element=element, stress_mask=mask)


@pytest.mark.parametrize("element", ["im", "plain"])
def test_nodal_stress_matches_oracle(gpu, element):
    from loadpath_amd.verify import nodal_stress, nodal_von_mises
    from oracle import stress as S
    from oracle.fem import Mesh

    rng = np.random.default_rng(7)
    g = rng.random((12, 9, 10)) < 0.75          # a voxel set with holes, odd sizes
    mesh = Mesh(np.argwhere(g) + np.array([3, 1, 2]), 0.8, (1.0, 2.0, 3.0))
    u = rng.standard_normal(mesh.n_dof)
    mod = rng.uniform(0.1, 3.0, mesh.n_elem)
    th = rng.uniform(-1e-2, 1e-2, mesh.n_elem)
    mask = rng.random(mesh.n_elem) < 0.8
    for kw, want in [
        (dict(modulus=mod, thermal=th, mask=mask), S.nodal_stress(mesh, u, mod, 0.3, element, thermal=th, mask=mask)),
        (dict(modulus=2.5, thermal=None, mask=None), S.nodal_stress(mesh, u, 2.5, 0.3, element)),
    ]:
        st = _stiff(mesh, 0.3, element, kw["mask"])
        got = nodal_stress(st, u, kw["modulus"], kw["thermal"])
        scale = np.abs(want).max()
        assert np.abs(got - want).max() <= 1e-12 * scale
        vm = nodal_von_mises(st, u, kw["modulus"], kw["thermal"])
        assert np.allclose(vm, S.von_mises(want), rtol=1e-12, atol=1e-12 * scale)


def test_nodal_stress_closed_forms(gpu):
    """The constant-strain patch reads C eps at every node, and the
    one-cell-deep strip in pure bending reads M c / I ("im")."""
    from loadpath_amd.verify import nodal_stress
    from oracle.fem import Mesh, isotropic_c
    from tests.test_oracle_stress import _pure_bending

    e, nu = 1000.0, 0.3
    mesh = Mesh(np.argwhere(np.ones((3, 3, 3), bool)), 0.7, (0.1, -0.2, 0.3))
    a = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4], [-1e-4, 3e-4, 7e-4]])
    u = (mesh.node_ijk * mesh.pitch @ a.T).ravel()
    eps = np.array([a[0, 0], a[1, 1], a[2, 2], a[0, 1] + a[1, 0], a[1, 2] + a[2, 1], a[2, 0] + a[0, 2]])
    sig = isotropic_c(e, nu) @ eps
    s = nodal_stress(_stiff(mesh, nu, "im"), u, e)
    assert np.allclose(s, sig[None], rtol=0, atol=1e-9 * np.abs(sig).max())
    for t in (1, 2):
        _, (m, ub, eb, nub) = _pure_bending(t, "im")
        s = nodal_stress(_stiff(m, nub, "im"), ub, eb)
        x = m.node_ijk.astype(float)
        top = np.isclose(x[:, 1], t) & (x[:, 0] > 2) & (x[:, 0] < 10)
        assert s[top, 0].mean() / (t / 2.0) == pytest.approx(1.0, abs=1e-6)


@pytest.mark.parametrize("radius", [0.9, 1.4, 2.6, 3.2, 3.7, 7.1])
def test_material_average_matches_oracle(gpu, radius):
    """verify._MaterialAverage on the device against the oracle's
    restatement: a holed voxel set, partial and empty fills, quarter-point
    samples (radius <= 1.5 cells), centre samples past that and blocks of
    2^3 and 3^3 cells past 3.3 cells."""
    from loadpath_amd.verify import MaterialAverage
    from oracle import stress as S
    from oracle.fem import Mesh

    rng = np.random.default_rng(11)
    mesh = Mesh(np.argwhere(rng.random((11, 8, 9)) < 0.8) + np.array([2, 0, 5]), 0.6, (1.0, -2.0, 0.5))
    xyz = mesh.origin + mesh.node_ijk * mesh.pitch
    fill = rng.uniform(0.0, 1.0, mesh.n_elem)
    fill[rng.random(mesh.n_elem) < 0.2] = 0.0
    sig = rng.standard_normal((mesh.n_nodes, 6))
    r = radius * mesh.pitch
    want = S.material_average(xyz, mesh.enode, fill, mesh.pitch, r, sig)
    avg = MaterialAverage(mesh, fill, r)
    for s in (sig, 3.0 * sig[::-1].copy()):
        got = avg(s)
        ref = want if s is sig else S.material_average(xyz, mesh.enode, fill, mesh.pitch, r, s)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())

