"""Pins oracle.mg with closed forms (sparse prolongation rows, assembled
Galerkin products), direct solves and eigen-solves."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle.fem import assemble, box_mesh, hex8_ke
from oracle.mg import GridOracle, coarse_matrix, prolong, prolong_rows, restrict
from tests.conftest import cantilever_bar


def test_transfers_adjoint_and_linear_exact():
    rng = np.random.default_rng(1)
    c = rng.random((3, 4, 5, 6))
    f = rng.random((3, 7, 9, 11))
    assert abs(np.vdot(prolong(c), f) - np.vdot(c, restrict(f))) < 1e-10
    x = np.arange(4.0)[None, :, None, None] * np.ones((3, 4, 5, 6))
    assert np.allclose(prolong(x)[0, :, 0, 0], np.arange(7) / 2.0)


def test_prolong_equals_closed_form_rows():
    rng = np.random.default_rng(2)
    cs = (4, 3, 5)
    c = rng.random(cs)
    fs = tuple(2 * n - 1 for n in cs)
    fijk = np.stack(np.unravel_index(np.arange(int(np.prod(fs))), fs), axis=1)
    p = prolong_rows(fijk, cs)
    assert np.allclose(prolong(c).ravel(), p @ c.ravel(), rtol=0, atol=1e-15)


def _vec_p(fine_nodes, coarse_nodes):
    fijk = np.stack(np.unravel_index(np.arange(int(np.prod(fine_nodes))), fine_nodes), axis=1)
    p = prolong_rows(fijk, coarse_nodes)
    return sp.block_diag([p, p, p]).tocsr()


def _level_matrix(lv):
    """Explicit sparse matrix of a level's unmasked K on its node grid
    (component-major dofs)."""
    n = 3 * int(np.prod(lv.node_shape[1:]))
    cols = []
    eye = np.eye(n)
    for j in range(n):
        cols.append(lv.apply_k(eye[j].reshape(lv.node_shape)).ravel())
    return np.array(cols).T


def test_galerkin_levels_equal_assembled_products():
    mesh, fixed, f, e = cantilever_bar(8, 4, 4, seed=3)
    g = GridOracle(mesh, 0.3, fixed, coarse_cells=10, vdtype=np.float64)
    g.set_moduli(e)
    assert len(g.levels) >= 3
    k0 = _level_matrix(g.levels[0])
    m0 = g.levels[0].mask.ravel()
    # brute force fine K from the assembled evaluator matrix
    ka = assemble(mesh, g.ke, e).toarray()
    perm = np.concatenate([np.arange(c, mesh.n_dof, 3) for c in range(3)])
    assert np.allclose(k0, ka[np.ix_(perm, perm)], atol=1e-13)
    # level 1 = P^T D K0 D P (first coarsening is masked, mg.py:609-642)
    fine_nodes = g.levels[0].node_shape[1:]
    c1 = g.levels[1].node_shape[1:]
    pad = tuple(2 * n - 1 for n in c1)
    p = _vec_p(pad, c1)
    sel = np.zeros((3,) + pad, bool)
    sel[:, :fine_nodes[0], :fine_nodes[1], :fine_nodes[2]] = True
    pf = p[np.flatnonzero(sel.ravel())]
    d = sp.diags(m0)
    k1 = (pf.T @ d @ sp.csr_matrix(k0) @ d @ pf).toarray()
    assert np.allclose(_level_matrix(g.levels[1]), k1, atol=1e-12 * np.abs(k1).max())
    # level 2 = P^T K1 P, unmasked
    c2 = g.levels[2].node_shape[1:]
    pad2 = tuple(2 * n - 1 for n in c2)
    p2 = _vec_p(pad2, c2)
    sel2 = np.zeros((3,) + pad2, bool)
    sel2[:, :c1[0], :c1[1], :c1[2]] = True
    p2 = p2[np.flatnonzero(sel2.ravel())]
    k2 = (p2.T @ sp.csr_matrix(k1) @ p2).toarray()
    assert np.allclose(_level_matrix(g.levels[2]), k2, atol=1e-12 * np.abs(k2).max())
    # the coarsest dense matrix is D K D + (I - D) of that level
    a = coarse_matrix(g.levels[-1])
    m = g.levels[-1].mask.ravel()
    want = m[:, None] * _level_matrix(g.levels[-1]) * m[None] + np.diag(1 - m)
    assert np.allclose(a, 0.5 * (want + want.T), atol=1e-12 * np.abs(want).max())


def _direct(mesh, fixed, f, e, ke):
    k = assemble(mesh, ke, e)
    free = ~fixed
    u = np.zeros(mesh.n_dof)
    u[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    return u


@pytest.mark.parametrize("smoother,contrast", [("chebyshev", True), ("chebyshev", False), ("jacobi", False)])
def test_pcg_matches_direct(smoother, contrast):
    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=2, contrast=contrast)
    ref = _direct(mesh, fixed, f, e, hex8_ke(0.3))
    g = GridOracle(mesh, 0.3, fixed, rtol=1e-9, coarse_cells=100, maxiter=2000, smoother=smoother)
    g.set_moduli(e)
    u = g.solve_many([f])[0]
    assert np.linalg.norm(u - ref) < 1e-6 * np.linalg.norm(ref)
    g.set_moduli(e * 1.1)
    w = g.solve_many([f], [u])[0]
    assert np.linalg.norm(w - ref / 1.1) < 1e-6 * np.linalg.norm(ref)
    assert g.last_iters <= 3                     # the scaled warm start is nearly exact


def test_lmax_bounds_the_spectrum():
    mesh, fixed, f, e = cantilever_bar(8, 4, 4, seed=5)
    g = GridOracle(mesh, 0.3, fixed, coarse_cells=10, vdtype=np.float64)
    g.set_moduli(e)
    for lv in g.vlevels[:2]:
        a = _level_matrix(lv)
        m = lv.mask.ravel()
        a = m[:, None] * a * m[None] + np.diag(1 - m)
        dinv = lv.dinv.ravel()
        lam = np.linalg.eigvals(dinv[:, None] * a).real.max()
        assert 1.3 * 0.9 * lam <= lv.lmax <= 1.3 * lam * 1.000001


def test_strain_energy_closed_form():
    mesh, fixed, f, e = cantilever_bar(6, 4, 4, seed=6)
    g = GridOracle(mesh, 0.3, fixed)
    u = np.random.default_rng(7).standard_normal(mesh.n_dof)
    ue = u[mesh.edof]
    want = np.einsum("ij,jk,ik->i", ue, g.ke, ue)
    assert np.allclose(g.strain_energy(u), want, rtol=1e-13)


def test_strict_hierarchy_converges():
    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=8)
    ref = _direct(mesh, fixed, f, e, hex8_ke(0.3))
    g = GridOracle(mesh, 0.3, fixed, rtol=1e-9, coarse_cells=100, maxiter=2000)
    g.strict = True
    g.set_moduli(e)
    u = g.solve_many([f])[0]
    assert np.linalg.norm(u - ref) < 1e-6 * np.linalg.norm(ref)
