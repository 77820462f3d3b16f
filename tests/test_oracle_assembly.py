"""The oracle's restatement of fem.Stiffness's CSR assembly (fem.py:377-526)
against an independent construction: scipy's COO -> CSR of every element's
scaled Ke (the same pattern, the values to rounding) and the Dirichlet
identity rows."""
import numpy as np

from oracle.fem import Mesh, assemble, hex8_ke, stiffness_csr


def _mesh(seed=6):
    rng = np.random.default_rng(seed)
    g = rng.random((6, 5, 4)) < 0.75
    g[0, 0, 0] = True
    return rng, Mesh(np.argwhere(g) + np.array([2, 0, 1]), 0.8)


def test_pattern_and_values_match_scipy():
    rng, mesh = _mesh()
    ke = hex8_ke(0.3) * mesh.pitch
    scale = rng.uniform(0.01, 2.0, mesh.n_elem)
    data, indices, indptr = stiffness_csr(mesh, ke, scale, np.zeros(mesh.n_dof, bool))
    k = assemble(mesh, ke, scale)
    k.sum_duplicates()
    k.sort_indices()
    assert np.array_equal(indptr, k.indptr) and np.array_equal(indices, k.indices)
    assert np.abs(data - k.data).max() <= 1e-14 * np.abs(k.data).max()


def test_dirichlet_rows_are_identity():
    rng, mesh = _mesh(8)
    ke = hex8_ke(0.3, "plain")
    scale = rng.uniform(0.01, 2.0, mesh.n_elem)
    fixed = rng.random(mesh.n_dof) < 0.1
    import scipy.sparse as sp

    data, indices, indptr = stiffness_csr(mesh, ke, scale, fixed)
    k = sp.csr_matrix((data, indices, indptr), shape=(mesh.n_dof, mesh.n_dof)).toarray()
    full = assemble(mesh, ke, scale).toarray()
    free = ~fixed
    assert np.abs(k[np.ix_(free, free)] - full[np.ix_(free, free)]).max() <= 1e-14 * np.abs(full).max()
    assert np.array_equal(k[fixed], np.eye(mesh.n_dof)[fixed])
    assert not k[:, fixed][free].any()
