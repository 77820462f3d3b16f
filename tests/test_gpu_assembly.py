"""fem.Stiffness's CSR assembly on the GPU (lp_stiffness_*, loadpath_amd.fem.
Stiffness) against the oracle's restatement of the reference's numpy path:
the same layout and the same values to the bit."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("element", ["im", "plain"])
def test_stiffness_matches_oracle_bit_for_bit(gpu, element):
    from loadpath_amd.fem import Stiffness
    from oracle.fem import Mesh, hex8_ke, stiffness_csr

    rng = np.random.default_rng(13)
    g = rng.random((9, 7, 8)) < 0.7
    g[0, 0, 0] = True
    mesh = Mesh(np.argwhere(g) + np.array([1, 3, 0]), 0.7)
    fixed = rng.random(mesh.n_dof) < 0.08
    st = Stiffness(mesh, 0.3, fixed, element)
    assert np.abs(st.ke - hex8_ke(0.3, element) * mesh.pitch).max() <= 1e-14 * np.abs(st.ke).max()
    for _ in range(2):
        scale = rng.uniform(1e-6, 2.0, mesh.n_elem)
        data, indices, indptr = stiffness_csr(mesh, st.ke, scale, fixed)
        assert np.array_equal(st.indptr, indptr) and np.array_equal(st.indices, indices)
        got = st.assemble(scale)
        assert np.array_equal(got.data, data)


def test_stiffness_at_the_route_limit(gpu):
    """A 100k-cell box (the assembled route's upper end): the layout and one
    assembly, timed; K applied to a rigid translation vanishes on the free
    rows away from the supports."""
    from loadpath_amd.fem import Stiffness, box_mesh

    mesh = box_mesh(64, 40, 40)
    fixed = np.zeros(mesh.n_dof, bool)
    held = np.flatnonzero(mesh.node_ijk[:, 0] == 0)
    for i in range(3):
        fixed[3 * held + i] = True
    t0 = time.perf_counter()
    st = Stiffness(mesh, 0.3, fixed, "im")
    t1 = time.perf_counter()
    scale = np.random.default_rng(1).uniform(0.1, 1.0, mesh.n_elem)
    st.assemble_data(scale)
    t2 = time.perf_counter()
    k = st.assemble(scale)
    t3 = time.perf_counter()
    print(f"\nlayout {t1 - t0:.3f} s, assemble {t3 - t2:.4f} s (nnz {st.nnz}, first call {t2 - t1:.4f} s)")
    u = np.zeros(mesh.n_dof)
    u[0::3] = 1.0
    r = k @ u
    away = np.repeat(mesh.node_ijk[:, 0] >= 2, 3)
    assert np.abs(r[away]).max() <= 1e-12 * np.abs(k.data).max()
