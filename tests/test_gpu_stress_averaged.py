"""The constrained loop's material-averaged stress reading on the GPU
(constrained._Averaged, constrained.py:146-252, read at 517-605;
lp_stress_averaged) against the FP64 oracle: the aggregate, Le's
normalization, the adjoint right-hand side and the explicit gradient, and
the whole gradient through the adjoint solve against a direct solve."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _holed(seed=4):
    from oracle.fem import Mesh

    rng = np.random.default_rng(seed)
    g = rng.random((10, 7, 6)) < 0.8
    g[0, 0, 0] = True
    return rng, Mesh(np.argwhere(g) + np.array([1, 0, 2]), 0.6)


@pytest.mark.parametrize("radius,bond", [(1.2, None), (2.3, None), (1.2, (0.5, (0.0, 0.4, 1.0), 0.7)),
                                         (3.1, (0.8, (1.0, 0.0, 0.0), 1.2))])
def test_averaged_parts_match_oracle(gpu, radius, bond):
    from loadpath_amd.stress import StressAveraged
    from oracle import stress as S

    rng, mesh = _holed()
    us = [rng.standard_normal(mesh.n_dof) for _ in range(2)]
    still = mesh.node_ijk[:, 0] < 4                       # part of the mesh does not move
    us[1].reshape(-1, 3)[still] = 0.0
    xe = rng.uniform(0.0, 1.0, mesh.n_elem)
    xe[rng.random(mesh.n_elem) < 0.15] = 0.0
    readable = rng.random(mesh.n_nodes) < 0.85
    fixed = np.flatnonzero(rng.random(mesh.n_dof) < 0.05)
    r = radius * mesh.pitch
    kw = dict(p=8.0, q=1.0, limit=2.5, bond=bond)
    agg = StressAveraged(mesh, 0.3, r, readable, fixed=fixed, **kw)
    for it, prev in ((1, None), (3, [1.7, 0.6])):
        got = agg.evaluate(us, xe, kappa=prev, it=it)
        want = S.averaged_pnorm(mesh, 0.3, r, readable, us, xe, kappa_prev=prev, it=it, fixed=fixed, **kw)
        for a, b in zip(got, want):
            for key in ("pm", "smax", "kappa", "g"):
                assert a[key] == pytest.approx(b[key], rel=1e-12), key
            assert np.abs(a["rhs"] - b["rhs"]).max() <= 1e-11 * np.abs(b["rhs"]).max()
            assert np.abs(a["gx"] - b["gx"]).max() <= 1e-11 * np.abs(b["gx"]).max()


def test_averaged_gradient_matches_oracle_direct(gpu):
    """dg/dxe with the adjoint solved by the device multigrid PCG (FP64,
    rtol 1e-12) against the same right-hand side solved directly."""
    import scipy.sparse.linalg as spla

    from loadpath_amd.stress import StressAveraged
    from oracle import stress as S
    from oracle.fem import Mesh, assemble, hex8_ke
    from tests.test_gpu_stress import _lbracket

    penal, emin = 3.0, 1e-9
    _, mesh, fixed, f, x, g = _lbracket(16, 8, 8)
    e = emin + x ** penal * (1 - emin)
    de = penal * x ** (penal - 1) * (1 - emin)
    g.set_moduli(e)
    u = g.solve_many([f])[0]
    r = 1.3 * mesh.pitch
    readable = np.ones(mesh.n_nodes, bool)
    fx = np.flatnonzero(fixed)
    agg = StressAveraged(mesh, 0.3, r, readable, limit=1.0, fixed=fx)
    ev, grads = agg.gradient(g, [u], x, de)
    om = Mesh(mesh.elem_ijk, mesh.pitch)
    want = S.averaged_pnorm(om, 0.3, r, readable, [u], x, fixed=fx)[0]
    ke = hex8_ke(0.3)
    k = assemble(om, ke, e)
    free = ~np.asarray(fixed, bool)
    lam = np.zeros(om.n_dof)
    lam[free] = spla.splu(k[free][:, free].tocsc()).solve(want["rhs"][free])
    grad = want["gx"] - de * S.element_product(lam, u, om.edof, ke)
    assert ev[0]["g"] == pytest.approx(want["g"], rel=1e-12)
    assert np.abs(grads[0] - grad).max() <= 1e-7 * np.abs(grad).max()
