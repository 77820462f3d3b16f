"""The FP64 CG's direction and solution updates fused into the next operator
launch (lp_tune param 21, the default): the operator stages p = z + beta
p_old as it loads each plane, writes p and A p, and adds the pending step
alpha p_old to x, with the separate update pass's arithmetic. The solve must
be bitwise the one with the separate pass (reference CG: mg.py:872-926),
also when the right-hand sides converge at different iterations (one
finished, its last step added by the next launch alone) and when the
iteration cap ends the loop with the updates still deferred."""
import numpy as np
import pytest

from tests.conftest import cantilever_bar

pytestmark = pytest.mark.gpu


def _loads(mesh, f, seed):
    rng = np.random.default_rng(seed)
    g = np.zeros_like(f)
    top = np.nonzero(mesh.node_ijk[:, 2] == mesh.node_ijk[:, 2].max())[0]
    g[3 * top + rng.integers(0, 3, len(top))] = rng.standard_normal(len(top))
    # three cases of different difficulty: they stop at different iterations
    return [f, g, 1e-3 * f + g]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("maxiter", [2000, 7])
def test_fused_cg_update_is_the_separate_pass(precision, maxiter):
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity

    mesh, fixed, f, e = cantilever_bar(24, 10, 9, seed=41)
    loads = _loads(mesh, f, 42)
    out = {}
    for fused in (1, 0):
        g = GridElasticity(mesh, 0.3, fixed, precision=precision, coarse_cells=30, rtol=1e-10, maxiter=maxiter)
        assert _lib.lib().lp_tune(g._h, 21, fused) == 1
        g.set_moduli(e)
        b = np.stack([g.to_grid(v) for v in loads])
        x, its, ok = g.solve_grids(b)
        out[fused] = (x, its.copy())
    xa, ia = out[1]
    xb, ib = out[0]
    assert np.isfinite(xa).all()
    assert np.array_equal(ia, ib), (ia, ib)
    if maxiter > 100:
        assert len(set(ia.tolist())) > 1, ia   # the cases stop at different iterations
    assert np.array_equal(xa, xb)
