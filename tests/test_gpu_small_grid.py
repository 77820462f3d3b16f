"""Small grids: the fine operator's launch shape and the coarsest dense
factor.

- The x-chunk of the fine-level kernels (node planes a block or wave
  marches) only partitions the work: every node's sums run in the same
  order whatever the chunk, so the products at 4, 8 and 16 planes, and at
  the untuned size-based default, are equal to the bit (FP64 vector kernel
  and FP16-MFMA wave kernel).
- The coarsest solve is a dense FP64 inverse (blocked Cholesky with 64-wide
  panels, per-row panel substitution, batched diagonal-block inverses): on a
  two-level grid the FP64 V-cycle's coarsest step is that solve, so the
  V-cycle must equal the oracle's (its coarsest step a Cholesky factor and
  triangular solves, mg._coarse_solve, mg.py:708-766) to FP64 rounding, for
  coarsest sizes below one panel, with a partial last panel and at 21 and
  26 panels.
"""
import numpy as np
import pytest

from tests.conftest import cantilever_bar

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fine_operator_xchunk_does_not_change_products(gpu, precision):
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity

    mesh, fixed, f, e = cantilever_bar(40, 10, 9, seed=21)
    g = GridElasticity(mesh, 0.3, fixed, precision=precision, coarse_cells=50)
    g.set_moduli(e)
    dt = np.float64 if precision == "fp64" else np.float32
    rng = np.random.default_rng(3)
    u = rng.standard_normal((2,) + g.levels[0].node_shape).astype(dt)
    base = g.apply(u, level=0, dtype=dt)          # the size-based default
    lib = _lib.lib()
    for xc in (4, 8, 16):
        assert lib.lp_tune(g._h, 1, xc) >= 0
        got = g.apply(u, level=0, dtype=dt)
        assert np.array_equal(got, base), xc


@pytest.mark.parametrize("cells", [(4, 4, 4), (8, 4, 4), (16, 12, 12), (20, 12, 12)])
def test_coarsest_dense_solve_in_the_vcycle_matches_oracle(gpu, cells):
    """Two-level grids whose coarsest level has 81, 135, 1,323 and 1,617
    dofs (one partial panel; three panels, the last partial; 21 panels; 26
    panels): the FP64 V-cycle, whose coarsest step is the dense solve,
    equals the oracle's (Cholesky and triangular solves) to FP64 rounding."""
    from loadpath_amd.mg import GridElasticity
    from oracle.mg import GridOracle

    mesh, fixed, f, e = cantilever_bar(*cells, seed=5)
    cc = int(np.prod(cells)) // 8
    g = GridElasticity(mesh, 0.3, fixed, precision="fp64", coarse_cells=cc)
    o = GridOracle(mesh, 0.3, fixed, vdtype=np.float64, coarse_cells=cc)
    g.set_moduli(e)
    o.set_moduli(e)
    assert len(g.levels) == 2 and len(o.levels) == 2 and o.coarse is not None
    rng = np.random.default_rng(9)
    r = rng.standard_normal((2,) + o.levels[0].node_shape) * o.levels[0].mask
    got = g.precondition(r)
    for k in range(2):
        want = o.precondition(r[k])
        assert np.abs(got[k] - want).max() <= 1e-9 * np.abs(want).max(), (cells, k)
