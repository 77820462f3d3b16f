"""The mixed V-cycle's coarse levels as FP16 cell matrices (lp_tune param 20:
2, the default, every coarse level, the first one's sum_q w_q M_q with
patch matrices on cells touching held dofs; 1, the Galerkin levels below
it): each cell's 576 entries rounded to FP16 after a power-of-two scale of
its own, applied with FP32 arithmetic. Checked against the FP64 oracle's levels (mg.py:609-677)
to FP16 tolerance, with a large common offset (each row subtracts its
rounded entries' sum times the cell's corner-0 value, so translations still
cancel), and through the V-cycle and the solve."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests.conftest import cantilever_bar

pytestmark = pytest.mark.gpu

# FP16 entries (11-bit significand) with FP32 accumulation
TOL_FP16 = 4e-3


def _holed(seed=13):
    mesh, fixed, f, e = cantilever_bar(16, 10, 8, seed=seed)
    rng = np.random.default_rng(seed + 1)
    e = e.copy()
    e[rng.random(mesh.n_elem) < 0.1] = 0.0
    e[mesh.elem_ijk[:, 0] == 0] = 1.0
    return mesh, fixed, f, e


def _grids(mesh, fixed, e, half=2, **kw):
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity
    from oracle.mg import GridOracle

    g = GridElasticity(mesh, 0.3, fixed, precision="mixed", **kw)
    assert _lib.lib().lp_tune(g._h, 20, half) == 2
    g.set_moduli(e)
    o = GridOracle(mesh, 0.3, fixed, vdtype=np.float64, **kw)
    o.set_moduli(e)
    return g, o


def test_half_levels_match_oracle():
    mesh, fixed, f, e = _holed()
    g, o = _grids(mesh, fixed, e, coarse_cells=1)
    assert len(o.levels) >= 3
    rng = np.random.default_rng(21)
    for level in range(1, len(o.levels)):
        lv = o.levels[level]
        u = rng.standard_normal((3,) + lv.node_shape).astype(np.float32)
        got = g.apply(u, level=level, dtype=np.float32)
        want = np.stack([lv.apply(v.astype(np.float64)) for v in u])
        assert np.isfinite(got).all(), level
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= TOL_FP16, (level, err)


@pytest.mark.parametrize("level", [1, 2])
def test_half_levels_cancel_rigid_offset(level):
    """u = 1000 + d: away from the supports (where the masked operand is no
    longer a translation) the product keeps FP16 accuracy relative to the
    product of d, not to that of the offset."""
    mesh, fixed, f, e = cantilever_bar(16, 12, 12, seed=23)
    g, o = _grids(mesh, fixed, e, coarse_cells=1)
    lv = o.levels[level]
    rng = np.random.default_rng(24)
    d = rng.standard_normal(lv.node_shape)
    u = (1000.0 + d).astype(np.float32)
    got = g.apply(u[None], level=level, dtype=np.float32)[0]
    want = lv.apply(u.astype(np.float64))
    inner = np.zeros(lv.node_shape, bool)
    inner[:, 2:] = True          # node planes x >= 2: no cell touches a held dof
    dprod = lv.apply(d * inner)  # the deformation's own product, for scale
    ref = np.abs(dprod[:, 3:]).max()
    assert np.abs(got - want)[:, 3:].max() <= TOL_FP16 * ref


def test_half_levels_vcycle_fused_matches_separate():
    """The fused Chebyshev launches (first step staging d = dinv b / theta,
    last step adding into x) on the FP16 matrix levels equal the separate
    vector passes to FP32 rounding."""
    from loadpath_amd import _lib

    mesh, fixed, f, e = _holed(seed=31)
    g, o = _grids(mesh, fixed, e, coarse_cells=1)
    r = np.random.default_rng(32).standard_normal((2,) + g.levels[0].node_shape) * g.level_mask(0)
    a = g.precondition(r)
    assert _lib.lib().lp_tune(g._h, 2, 0) == 1
    b = g.precondition(r)
    _lib.lib().lp_tune(g._h, 2, 1)
    assert np.isfinite(a).all()
    assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max()


@pytest.mark.parametrize("half", [2, 1])
def test_cheb_residual_only_step(half):
    """Degree-2 smoothers from a nonzero x (lp_tune param 24): the residual
    launch storing r only and the last step staging d = (dinv r) / theta
    from it equal the residual launch storing r and d, to FP32 rounding of
    d (then FP16 operands), on every level kind: the fine FP16-MFMA level,
    the first coarse level as FP16 matrices (half 2) or the FP16-MFMA
    weights kernel with its patch products (half 1), the Galerkin levels;
    and the solve still lands on the direct solve."""
    from loadpath_amd import _lib
    from oracle.fem import assemble, hex8_ke

    lib = _lib.lib()
    mesh, fixed, f, e = _holed(seed=51)
    g, o = _grids(mesh, fixed, e, half=half, coarse_cells=1)
    r = np.random.default_rng(52).standard_normal((3,) + g.levels[0].node_shape) * g.level_mask(0)
    a = g.precondition(r)
    assert lib.lp_tune(g._h, 24, 0) == 1
    b = g.precondition(r)
    assert lib.lp_tune(g._h, 24, 1) == 0
    assert np.isfinite(a).all()
    assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max()
    mesh, fixed, f, e = cantilever_bar(20, 10, 10, seed=53)
    k = assemble(mesh, hex8_ke(0.3), e)
    free = ~fixed
    ref = np.zeros(mesh.n_dof)
    ref[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    g, _ = _grids(mesh, fixed, e, half=half, coarse_cells=30, rtol=1e-10, maxiter=2000)
    x = g.solve_many([f])
    assert np.linalg.norm(x[0] - ref) <= 1e-7 * np.linalg.norm(ref)


def test_half_levels_solve_matches_direct():
    """A mixed solve through the FP16 matrix levels lands on the direct solve,
    with the CG count of the FP16-MFMA weights / FP32 levels within 2; so
    do the FP16 Galerkin levels under the FP16-MFMA weights level."""
    from oracle.fem import assemble, hex8_ke

    mesh, fixed, f, e = cantilever_bar(20, 10, 10, seed=33)
    kw = dict(coarse_cells=30, rtol=1e-10, maxiter=2000)
    k = assemble(mesh, hex8_ke(0.3), e)
    free = ~fixed
    ref = np.zeros(mesh.n_dof)
    ref[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    its = []
    for half in (2, 0, 1):
        g, _ = _grids(mesh, fixed, e, half=half, **kw)
        x = g.solve_many([f, 2.0 * f])
        assert np.linalg.norm(x[0] - ref) <= 1e-7 * np.linalg.norm(ref), half
        assert np.linalg.norm(x[1] - 2.0 * ref) <= 1e-7 * np.linalg.norm(2.0 * ref), half
        its.append(g.last_iters)
    assert abs(its[0] - its[1]) <= 2 and abs(its[2] - its[1]) <= 2, its


@pytest.mark.parametrize("brick", [40808, 50603, 21302])
def test_half_brick_matches_gathers(brick):
    """The LDS-staged brick form (lp_tune param 22: ty * 10000 + tz * 100 +
    x-chunk) makes the same FP16 FMAs in the same order as the
    node-per-thread gathers, so every level's product, the fused Chebyshev
    launches of a V-cycle and a batch of 10 right-hand sides (two groups of
    8) equal the gathers' bit for bit; bricks that do not tile the level
    (65 = 8 x 8 + 1 nodes along z) and x-chunks of 2 planes included."""
    from loadpath_amd import _lib

    lib = _lib.lib()
    mesh, fixed, f, e = cantilever_bar(34, 26, 20, seed=41)
    e = e.copy()
    e[np.random.default_rng(42).random(mesh.n_elem) < 0.1] = 0.0
    e[mesh.elem_ijk[:, 0] == 0] = 1.0
    from loadpath_amd.mg import GridElasticity

    g = GridElasticity(mesh, 0.3, fixed, precision="mixed", coarse_cells=1, max_rhs=10)
    assert lib.lp_tune(g._h, 23, 0) == 512     # every level, however small
    assert lib.lp_tune(g._h, 22, 0) == 40804   # the default brick; gathers first
    g.set_moduli(e)
    rng = np.random.default_rng(43)
    for level in range(1, len(g.levels)):
        shape = g.levels[level].node_shape
        u = rng.standard_normal((10,) + tuple(shape)).astype(np.float32)
        a = g.apply(u, level=level, dtype=np.float32)
        assert lib.lp_tune(g._h, 22, brick) == 0
        b = g.apply(u, level=level, dtype=np.float32)
        assert lib.lp_tune(g._h, 22, 0) == brick
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), (level, np.abs(a - b).max())
    r = rng.standard_normal((10,) + g.levels[0].node_shape) * g.level_mask(0)
    a = g.precondition(r)
    lib.lp_tune(g._h, 22, brick)
    b = g.precondition(r)
    lib.lp_tune(g._h, 22, 0)
    assert np.array_equal(a, b), np.abs(a - b).max()


def test_half_brick_solve_and_shapes():
    """A mixed solve with the brick form on from the hierarchy build (its
    power-iteration partials per brick) lands on the direct solve with the
    gathers' CG count within 1; shapes the kernel cannot take are refused."""
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity
    from oracle.fem import assemble, hex8_ke

    lib = _lib.lib()
    mesh, fixed, f, e = cantilever_bar(20, 10, 10, seed=44)
    kw = dict(coarse_cells=30, rtol=1e-10, maxiter=2000)
    k = assemble(mesh, hex8_ke(0.3), e)
    free = ~fixed
    ref = np.zeros(mesh.n_dof)
    ref[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    its = []
    for brick in (0, 40808):
        g = GridElasticity(mesh, 0.3, fixed, precision="mixed", **kw)
        assert lib.lp_tune(g._h, 23, 0) == 512
        assert lib.lp_tune(g._h, 22, brick) == 40804
        g.set_moduli(e)
        x = g.solve_many([f, 2.0 * f])
        assert np.linalg.norm(x[0] - ref) <= 1e-7 * np.linalg.norm(ref), brick
        assert np.linalg.norm(x[1] - 2.0 * ref) <= 1e-7 * np.linalg.norm(2.0 * ref), brick
        its.append(g.last_iters)
    assert abs(its[0] - its[1]) <= 1, its
    for bad in (-1, 40800, 40008, 50708, 80808, 13201):
        assert lib.lp_tune(g._h, 22, bad) < 0, bad
    assert lib.lp_tune(g._h, 22, 0) == 40808
