"""The mixed-precision path: FP16-MFMA fine-level contraction inside an FP32
V-cycle under the FP64 CG. The operator is checked against the FP64 oracle
to FP16 tolerance (with and without a large rigid offset, which the
per-cell translation removal must cancel); the solve and the SIMP loop must
land on the FP64 answers to the solve tolerance."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests.conftest import cantilever_bar

pytestmark = pytest.mark.gpu

# FP16 operands (11-bit significand) with FP32 accumulation: relative error
# of a Ke.u product ~ sqrt(24) * 2^-11 ~ 2.4e-3 of its magnitude at worst
TOL_FP16 = 4e-3


def _grids(mesh, fixed, e, **kw):
    from loadpath_amd.mg import GridElasticity
    from oracle.mg import GridOracle

    g = GridElasticity(mesh, 0.3, fixed, precision="mixed", **kw)
    o = GridOracle(mesh, 0.3, fixed, vdtype=np.float64, **kw)
    g.set_moduli(e)
    o.set_moduli(e)
    return g, o


def test_mfma_operator_matches_fp64():
    mesh, fixed, f, e = cantilever_bar(12, 10, 9, seed=11)
    g, o = _grids(mesh, fixed, e, coarse_cells=20)
    rng = np.random.default_rng(1)
    shape = o.levels[0].node_shape
    u = rng.standard_normal((2,) + shape).astype(np.float32)
    got = g.apply(u, level=0, dtype=np.float32)
    want = np.stack([o.levels[0].apply(v.astype(np.float64)) for v in u])
    assert np.abs(got - want).max() <= TOL_FP16 * np.abs(want).max()


def test_mfma_operator_cancels_rigid_offset():
    """A large common offset on an unsupported grid (every dof free): each
    cell's FP16 operand is taken relative to its corner 0, so the product
    keeps FP16 accuracy relative to the deformation, not to the offset."""
    mesh, fixed, f, e = cantilever_bar(12, 10, 8, seed=12)
    free = np.zeros_like(fixed)
    # a single level: the unsupported operator is singular (rigid modes), so
    # no coarsest factorization; only the fine operator is exercised
    g, o = _grids(mesh, free, e, coarse_cells=10 ** 9)
    assert o.levels[0].mask.all()
    rng = np.random.default_rng(2)
    shape = o.levels[0].node_shape
    d = rng.standard_normal(shape)
    u = (1000.0 + d).astype(np.float32)
    got = g.apply(u[None], level=0, dtype=np.float32)[0]
    want = o.levels[0].apply(u.astype(np.float64))
    assert np.abs(got - want).max() <= TOL_FP16 * np.abs(want).max()


@pytest.mark.parametrize("level", [0, 1, 2])
def test_mfma_operator_tiny_magnitudes(level):
    """Cells whose corner differences are below 2^-113 (FP32 vectors of the
    V-cycle in void regions reach such values) must keep their FP16
    accuracy: the per-cell power-of-two scale stays finite, so no NaN
    comes from the exact zero of corner 0. Right-hand side 0 is tiny
    everywhere, right-hand side 1 spans magnitudes 1 ... 1e-36 plane by
    plane."""
    mesh, fixed, f, e = cantilever_bar(16, 10, 8, seed=15)
    g, o = _grids(mesh, fixed, e, coarse_cells=20)
    lv = o.levels[level]
    rng = np.random.default_rng(16)
    u = rng.standard_normal((2,) + lv.node_shape)
    u[0] *= 2e-36
    nxp = lv.node_shape[1]
    u[1] *= (10.0 ** -np.linspace(0, 36, nxp))[None, :, None, None]
    got = g.apply(u.astype(np.float32), level=level, dtype=np.float32)
    assert np.isfinite(got).all()
    for k in range(2):
        want = lv.apply(u[k].astype(np.float32).astype(np.float64))
        assert np.abs(got[k] - want).max() <= TOL_FP16 * np.abs(want).max()


def test_wave_tile_kernel_matches_block_tile_kernel():
    """The two FP16-MFMA kernels (node sums in registers across lanes, or
    through LDS) contract the same FP16 operands: the operators on the fine
    and weights levels and the whole mixed V-cycle agree to FP32 rounding.
    The grid is not a multiple of the 15-node tiles in y or z, and the
    x-chunks leave a remainder."""
    from loadpath_amd import _lib

    mesh, fixed, f, e = cantilever_bar(33, 17, 19, seed=17)
    g, o = _grids(mesh, fixed, e, coarse_cells=40)
    lib = _lib.lib()
    lib.lp_tune(g._h, 20, 0)   # the weights level as the FP16-MFMA kernels
    g.set_moduli(e)
    rng = np.random.default_rng(18)
    out = {}
    for wave in (2, 0):
        lib.lp_tune(g._h, 7, wave)
        res = []
        for level in (0, 1):
            u = rng.standard_normal((3,) + o.levels[level].node_shape).astype(np.float32)
            res.append(g.apply(u, level=level, dtype=np.float32))
        r = rng.standard_normal((2,) + g.levels[0].node_shape) * g.level_mask(0)
        res.append(g.precondition(r))
        out[wave] = res
        rng = np.random.default_rng(18)
    lib.lp_tune(g._h, 7, 1)
    for k, (a, b) in enumerate(zip(out[2], out[0])):
        assert np.isfinite(a).all(), k
        err = np.abs(a - b).max() / np.abs(b).max()
        # the fine level: the same FP16 operands, FP32 accumulation order
        # only; the weights level: the wave-tile kernel rounds the child
        # weights into the FP16 operand (FP16 tolerance); the V-cycle
        # carries those differences through its smoothers and coarse solves
        assert err <= (1e-5, TOL_FP16, 1e-2)[k], (k, err)


@pytest.mark.parametrize("wave", [1, 2])
def test_weights_level_matches_oracle(wave):
    """The first coarse level through the FP16-MFMA weights operator (sum of
    w_q M_q, explicit patches for cells touching held dofs) equals the
    oracle's level (FP64) to FP16 tolerance, on a grid with supports and
    removed cells (patch cells present), in the block-tile kernel (wave 1)
    and the wave-tile kernel (wave 2: child weights in the FP16 operand)."""
    from loadpath_amd import _lib

    mesh, fixed, f, e = cantilever_bar(16, 10, 8, seed=13)
    rng = np.random.default_rng(14)
    e = e.copy()
    e[rng.random(mesh.n_elem) < 0.1] = 0.0
    e[mesh.elem_ijk[:, 0] == 0] = 1.0
    g, o = _grids(mesh, fixed, e, coarse_cells=20)
    _lib.lib().lp_tune(g._h, 20, 0)   # the weights level as the FP16-MFMA kernels
    g.set_moduli(e)
    _lib.lib().lp_tune(g._h, 7, wave)
    lv = o.levels[1]
    u = rng.standard_normal((2,) + lv.node_shape).astype(np.float32)
    got = g.apply(u, level=1, dtype=np.float32)
    want = np.stack([lv.apply(v.astype(np.float64)) for v in u])
    assert np.abs(got - want).max() <= TOL_FP16 * np.abs(want).max()


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_explicit32_first_coarse_level(precision):
    """lp_tune param 18: the first coarse level of the FP32 / mixed V-cycle
    as its formed FP32 cell matrices (patch cells included) equals the
    oracle's level to FP32 rounding, and a solve through it lands on the
    direct solve with the CG count of the matrix-free level."""
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity
    from oracle.fem import assemble, hex8_ke
    from oracle.mg import GridOracle

    mesh, fixed, f, e_solid = cantilever_bar(16, 10, 8, seed=13)
    rng = np.random.default_rng(14)
    e = e_solid.copy()
    e[rng.random(mesh.n_elem) < 0.1] = 0.0
    e[mesh.elem_ijk[:, 0] == 0] = 1.0
    kw = dict(coarse_cells=20, rtol=1e-10, maxiter=2000)
    g = GridElasticity(mesh, 0.3, fixed, precision=precision, **kw)
    assert _lib.lib().lp_tune(g._h, 18, 1) == 0
    g.set_moduli(e)
    o = GridOracle(mesh, 0.3, fixed, vdtype=np.float64, **kw)
    o.set_moduli(e)
    lv = o.levels[1]
    u = rng.standard_normal((2,) + lv.node_shape).astype(np.float32)
    got = g.apply(u, level=1, dtype=np.float32)
    want = np.stack([lv.apply(v.astype(np.float64)) for v in u])
    assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()

    # the solve on the design without holes (a direct solve needs no
    # floating cells)
    g.set_moduli(e_solid)
    k = assemble(mesh, hex8_ke(0.3), e_solid)
    free = ~fixed
    ref = np.zeros(mesh.n_dof)
    ref[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    x = g.solve_many([f])[0]
    assert np.linalg.norm(x - ref) <= 1e-7 * np.linalg.norm(ref)
    g0 = GridElasticity(mesh, 0.3, fixed, precision=precision, **kw)
    g0.set_moduli(e_solid)
    g0.solve_many([f])
    assert abs(g.last_iters - g0.last_iters) <= 2, (g.last_iters, g0.last_iters)


@pytest.mark.parametrize("degree", [2, 3])
def test_mixed_vcycle_tracks_fp32(degree):
    """The mixed V-cycle (FP16-MFMA operators on the two finest levels, the
    Chebyshev steps fused into them) against the FP32 V-cycle of the same
    hierarchy: the same preconditioner up to FP16 rounding."""
    from loadpath_amd.mg import GridElasticity

    mesh, fixed, f, e = cantilever_bar(24, 12, 12, seed=5)
    gm = GridElasticity(mesh, 0.3, fixed, precision="mixed", coarse_cells=100, degree=degree)
    g32 = GridElasticity(mesh, 0.3, fixed, precision="fp32", coarse_cells=100, degree=degree)
    gm.set_moduli(e)
    g32.set_moduli(e)
    rng = np.random.default_rng(6)
    r = rng.standard_normal((2,) + gm.levels[0].node_shape) * gm.level_mask(0)
    zm = gm.precondition(r)
    z32 = g32.precondition(r)
    assert np.linalg.norm(zm - z32) <= 1e-2 * np.linalg.norm(z32)


def test_fused_chebyshev_matches_separate_passes():
    """The fused Chebyshev steps (start, update and last-step passes folded
    into the FP16-MFMA operator launches) give the V-cycle of the separate
    passes up to FP32 rounding."""
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity

    mesh, fixed, f, e = cantilever_bar(24, 12, 12, seed=7)
    for degree in (2, 3):
        g = GridElasticity(mesh, 0.3, fixed, precision="mixed", coarse_cells=100, degree=degree)
        g.set_moduli(e)
        r = np.random.default_rng(8).standard_normal((2,) + g.levels[0].node_shape) * g.level_mask(0)
        z_fused = g.precondition(r)
        assert _lib.lib().lp_tune(g._h, 2, 0) == 1
        z_sep = g.precondition(r)
        assert np.linalg.norm(z_fused - z_sep) <= 1e-5 * np.linalg.norm(z_sep), degree


def test_mixed_solve_matches_direct():
    from oracle.fem import assemble, hex8_ke

    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=2)
    k = assemble(mesh, hex8_ke(0.3), e)
    free = ~fixed
    ref = np.zeros(mesh.n_dof)
    ref[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    from loadpath_amd.mg import GridElasticity

    g = GridElasticity(mesh, 0.3, fixed, precision="mixed", rtol=1e-10, coarse_cells=100, maxiter=2000)
    g.set_moduli(e)
    u = g.solve_many([f, 3.0 * f])
    assert np.linalg.norm(u[0] - ref) <= 1e-7 * np.linalg.norm(ref)
    assert np.linalg.norm(u[1] - 3.0 * ref) <= 1e-7 * np.linalg.norm(3.0 * ref)
    g32 = GridElasticity(mesh, 0.3, fixed, precision="fp32", rtol=1e-10, coarse_cells=100, maxiter=2000)
    g32.set_moduli(e)
    g32.solve_many([f])
    assert g.last_iters <= 1.3 * g32.last_iters + 2, (g.last_iters, g32.last_iters)


def test_mixed_fallback_repeats_with_fp32():
    """The mixed path's safety net: a solve the FP16 levels fail is repeated
    with FP32 operators on every level, warm from the iterates. Forced here
    after one iteration (lp_tune param 5 = 2): the answer is the direct
    solve's, one repeat is counted, and the repeat needs no more iterations
    than a plain FP32 solve plus the forced one."""
    from oracle.fem import assemble, hex8_ke
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity

    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=3)
    k = assemble(mesh, hex8_ke(0.3), e)
    free = ~fixed
    ref = np.zeros(mesh.n_dof)
    ref[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    g = GridElasticity(mesh, 0.3, fixed, precision="mixed", rtol=1e-10, coarse_cells=100, maxiter=2000)
    g.set_moduli(e)
    lib = _lib.lib()
    assert lib.lp_tune(g._h, 5, 2) == 1
    u = g.solve_many([f, 2.0 * f])
    assert lib.lp_tune(g._h, 6, 0) == 1
    assert np.linalg.norm(u[0] - ref) <= 1e-7 * np.linalg.norm(ref)
    assert np.linalg.norm(u[1] - 2.0 * ref) <= 1e-7 * np.linalg.norm(2.0 * ref)
    g32 = GridElasticity(mesh, 0.3, fixed, precision="fp32", rtol=1e-10, coarse_cells=100, maxiter=2000)
    g32.set_moduli(e)
    g32.solve_many([f])
    assert g.last_iters <= g32.last_iters + 3, (g.last_iters, g32.last_iters)
    # a new hierarchy uses the FP16 levels again; unforced, the healthy
    # solve needs no repeat
    assert lib.lp_tune(g._h, 5, 1) == 2
    g.set_moduli(e)
    g.solve_many([f])
    assert lib.lp_tune(g._h, 6, 0) == 1


def test_mixed_simp_tracks_fp64():
    from loadpath_amd.problems import multiload_256x128x128
    from loadpath_amd.simp import GridSimp

    p = multiload_256x128x128(32, 16, 16)
    a = GridSimp(p, precision="mixed", rtol=1e-9, coarse_cells=0, max_levels=4)
    b = GridSimp(p, precision="fp64", rtol=1e-9, coarse_cells=0, max_levels=4)
    ha, hb = a.iterate(3), b.iterate(3)
    for sa, sb in zip(ha, hb):
        assert sa["compliance"] == pytest.approx(sb["compliance"], rel=1e-7)
    assert np.abs(a.xphys - b.xphys).max() < 1e-6


@pytest.mark.parametrize("rtol", [1e-8, 1e-11])
def test_fp32_cg_reliable_updates_match_direct(rtol):
    """The mixed path's CG in FP32 with reliable FP64 updates (lp_tune param
    13): the answer meets the tolerance in FP64 (it is checked on the FP64
    residual), warm starts included, in about the FP64 CG's iterations."""
    from oracle.fem import assemble, hex8_ke
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity

    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=2)
    k = assemble(mesh, hex8_ke(0.3), e)
    free = ~fixed
    ref = np.zeros(mesh.n_dof)
    ref[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    lib = _lib.lib()
    res = {}
    for cg32 in (0, 1):
        g = GridElasticity(mesh, 0.3, fixed, precision="mixed", rtol=rtol, coarse_cells=100, maxiter=2000)
        lib.lp_tune(g._h, 13, cg32)
        g.set_moduli(e)
        u = g.solve_many([f, 3.0 * f])
        it = g.last_iters
        # the true residual of the answer, in FP64
        for j, s in enumerate((1.0, 3.0)):
            r = (s * f - k @ u[j])[free]
            # FP32 CG: stopped on its FP64 residual; FP64 CG: on its recursion
            bound = 1.05 if cg32 else 10.0
            assert np.linalg.norm(r) <= bound * rtol * np.linalg.norm(s * f[free]), (cg32, j)
        assert np.linalg.norm(u[0] - ref) <= max(1e3 * rtol, 1e-7) * np.linalg.norm(ref)
        g.set_moduli(e * 1.1)
        w = g.solve_many([f], [u[0]])[0]
        assert np.linalg.norm(w - ref / 1.1) <= max(1e3 * rtol, 1e-7) * np.linalg.norm(ref / 1.1)
        res[cg32] = (it, int(lib.lp_tune(g._h, 14, 0)), int(lib.lp_tune(g._h, 6, 0)))
    (i64, _, _), (i32, upd, rep) = res[0], res[1]
    assert upd >= 1 and rep == 0, res
    assert i32 <= 1.15 * i64 + 3, res


@pytest.mark.parametrize("rtol", [1e-8, 1e-11])
def test_fp32_refinement_matches_direct(rtol):
    """Iterative refinement (lp_tune params 13 + 19, inner tolerance 3e-3 by
    param 15): the FP32 PCG restarts from each FP64 residual, as
    gpu.py:429-481's solve_fp32 does; the answer meets the tolerance on its
    FP64 residual, warm starts included."""
    from oracle.fem import assemble, hex8_ke
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity

    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=2)
    k = assemble(mesh, hex8_ke(0.3), e)
    free = ~fixed
    ref = np.zeros(mesh.n_dof)
    ref[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    lib = _lib.lib()
    g = GridElasticity(mesh, 0.3, fixed, precision="mixed", rtol=rtol, coarse_cells=100, maxiter=2000)
    lib.lp_tune(g._h, 13, 1)
    lib.lp_tune(g._h, 15, 3)
    assert lib.lp_tune(g._h, 19, 1) == 0
    g.set_moduli(e)
    u = g.solve_many([f, 3.0 * f])
    for j, s in enumerate((1.0, 3.0)):
        r = (s * f - k @ u[j])[free]
        assert np.linalg.norm(r) <= 1.05 * rtol * np.linalg.norm(s * f[free]), j
    assert np.linalg.norm(u[0] - ref) <= max(1e3 * rtol, 1e-7) * np.linalg.norm(ref)
    g.set_moduli(e * 1.1)
    w = g.solve_many([f], [u[0]])[0]
    assert np.linalg.norm(w - ref / 1.1) <= max(1e3 * rtol, 1e-7) * np.linalg.norm(ref / 1.1)
    assert int(lib.lp_tune(g._h, 14, 0)) >= 2 and int(lib.lp_tune(g._h, 6, 0)) == 0


def test_fp32_cg_simp_tracks_fp64():
    from loadpath_amd import _lib
    from loadpath_amd.problems import multiload_256x128x128
    from loadpath_amd.simp import GridSimp

    p = multiload_256x128x128(32, 16, 16)
    a = GridSimp(p, precision="mixed", rtol=1e-9, coarse_cells=0, max_levels=4)
    _lib.lib().lp_tune(a.grid._h, 13, 1)
    b = GridSimp(p, precision="fp64", rtol=1e-9, coarse_cells=0, max_levels=4)
    ha, hb = a.iterate(3), b.iterate(3)
    for sa, sb in zip(ha, hb):
        assert sa["compliance"] == pytest.approx(sb["compliance"], rel=1e-7)
    assert np.abs(a.xphys - b.xphys).max() < 1e-6


# BF16 operands (8-bit significand), FP32 accumulation
TOL_BF16 = 3e-2


def _bf16(g):
    from loadpath_amd import _lib

    assert _lib.lib().lp_tune(g._h, 16, 1) == 0
    return g


def test_bf16_operator_matches_fp64():
    mesh, fixed, f, e = cantilever_bar(12, 10, 9, seed=11)
    g, o = _grids(mesh, fixed, e, coarse_cells=20)
    _bf16(g)
    rng = np.random.default_rng(1)
    shape = o.levels[0].node_shape
    u = rng.standard_normal((2,) + shape).astype(np.float32)
    got = g.apply(u, level=0, dtype=np.float32)
    want = np.stack([o.levels[0].apply(v.astype(np.float64)) for v in u])
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err <= TOL_BF16, err
    # the rounding is the operand's, not a scale's: FP16 is tighter
    from loadpath_amd import _lib

    _lib.lib().lp_tune(g._h, 16, 0)
    got16 = g.apply(u, level=0, dtype=np.float32)
    assert np.abs(got16 - want).max() / np.abs(want).max() < err


def test_bf16_operator_cancels_rigid_offset_and_tiny_values():
    mesh, fixed, f, e = cantilever_bar(12, 10, 8, seed=12)
    free = np.zeros_like(fixed)
    g, o = _grids(mesh, free, e, coarse_cells=10 ** 9)
    _bf16(g)
    rng = np.random.default_rng(2)
    shape = o.levels[0].node_shape
    d = rng.standard_normal(shape)
    for u in ((1000.0 + d), d * 2e-36):
        u = u.astype(np.float32)
        got = g.apply(u[None], level=0, dtype=np.float32)[0]
        want = o.levels[0].apply(u.astype(np.float64))
        assert np.isfinite(got).all()
        assert np.abs(got - want).max() <= TOL_BF16 * np.abs(want).max()


def test_bf16_solve_and_simp_track_fp64():
    from oracle.fem import assemble, hex8_ke
    from loadpath_amd.mg import GridElasticity
    from loadpath_amd.problems import multiload_256x128x128
    from loadpath_amd.simp import GridSimp

    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=2)
    k = assemble(mesh, hex8_ke(0.3), e)
    free = ~fixed
    ref = np.zeros(mesh.n_dof)
    ref[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    g = _bf16(GridElasticity(mesh, 0.3, fixed, precision="mixed", rtol=1e-10, coarse_cells=100, maxiter=2000))
    g.set_moduli(e)
    u = g.solve_many([f])[0]
    assert np.linalg.norm(u - ref) <= 1e-7 * np.linalg.norm(ref)
    p = multiload_256x128x128(32, 16, 16)
    a = GridSimp(p, precision="mixed", rtol=1e-9, coarse_cells=0, max_levels=4)
    _bf16(a.grid)
    b = GridSimp(p, precision="fp64", rtol=1e-9, coarse_cells=0, max_levels=4)
    ha, hb = a.iterate(3), b.iterate(3)
    for sa, sb in zip(ha, hb):
        assert sa["compliance"] == pytest.approx(sb["compliance"], rel=1e-7)
    assert np.abs(a.xphys - b.xphys).max() < 1e-6
