"""The strict hierarchy and the STALL -> strict fallback on the device
(mg.py:561-586 rules, mg.py:928-965 fallback; reference tests/test_mg.py:
234-251), against the oracle's strict hierarchy and a direct solve."""
import numpy as np
import pytest

from tests.conftest import cantilever_bar
from tests.test_gpu_parity import _voided

pytestmark = pytest.mark.gpu


def _direct(mesh, fixed, f, e, ke):
    import scipy.sparse.linalg as spla

    from oracle.fem import assemble

    k = assemble(mesh, ke, e)
    free = np.flatnonzero(~fixed)
    u = np.zeros(mesh.n_dof)
    u[free] = spla.spsolve(k[free][:, free].tocsc(), f[free])
    return u


def test_strict_levels_bit_exact(gpu):
    """Strict rules (held injection, kept >= 1) give the oracle's masks on
    every level, on a voided bar with random moduli."""
    from loadpath_amd.mg import GridElasticity
    from oracle.mg import GridOracle

    mesh, fixed, f, e = _voided()
    g = GridElasticity(mesh, 0.3, fixed, coarse_cells=20)
    o = GridOracle(mesh, 0.3, fixed, coarse_cells=20)
    g.strict = True
    o.strict = True
    g.set_moduli(e)
    o.set_moduli(e)
    assert g.strict and o.strict
    assert len(g.levels) == len(o.levels) >= 3
    for i, lv in enumerate(o.levels):
        assert g.levels[i].dims == lv.dims
        assert np.array_equal(g.level_mask(i), lv.mask), i
    # the strict levels differ from the plain ones (else the test pins nothing)
    p = GridOracle(mesh, 0.3, fixed, coarse_cells=20)
    p.set_moduli(e)
    assert any(not np.array_equal(a.mask, b.mask) for a, b in zip(p.levels[1:], o.levels[1:]))


def test_strict_solve_matches_direct_and_oracle(gpu):
    from loadpath_amd.mg import GridElasticity
    from oracle.fem import hex8_ke
    from oracle.mg import GridOracle

    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=8)
    ref = _direct(mesh, fixed, f, e, hex8_ke(0.3))
    g = GridElasticity(mesh, 0.3, fixed, rtol=1e-9, coarse_cells=100, maxiter=2000)
    o = GridOracle(mesh, 0.3, fixed, rtol=1e-9, coarse_cells=100, maxiter=2000)
    g.strict = True
    o.strict = True
    g.set_moduli(e)
    o.set_moduli(e)
    u = g.solve_many([f])[0]
    uo = o.solve_many([f])[0]
    assert np.linalg.norm(u - ref) < 1e-6 * np.linalg.norm(ref)
    assert abs(g.last_iters - o.last_iters) <= 2, (g.last_iters, o.last_iters)
    assert np.linalg.norm(u - uo) < 1e-7 * np.linalg.norm(uo)


def test_stall_moves_to_strict(gpu, monkeypatch):
    """A first attempt that does not converge in the plain hierarchy's
    allowance builds the strict one, keeps it, and solves again from the
    start; the oracle with the same allowance takes the same path. The
    allowance is cut to one CG iteration on both sides so the switch
    happens on a small grid."""
    import oracle.mg as om
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity
    from oracle.fem import hex8_ke

    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=3)
    ref = _direct(mesh, fixed, f, e, hex8_ke(0.3))
    g = GridElasticity(mesh, 0.3, fixed, rtol=1e-9, coarse_cells=100, maxiter=2000)
    assert _lib.lib().lp_tune(g._h, 9, 2) == 1
    g.set_moduli(e)
    assert not g.strict
    u = g.solve_many([f, 2 * f])
    assert g.strict
    assert _lib.lib().lp_tune(g._h, 11, 0) == 1
    monkeypatch.setattr(om, "STALL", 1)
    o = om.GridOracle(mesh, 0.3, fixed, rtol=1e-9, coarse_cells=100, maxiter=2000)
    o.set_moduli(e)
    uo = o.solve_many([f])[0]
    assert o.strict
    assert np.linalg.norm(u[0] - ref) < 1e-6 * np.linalg.norm(ref)
    assert np.linalg.norm(u[1] - 2 * ref) < 1e-6 * np.linalg.norm(2 * ref)
    assert abs(g.last_iters - o.last_iters) <= 2, (g.last_iters, o.last_iters)
    # strict is sticky: the next build is strict and solves without a switch
    g.set_moduli(e * 1.5)
    assert g.strict
    w = g.solve_many([f], [u[0]])[0]
    assert np.linalg.norm(w - ref / 1.5) < 1e-6 * np.linalg.norm(ref / 1.5)
    assert _lib.lib().lp_tune(g._h, 11, 0) == 1


def test_simp_loop_stall_moves_to_strict(gpu):
    """The device top3d loop goes through the same fallback: with the
    allowance cut to one iteration its first solve moves to the strict
    hierarchy, and its iterates still equal the oracle's direct-solve loop."""
    from loadpath_amd import _lib
    from loadpath_amd.problems import cantilever_3d
    from loadpath_amd.simp import GridSimp
    from oracle import simp as osimp

    p = cantilever_3d(16, 4, 8)
    s = GridSimp(p, element="plain", rtol=1e-11, maxiter=2000, coarse_cells=50)
    _lib.lib().lp_tune(s.grid._h, 9, 2)
    st = s.iterate(3)
    assert s.grid.strict
    q = osimp.Problem(p.name, p.dims, p.volfrac, p.rmin, p.fixed, p.cases)
    ro = osimp.optimize_top3d(q, element="plain", iterations=3)
    assert abs(st[-1]["compliance"] - ro["compliance_reported"]) <= 1e-8 * ro["compliance_reported"]
    assert np.abs(s.xphys - ro["x"]).max() < 1e-8
