"""The constrained minimum-mass loop on the device (loadpath_amd.
constrained.optimize_constrained) against the oracle's restatement of
constrained._constrained_run (oracle/constrained.py): the same volumes,
constraint readings, move limits and beta steps iteration by iteration,
with the corner-sampled stress, the averaged stress, the layer bond and
the contact band (corners near the supports and loads not read)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem():
    from loadpath_amd.problems import cantilever_3d

    return cantilever_3d(8, 2, 4)


def _oracle_problem(p):
    from oracle import simp as os_

    return os_.Problem(p.name, p.dims, p.volfrac, p.rmin, list(p.fixed), [list(c) for c in p.cases])


@pytest.mark.parametrize("radius,bond,band", [(None, None, None), (1.3, None, None),
                                              (2.2, (0.6, (0.0, 0.0, 1.0), 0.8), None), (None, None, 1.5),
                                              (1.3, None, 1.5)])
def test_constrained_loop_matches_oracle(gpu, radius, bond, band):
    from loadpath_amd.constrained import Limits, optimize_constrained
    from loadpath_amd.simp import SimpOptions
    from oracle.constrained import Limits as OLimits
    from oracle.constrained import optimize_constrained as oracle_loop
    from oracle.loop import LoopOptions

    p = _problem()
    pres = np.zeros(p.dims, bool)
    pres[-1, :, :2] = True
    n_it = 24
    kw = dict(volume_fraction=0.5, filter_radius=1.5, max_iterations=n_it, beta_every=6, beta_max=8.0)
    r = optimize_constrained(p, 0.3, SimpOptions(solve_rtol=1e-11, **kw),
                             Limits(stress=80.0, deflection=5.0, modulus=1000.0, radius=radius, bond=bond,
                                    contact_band=band),
                             preserve=pres, precision="fp64", maxiter=4000, coarse_cells=50)
    ro = oracle_loop(_oracle_problem(p), LoopOptions(**kw),
                     OLimits(stress=80.0, deflection=5.0, modulus=1000.0, radius=radius, bond=bond,
                             contact_band=band),
                     preserve=pres.ravel())
    assert len(r.history) == len(ro["history"])
    for a, b in zip(r.history, ro["history"]):
        assert a["beta"] == b["beta"]
        assert a["volume"] == pytest.approx(b["volume"], abs=1e-6)
        assert np.allclose(a["constraints"], b["constraints"], rtol=0, atol=1e-5)
        assert a["move"] == pytest.approx(b["move"], rel=1e-12)
    assert np.abs(r.x - ro["x"]).max() < 1e-5
    assert (r.delivered is None) == (ro["delivered"] is None)


def test_constrained_rejects_what_it_does_not_build(gpu):
    from loadpath_amd.constrained import Limits, optimize_constrained
    from loadpath_amd.simp import SimpOptions

    p = _problem()
    o = SimpOptions(volume_fraction=0.5, filter_radius=1.5, max_iterations=2)
    with pytest.raises(NotImplementedError):
        optimize_constrained(p, 0.3, o, Limits(stress=80.0, deflection=None, modulus=1.0, buckling=2.0))
    with pytest.raises(ValueError):
        optimize_constrained(p, 0.3, o, Limits(stress=None, deflection=None, modulus=1.0))


def _tie_problem():
    """A cantilever 8 x 4 x 4 loaded off its y centre and obliquely, so
    neither the supports nor the loads are symmetric in y: a y mirror and z
    extrusion columns are each a real tie."""
    from loadpath_amd.problems import BoxProblem, nodes

    nx, ny, nz = 8, 4, 4
    return BoxProblem("cantilever-offset", (nx, ny, nz), 0.5, 1.5,
                      [(nodes(0, range(ny + 1), range(nz + 1)), "xyz")],
                      [[(nodes(nx, 1, range(nz + 1)), "z", -0.2), (nodes(nx, 1, range(nz + 1)), "y", 0.1)]])


def test_constrained_ties_match_oracle(gpu):
    """A y mirror and extrusion along z (constrained.py:775-780, 796-799)
    on a problem that is neither: the device loop's iterates equal the
    oracle's, and every iterate satisfies the ties themselves (x equals x
    at the mirror index, and each z column is one value)."""
    from loadpath_amd.constrained import Limits, optimize_constrained
    from loadpath_amd.simp import SimpOptions
    from oracle.constrained import Limits as OLimits
    from oracle.constrained import optimize_constrained as oracle_loop
    from oracle.loop import LoopOptions

    p = _tie_problem()
    kw = dict(volume_fraction=0.5, filter_radius=1.5, max_iterations=16, beta_every=6, beta_max=8.0)
    r = optimize_constrained(p, 0.3, SimpOptions(solve_rtol=1e-11, **kw),
                             Limits(stress=80.0, deflection=5.0, modulus=1000.0), precision="fp64",
                             mirror="y", extrusion=2, maxiter=4000, coarse_cells=50)
    ro = oracle_loop(_oracle_problem(p), LoopOptions(**kw), OLimits(stress=80.0, deflection=5.0, modulus=1000.0),
                     mirror="y", extrusion=2)
    assert len(r.history) == len(ro["history"])
    for a, b in zip(r.history, ro["history"]):
        assert a["volume"] == pytest.approx(b["volume"], abs=1e-6)
        assert np.allclose(a["constraints"], b["constraints"], rtol=0, atol=1e-5)
    x = r.x.reshape(p.dims)
    assert np.array_equal(x, x[:, ::-1, :])                 # x[i, j, k] == x[i, ny-1-j, k]
    assert np.array_equal(x, np.broadcast_to(x[:, :, :1], x.shape))   # one value per z column
    assert np.abs(r.x - ro["x"]).max() < 1e-5
    # without the ties the same problem's design is not symmetric in y
    r0 = optimize_constrained(p, 0.3, SimpOptions(solve_rtol=1e-11, **dict(kw, max_iterations=4)),
                              Limits(stress=80.0, deflection=5.0, modulus=1000.0), precision="fp64", maxiter=4000,
                              coarse_cells=50)
    x0 = r0.x.reshape(p.dims)
    assert np.abs(x0 - x0[:, ::-1, :]).max() > 1e-3