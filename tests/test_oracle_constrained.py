"""The oracle's constrained minimum-mass loop (oracle/constrained.py,
constrained._constrained_run on a box): its constraint gradients against
central differences through the direct solve, and the loop's invariants."""
import numpy as np
import pytest

from oracle.constrained import Constrained, Limits, optimize_constrained
from oracle.loop import LoopOptions
from oracle.simp import cantilever_3d


def _setup(radius=None, bond=None, band=None):
    p = cantilever_3d(8, 2, 4)
    opts = LoopOptions(volume_fraction=0.5, filter_radius=1.5, max_iterations=50, beta_every=6, beta_max=8.0)
    return p, opts, Limits(stress=80.0, deflection=5.0, modulus=1000.0, radius=radius, bond=bond, contact_band=band)


@pytest.mark.parametrize("radius,bond,band", [(None, None, None), (1.3, None, None),
                                              (2.2, (0.6, (0.0, 0.0, 1.0), 0.8), None), (None, None, 1.5),
                                              (1.3, None, 1.5)])
def test_constraint_gradients_against_central_differences(radius, bond, band):
    """Deflection: d(defl / limit)/dx; stress: kappa dpm/dx (kappa is held
    as the loop holds it), each through filter, projection, erosion and the
    solve, along random directions on the design cells."""
    p, opts, lim = _setup(radius, bond, band)
    pres = np.zeros(p.dims, bool)
    pres[-1, :, :2] = True
    L = Constrained(p, opts, lim, preserve=pres.ravel())
    rng = np.random.default_rng(2)
    x = rng.uniform(0.3, 0.9, L.n)
    x[L.pres] = 1.0
    beta = 2.0
    cons, grads, rec, kappa, _, _ = L.constraints(x, beta, [1.0], 1)
    h = 1e-6
    for _ in range(2):
        d = rng.standard_normal(L.n) * L.design
        cp, _, rp, _, _, _ = L.constraints(x + h * d, beta, [1.0], 1)
        cm, _, rm, _, _, _ = L.constraints(x - h * d, beta, [1.0], 1)
        fd_defl = (cp[0] - cm[0]) / (2 * h)
        assert grads[0] @ d == pytest.approx(fd_defl, rel=1e-6)
        fd_pm = (rp["pm.0"] - rm["pm.0"]) / (2 * h)
        assert grads[1] @ d == pytest.approx(kappa[0] * fd_pm, rel=1e-5)


def test_loop_invariants():
    """Beta only rises, the move stays within [0.02, 0.1], the delivered
    design is the lightest feasible one at the sharpest projection and
    meets every limit to the loop's slack."""
    p, opts, lim = _setup()
    r = optimize_constrained(p, opts, lim)
    h = r["history"]
    betas = [q["beta"] for q in h]
    assert all(b1 >= b0 for b0, b1 in zip(betas, betas[1:]))
    assert all(0.02 - 1e-15 <= q["move"] <= 0.1 + 1e-15 for q in h)
    d = r["delivered"]
    assert d is not None and max(d["constraints"]) <= 0.02
    sharp = [q for q in h if q["beta"] >= opts.beta_max and max(q["constraints"]) <= 0.02]
    assert d["volume"] == min(q["volume"] for q in sharp)
    assert 0.0 < d["volume"] < 1.0


def test_contact_band_masks_corners():
    """The band drops the corners near the supports and the load, not whole
    cells: some cells keep part of their corners."""
    p, opts, lim = _setup(band=1.5)
    L = Constrained(p, opts, lim)
    pc = L.pt_counted
    partial = pc.any(axis=1) & ~pc.all(axis=1)
    assert partial.any() and (~pc).any() and pc.any()
    assert np.array_equal(L.counted, pc.any(axis=1))


def test_tie_restatement_against_index_maps():
    """The oracle's mirror and column averages against the reference's
    formulation on index maps (domain.py:1981-1988: lut of the mirrored
    ijk; simp._symmetrize's pairwise means; manufacture.tie_groups' column
    labels with simp._extrude's bincount means), on random fields with
    preserved cells, to the bit."""
    from oracle.loop import column_average, mirror_average

    rng = np.random.default_rng(4)
    dims = (5, 4, 6)
    n = int(np.prod(dims))
    ijk = np.stack(np.unravel_index(np.arange(n), dims), axis=1)
    v = rng.standard_normal(n)
    design = rng.random(n) > 0.2
    for axes in ("y", "xz", "zyx"):
        want = v.copy()
        for ax in axes:
            a = "xyz".index(ax)
            m_ijk = ijk.copy()
            m_ijk[:, a] = dims[a] - 1 - m_ijk[:, a]
            m = np.ravel_multi_index(tuple(m_ijk.T), dims)
            nxt = want.copy()
            nxt[:] = 0.5 * (want + want[m])
            want = nxt
        assert np.array_equal(mirror_average(v, dims, axes), want), axes
    for a in range(3):
        others = [i for i in range(3) if i != a]
        _, col = np.unique(ijk[:, others[0]] * (dims[others[1]] + 1) + ijk[:, others[1]], return_inverse=True)
        c = col[design]
        mean = np.bincount(c, weights=v[design], minlength=col.max() + 1) / np.maximum(
            np.bincount(c, minlength=col.max() + 1), 1)
        want = v.copy()
        want[design] = mean[c]
        assert np.array_equal(column_average(v, dims, a, design), want), a


def test_mirror_and_extrusion_ties_hold():
    """With a y mirror and extrusion along z every iterate is symmetric and
    one value along each z column, on a problem that is neither."""
    from oracle.simp import Problem, nodes

    p0, opts, lim = _setup()
    p = Problem("offset", (8, 4, 4), 0.5, 1.5, [(nodes(0, range(5), range(5)), "xyz")],
                [[(nodes(8, 1, range(5)), "z", -0.2), (nodes(8, 1, range(5)), "y", 0.1)]])
    opts.max_iterations = 12
    seen = []
    r = optimize_constrained(p, opts, lim, mirror="y", extrusion=2, on_iter=lambda rec, xp: seen.append(xp.copy()))
    x = r["x"].reshape(p.dims)
    assert np.array_equal(x, x[:, ::-1, :])
    assert np.array_equal(x, np.broadcast_to(x[:, :, :1], x.shape))
    assert len(seen) == 12