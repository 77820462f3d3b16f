"""The production-loop oracle (oracle/loop.py) pinned without the GPU:
finite differences for the chain rule, a brute-force root for the volume
multiplier, a quadratic for the measured step, and the loop's invariants
on a small cantilever (the properties tests/test_loadpath.py:106-118 and
:1190-1223 assert of simp.optimize)."""
import numpy as np
import pytest

from oracle.loop import LoopOptions, Loop, beta_hold, optimize_loop, step_fraction, volume_root
from oracle.simp import cantilever_3d


def _tiny():
    p = cantilever_3d(nx=4, ny=2, nz=3, volfrac=0.4, rmin=1.5)
    pres = np.zeros(p.dims, bool)
    pres[-1, :, 0] = True          # a preserved pad under the load
    return p, pres.ravel()


# at beta 8 the eroded part is soft (c ~ 2e6) and the direct solve's
# rounding shows in the differences at ~1e-6
@pytest.mark.parametrize("beta,tol", [(1.0, 1e-6), (4.0, 1e-6), (8.0, 5e-5)])
def test_chain_rule_matches_finite_differences(beta, tol):
    p, pres = _tiny()
    L = Loop(p, LoopOptions(volume_fraction=0.4, filter_radius=1.5), preserve=pres, weights=None)
    rng = np.random.default_rng(3)
    # densities above the eroded threshold: a stiff enough part that the
    # direct solve's rounding stays below the differences' truncation
    x = rng.uniform(0.55, 0.85, L.n)
    x[pres] = 1.0
    comp, dc, dv, _, _ = L.analyse(x, beta)

    def vol(xv):
        return float(L.physical(xv, beta)[1] @ L.wdesign)

    def d4(fun, j, h=1e-3):
        # fourth-order central difference: truncation O(h^4) at h where the
        # direct solve's rounding is still far below it
        def at(t):
            xv = x.copy()
            xv[j] += t
            return fun(xv)
        return (8 * (at(h) - at(-h)) - (at(2 * h) - at(-2 * h))) / (12 * h)

    for j in [0, 5, 11, 17, 23]:
        if pres[j]:
            continue
        fd_c = d4(lambda xv: L.analyse(xv, beta)[0], j)
        fd_v = d4(vol, j)
        assert abs(fd_c - dc[j]) <= tol * max(1.0, abs(dc[j])), (j, fd_c, dc[j])
        assert abs(fd_v - dv[j]) <= 1e-8 * max(1.0, abs(dv[j])), (j, fd_v, dv[j])


def test_volume_root_matches_a_bisection():
    p, pres = _tiny()
    L = Loop(p, LoopOptions(volume_fraction=0.4, filter_radius=1.5), preserve=pres)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.2, 0.8, L.n)
    x[pres] = 1.0
    beta = 8.0
    comp, dc, dv, _, _ = L.analyse(x, beta)
    dc, dv = np.minimum(dc, 0.0), np.maximum(dv, 1e-12)
    lo, hi = np.maximum(0, x - 0.2), np.minimum(1, x + 0.2)
    ratio = np.sqrt(-dc / dv)

    def stepped(lm):
        xn = np.clip(x * ratio / np.sqrt(lm), lo, hi)
        xn[pres] = 1.0
        return xn, float(L.physical(xn, beta)[1] @ L.wdesign) - L.target

    for lam_last in (None, 1e-3, 50.0):
        _, lm, f, f1, f2, n = volume_root(stepped, L.target, lam_last)
        assert f1 is not None and f2 is not None
        assert abs(f) <= 1e-13 * L.target * 1.0001 or n >= 2
        # brute force: the volume falls as lambda rises; bisect log lambda
        a, b = np.log(1e-12), np.log(1e12)
        for _ in range(200):
            m = 0.5 * (a + b)
            if stepped(np.exp(m))[1] > 0:
                a = m
            else:
                b = m
        assert abs(np.log(lm) - 0.5 * (a + b)) < 1e-9, (lm, np.exp(0.5 * (a + b)))
        assert abs(stepped(lm)[1]) <= 1e-12 * L.target


def test_step_fraction_is_the_line_minimum_of_a_quadratic():
    rng = np.random.default_rng(1)
    m = rng.standard_normal((6, 6))
    a = m @ m.T + 6 * np.eye(6)
    b = rng.standard_normal(6)
    x0 = rng.standard_normal(6)
    g = lambda x: a @ x - b
    dx = -0.3 * g(x0)
    t_exact = -(g(x0) @ dx) / (dx @ a @ dx)
    a_s, b_s = g(x0) @ dx, g(x0 + dx) @ dx
    assert a_s < 0
    assert abs(step_fraction(a_s, b_s) - t_exact) < 1e-12 * max(1.0, t_exact)
    # a slope that did not rise: no curvature seen, the step may grow
    assert step_fraction(-1.0, -2.0) == np.inf


def test_beta_hold_schedule():
    o = LoopOptions(volume_fraction=0.3, filter_radius=1.5, beta_every=10, beta_schedule=(20, 15))
    assert [beta_hold(o, b) for b in (1, 2, 4, 8, 32)] == [20, 15, 10, 10, 10]
    assert beta_hold(LoopOptions(volume_fraction=0.3, filter_radius=1.5), 4.0) == 10


def test_loop_invariants_on_a_small_cantilever():
    p = cantilever_3d(nx=16, ny=4, nz=6, volfrac=0.3, rmin=1.5)
    opts = LoopOptions(volume_fraction=0.3, filter_radius=1.5, max_iterations=70)
    r = optimize_loop(p, opts)
    h = r["history"]
    assert h[-1]["beta"] == 32.0 or r["beta"] == 32.0
    betas = [x["beta"] for x in h]
    assert betas == sorted(betas)                       # continuation only sharpens
    assert all(x["move"] <= 0.2 + 1e-12 for x in h)
    held = [x for x in h[1:] if x["volHeld"]]
    assert held
    # the volume of the design a step produced is on target to rounding (read
    # at the next iteration, at the same beta: a doubling moves it)
    pairs = [(a, b) for a, b in zip(h, h[1:]) if a["volHeld"] and a["beta"] == b["beta"]]
    assert len(pairs) > 10
    assert max(abs(b["volume"] - 0.3) for a, b in pairs) < 1e-10
    # (on a grid this small the eroded members are thin and the compliance
    # of the analysed design rises at each doubling; within the sharpest
    # stage it falls)
    assert r["converged"]
    last = [x for x in h if x["beta"] >= 32][1:]
    assert len(last) >= 5
    rise = max(b["compliance"] / a["compliance"] - 1 for a, b in zip(last, last[1:]))
    assert rise < 1e-3, rise
    assert last[-1]["compliance"] < last[0]["compliance"]


def test_keep_cores_against_a_brute_force_ball():
    """simp._keep_cores: max(opened, blueprint) on the cells whose centre is
    within radius / 2.6 + 0.5 of an analysed-solid cell, else the opening."""
    from oracle.loop import keep_cores

    rng = np.random.default_rng(4)
    dims = (9, 7, 8)
    n = int(np.prod(dims))
    opened = rng.random(n)
    analysed = np.where(rng.random(n) < 0.1, 0.9, 0.1)
    analysed[0] = 0.9
    opened[0] = 0.2                      # a lost cell, so the rule applies
    blueprint = rng.random(n)
    radius = 2.5
    got = keep_cores(dims, opened, analysed, blueprint, radius)
    reach = radius / 2.6 + 0.5
    ijk = np.argwhere(np.ones(dims, bool))
    solid = ijk[analysed > 0.5]
    want = opened.copy()
    for i, c in enumerate(ijk):
        if (((solid - c) ** 2).sum(axis=1) <= reach ** 2 * (1 + 1e-9)).any():
            want[i] = max(opened[i], blueprint[i])
    assert np.array_equal(got, want)
    # nothing lost: the opening itself
    assert keep_cores(dims, np.ones(n), analysed, blueprint, radius) is not None
    assert np.array_equal(keep_cores(dims, np.ones(n), analysed, blueprint, radius), np.ones(n))


def test_prune_removes_a_dead_tendril():
    """A cantilever bar carrying the load plus a free side branch of 5 cells:
    the branch carries no energy and is pruned, the bar stays
    (simp._prune_dead, simp.py:1244-1314)."""
    from oracle.loop import Loop, LoopOptions, prune_dead, E_MIN, PENALTY

    p = cantilever_3d(nx=12, ny=2, nz=8, volfrac=0.3, rmin=1.5)
    L = Loop(p, LoopOptions(volume_fraction=0.3, filter_radius=1.5))
    xp = np.zeros(p.dims)
    xp[:, :, 0:2] = 1.0                  # the bar along the bottom, under the tip load
    xp[5, :, 2:7] = 1.0                  # a branch standing up from it, free at the top
    xp = xp.ravel()
    us = L.solve(E_MIN + xp ** PENALTY * (1 - E_MIN))
    out, _ = prune_dead(L, xp, us)
    out = out.reshape(p.dims)
    assert (out[:, :, 0:2] == 1.0).all()
    assert out[5, :, 3:7].sum() < xp.reshape(p.dims)[5, :, 3:7].sum()
