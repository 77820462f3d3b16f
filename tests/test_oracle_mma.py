"""Pins the MMA oracle (oracle/mma.py): the reference's own known answer
(Svanberg's two-sphere problem, tests/test_loadpath.py:376-388 of the
reference), the subproblem's KKT conditions, and a closed-form optimum."""
import numpy as np
import pytest

from oracle.mma import MMA, _kkt, subsolve


def _spheres(x):
    f = np.array([((x - [5, 2, 1]) ** 2).sum() - 9, ((x - [3, 4, 3]) ** 2).sum() - 9])
    df = np.array([2 * (x - [5, 2, 1]), 2 * (x - [3, 4, 3])])
    return (x ** 2).sum(), 2 * x, f, df


def test_svanberg_toy_problem_known_answer():
    opt = MMA(3, 2, xmin=0, xmax=5, move=1.0)
    x = np.array([4.0, 3.0, 2.0])
    for _ in range(40):
        xn = opt.update(x, *_spheres(x))
        if np.abs(xn - x).max() < 1e-8:
            break
        x = xn
    assert np.allclose(x, [2.0175, 1.7800, 1.2375], atol=2e-4), x


def test_subproblem_solution_satisfies_kkt():
    rng = np.random.default_rng(0)
    n, m = 40, 3
    opt = MMA(n, m)
    x = rng.uniform(0.2, 0.8, n)
    df0 = rng.standard_normal(n)
    df = rng.standard_normal((m, n))
    f = rng.standard_normal(m)
    opt.iter = 1
    opt._asymptotes(x)
    alfa, beta, p0, q0, P, Q, b = opt.approximation(x, df0, f, df)
    xs, lam, st = subsolve(opt.low, opt.upp, alfa, beta, p0, q0, P, Q, b, opt.a0, opt.a, opt.c, opt.d,
                           return_state=True)
    r = _kkt(st, 1e-7, opt.low, opt.upp, alfa, beta, p0, q0, P, Q, b, opt.a0, opt.a, opt.c, opt.d)
    assert np.abs(r).max() < 1e-6
    assert np.all(xs >= alfa) and np.all(xs <= beta)
    assert np.all(lam >= 0)


def test_linear_problem_reaches_closed_form_vertex():
    """min c.x s.t. mean(x) <= 0.3, 0 <= x <= 1: x = 1 on the 3 most negative
    costs of 10, 0 elsewhere."""
    n = 10
    cost = -np.arange(1.0, n + 1)
    opt = MMA(n, 1, move=0.2)
    x = np.full(n, 0.3)
    for _ in range(200):
        f = np.array([x.mean() / 0.3 - 1.0])
        df = np.full((1, n), 1.0 / (0.3 * n))
        xn = opt.update(x, cost @ x, cost, f, df)
        if np.abs(xn - x).max() < 1e-9:
            break
        x = xn
    want = np.zeros(n)
    want[-3:] = 1.0
    assert np.abs(x - want).max() < 1e-3, x


def test_nonfinite_inputs_read_as_violations():
    opt = MMA(4, 1)
    x = np.array([0.5, np.nan, 0.2, 0.9])
    xn = opt.update(x, 1.0, np.array([1.0, np.inf, -np.inf, np.nan]), np.array([np.nan]), np.ones((1, 4)))
    assert np.all(np.isfinite(xn)) and np.all((xn >= 0) & (xn <= 1))
