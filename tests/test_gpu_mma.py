"""The device MMA (lp_mma_*) against the MMA oracle, step by step, and on
the reference's known-answer problem."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spheres(x):
    f = np.array([((x - [5, 2, 1]) ** 2).sum() - 9, ((x - [3, 4, 3]) ** 2).sum() - 9])
    df = np.array([2 * (x - [5, 2, 1]), 2 * (x - [3, 4, 3])])
    return (x ** 2).sum(), 2 * x, f, df


def test_svanberg_toy_problem_on_device(gpu):
    from loadpath_amd.mma import MMA

    opt = MMA(3, 2, xmin=0, xmax=5, move=1.0)
    x = np.array([4.0, 3.0, 2.0])
    for _ in range(40):
        xn = opt.update(x, *_spheres(x))
        if np.abs(xn - x).max() < 1e-8:
            break
        x = xn
    assert np.allclose(x, [2.0175, 1.7800, 1.2375], atol=2e-4), x


@pytest.mark.parametrize("n,m", [(5000, 1), (20000, 3)])
def test_updates_track_oracle(gpu, n, m):
    from loadpath_amd.mma import MMA
    from oracle.mma import MMA as OracleMMA

    rng = np.random.default_rng(n + m)
    dev, ref = MMA(n, m, move=0.2), OracleMMA(n, m, move=0.2)
    x = rng.uniform(0.1, 0.9, n)
    # a smooth separable problem: quadratic objective, linear constraints
    target = rng.uniform(0, 1, n)
    A = rng.uniform(0.5, 1.5, (m, n)) / n
    for k in range(6):
        f0 = ((x - target) ** 2).sum()
        df0 = 2 * (x - target)
        f = A @ x - 0.3
        xd = dev.update(x, f0, df0, f, A)
        xr = ref.update(x, f0, df0, f, A)
        assert np.abs(xd - xr).max() <= 1e-9, (k, np.abs(xd - xr).max())
        assert np.allclose(dev.lam, ref.lam, rtol=1e-7, atol=1e-12)
        assert np.allclose(dev.low, ref.low, rtol=0, atol=1e-12)
        assert np.allclose(dev.upp, ref.upp, rtol=0, atol=1e-12)
        x = xr
    dev.tighten(0.5)
    ref.tighten(0.5)
    assert np.allclose(dev.low, ref.low, atol=1e-12) and np.allclose(dev.upp, ref.upp, atol=1e-12)


def test_nonfinite_inputs_on_device(gpu):
    from loadpath_amd.mma import MMA
    from oracle.mma import MMA as OracleMMA

    x = np.array([0.5, np.nan, 0.2, 0.9])
    args = (1.0, np.array([1.0, np.inf, -np.inf, np.nan]), np.array([np.nan]), np.ones((1, 4)))
    xd = MMA(4, 1).update(x, *args)
    xr = OracleMMA(4, 1).update(x, *args)
    assert np.all(np.isfinite(xd)) and np.abs(xd - xr).max() <= 1e-9
