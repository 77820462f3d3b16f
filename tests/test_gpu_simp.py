"""The device SIMP loop against the oracle's top3d loop and the reference's
golden design (benchmarks/reference/cantilever-3d.top3d.npz, copied to
tests/golden/)."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _oracle_problem(p):
    from oracle import simp as os_

    fixed = [(nd, comps) for nd, comps in p.fixed]
    cases = [[(nd, c, v) for nd, c, v in case] for case in p.cases]
    return os_.Problem(p.name, p.dims, p.volfrac, p.rmin, fixed, cases)


def test_first_iterations_match_oracle(gpu):
    from loadpath_amd.problems import cantilever_3d
    from loadpath_amd.simp import GridSimp
    from oracle import simp as os_

    p = cantilever_3d(16, 4, 8)
    s = GridSimp(p, element="plain", rtol=1e-11, maxiter=2000, coarse_cells=50)
    hist_o = []
    ro = os_.optimize_top3d(_oracle_problem(p), element="plain", iterations=3,
                            on_iter=lambda it, c, ch: hist_o.append((c, ch)))
    hist = s.iterate(3)
    for (co, cho), st in zip(hist_o, hist):
        assert st["compliance"] == pytest.approx(co, rel=1e-8)
        assert st["change"] == pytest.approx(cho, rel=1e-6, abs=1e-9)
    assert np.allclose(s.xphys, ro["x"], rtol=0, atol=1e-8)


def test_top3d_golden_cantilever(gpu):
    """The golden run stops at iteration 133 on change <= 0.01, a rule that
    is sensitive to solver rounding (the oracle itself runs to the 200 cap
    with an MMD-ordered LU instead of the default one). So the device loop
    runs the golden's 133 iterations and must land on the same design: the
    same compliance to 1e-3, volume 0.3 and a near-identical density field."""
    from loadpath_amd.problems import cantilever_3d
    from loadpath_amd.simp import optimize_top3d

    z = np.load(GOLDEN / "cantilever-3d.top3d.npz")
    r = optimize_top3d(cantilever_3d(), iterations=int(z["iterations"]), element="plain", rtol=1e-11,
                       maxiter=3000, coarse_cells=500)
    print("device c", r["compliance_reported"], "golden", float(z["compliance_reported"]),
          "max|dx|", float(np.abs(r["x"] - z["x"]).max()), "mean|dx|", float(np.abs(r["x"] - z["x"]).mean()))
    assert r["compliance_reported"] == pytest.approx(float(z["compliance_reported"]), rel=1e-3)
    assert abs(r["x"].mean() - 0.3) < 2e-3
    # a few grey cells still chatter between runs of the loop (the change
    # metric never drops below 0.01 here): the field agrees on average
    assert np.abs(r["x"] - z["x"]).mean() < 3e-2
    assert ((r["x"] > 0.5) != (z["x"] > 0.5)).mean() < 0.03


def test_mma_iterations_match_oracle(gpu):
    """The device loop with the MMA update (lp_simp_set_optimizer) against
    the oracle's MMA loop (direct solves, numpy MMA): the same compliance,
    volume and design over the first iterations."""
    from loadpath_amd.problems import cantilever_3d
    from loadpath_amd.simp import GridSimp
    from oracle import simp as os_

    p = cantilever_3d(16, 4, 8)
    s = GridSimp(p, element="plain", rtol=1e-12, maxiter=3000, coarse_cells=50, optimizer="mma")
    hist_o, xo, xpo = os_.optimize_mma(_oracle_problem(p), iterations=4, element="plain")
    hist = s.iterate(4)
    for (co, vo, cho), st in zip(hist_o, hist):
        assert st["compliance"] == pytest.approx(co, rel=1e-8)
        assert st["volume"] == pytest.approx(vo, rel=1e-10)
        assert st["change"] == pytest.approx(cho, rel=1e-6, abs=1e-9)
    assert np.allclose(s.x.ravel(), xo, rtol=0, atol=1e-7)
    assert np.allclose(s.xphys.ravel(), xpo, rtol=0, atol=1e-7)


def test_projection_matches_oracle(gpu):
    """lp_project against simp._project / _dproject restated (simp.py:1443-1450),
    through numpy and through the torch.autograd binding on CUDA tensors."""
    import torch

    from loadpath_amd.autograd import project as tproject
    from loadpath_amd.simp import dproject, project
    from oracle.simp import dproject as odp
    from oracle.simp import project as op

    x = np.random.default_rng(11).uniform(0, 1, 20000)
    for beta, eta in ((1.0, 0.5), (8.0, 0.3), (64.0, 0.55)):
        assert np.allclose(project(x, beta, eta), op(x, beta, eta), rtol=1e-14, atol=1e-15)
        # 1 - tanh^2 cancels far from eta: a 1-ulp tanh difference is an
        # absolute error of a few ulp of the derivative's scale beta / den
        d = odp(x, beta, eta)
        assert np.abs(dproject(x, beta, eta) - d).max() <= 1e-13 * np.abs(d).max()
    xt = torch.tensor(x, dtype=torch.float64, device="cuda", requires_grad=True)
    y = tproject(xt, 8.0, 0.3)
    (y * 2.0).sum().backward()
    assert np.allclose(y.detach().cpu().numpy(), op(x, 8.0, 0.3), rtol=1e-14, atol=1e-15)
    d = 2.0 * odp(x, 8.0, 0.3)
    assert np.abs(xt.grad.cpu().numpy() - d).max() <= 1e-13 * np.abs(d).max()
