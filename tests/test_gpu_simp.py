"""The device SIMP loop against the oracle's top3d loop and the reference's
golden designs (benchmarks/reference/{cantilever-3d,bridge-3d}.top3d.npz,
copied to tests/golden/, with the oracle's histories of them)."""
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


@pytest.mark.parametrize("name", ["cantilever-3d", "bridge-3d"])
def test_top3d_golden_runs(gpu, name):
    """The reference's golden top3d runs (plain element) through the
    oracle's history of each: the oracle's loop with the reference
    evaluator's own direct solve lands on the golden exactly (stop
    iteration, compliance to 1e-12, design to 1e-6; asserted when the
    fixture was written, and in test_oracle_simp.py), and its per-iteration
    compliance and designs are the fixture
    (tests/golden/make_top3d_histories.py). The device loop (multigrid PCG
    to rtol 1e-12) runs the golden's iteration count.

    The loop amplifies noise: two direct solves of the first system (default
    and MMD-ordered LU) already differ by ~1e-11, and that difference grows
    to 2.6e-4 in compliance by iteration 133 (the MMD run also misses the
    golden's change <= 0.01 stop). So the device is held to the oracle
    tightly where the loop has not amplified yet (compliance to 1e-9 over
    the first 10 iterations, 1e-8 over the first 50, the design after 10 to
    1e-5), and at the golden's iteration count to the spread the
    amplification produces (the device's own runs at rtol 1e-12 and 1e-14
    land 1.1e-2 and 4.5e-4 from the golden in mean |dx|)."""
    from loadpath_amd.problems import bridge_3d, cantilever_3d
    from loadpath_amd.simp import GridSimp

    z = np.load(GOLDEN / f"{name}.top3d.npz")
    if not (GOLDEN / f"{name}.history.npz").exists():
        pytest.skip(f"{name}.history.npz not written (tests/golden/make_top3d_histories.py {name})")
    h = np.load(GOLDEN / f"{name}.history.npz")
    p = {"cantilever-3d": cantilever_3d, "bridge-3d": bridge_3d}[name]()
    n = int(z["iterations"])
    s = GridSimp(p, element="plain", rtol=1e-12, maxiter=4000, coarse_cells=500, precision="fp64")
    hist = s.iterate(10)
    x10 = s.xphys.copy()
    hist += s.iterate(40)
    x50 = s.xphys.copy()
    hist += s.iterate(n - 50)
    c = np.array([st["compliance"] for st in hist])
    rel = np.abs(c - h["compliance"][:n]) / h["compliance"][:n]
    dx10 = float(np.abs(x10 - h["xphys_10"]).max())
    dx50 = float(np.abs(x50 - h["xphys_50"]).max())
    dxg = np.abs(s.xphys - z["x"])
    dcg = abs(c[-1] - float(z["compliance_reported"])) / float(z["compliance_reported"])
    flip = float(((s.xphys > 0.5) != (z["x"] > 0.5)).mean())
    print(f"{name}: compliance rel err its 1-10 {rel[:10].max():.2e}, 1-50 {rel[:50].max():.2e}, 1-{n} "
          f"{rel.max():.2e}; max|dx| it 10 {dx10:.2e}, it 50 {dx50:.2e}; at {n}: compliance vs golden {dcg:.2e}, "
          f"mean|dx| {dxg.mean():.2e}, max|dx| {dxg.max():.2e}, flipped {flip:.4f}; the oracle's multigrid "
          f"loop: compliance {abs(float(h['mg_compliance_reported']) / float(z['compliance_reported']) - 1):.2e}, "
          f"mean|dx| {float(h['mg_mean_dx']):.2e}, flipped {float(h['mg_flipped']):.4f}; MMD LU: mean|dx| "
          f"{float(h['mmd_mean_dx']):.2e}")
    if name == "bridge-3d":
        # this loop does not amplify the solves' noise (the oracle's MMD-LU
        # and multigrid runs land within 3e-8 of the golden design): held to
        # the golden itself, its stop iteration included
        ch = np.array([st["change"] for st in hist])
        assert rel.max() <= 1e-9 and dcg <= 1e-9 and dxg.max() <= 1e-6
        assert int(np.argmax(ch <= 0.01)) == n - 1 and ch[n - 1] <= 0.01
    assert rel[:10].max() <= 1e-9
    assert rel[:50].max() <= 1e-8
    assert dx10 <= 1e-5 and dx50 <= 2e-3
    assert dcg <= 1e-4
    assert dxg.mean() <= 2e-2
    assert flip <= 0.02
    assert abs(s.xphys.mean() - p.volfrac) < 2e-3


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
