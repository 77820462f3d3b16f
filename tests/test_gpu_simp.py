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
    from loadpath_amd.problems import cantilever_3d
    from loadpath_amd.simp import optimize_top3d

    z = np.load(GOLDEN / "cantilever-3d.top3d.npz")
    r = optimize_top3d(cantilever_3d(), element="plain", rtol=1e-11, maxiter=3000, coarse_cells=500)
    assert abs(r["iterations"] - int(z["iterations"])) <= 2, r["iterations"]
    assert r["compliance_reported"] == pytest.approx(float(z["compliance_reported"]), rel=1e-5)
    assert np.abs(r["x"] - z["x"]).max() < 5e-3
    assert abs(r["x"].mean() - 0.3) < 2e-3
