"""Concurrent studies and the graph-replayed V-cycle give the same numbers
as one-at-a-time runs and kernel-by-kernel V-cycles (bitwise: every study
owns its handle and stream, and reductions run in a fixed order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed"])
def test_graph_vcycle_matches_launches(gpu, precision):
    from loadpath_amd import _lib
    from loadpath_amd.mg import GridElasticity
    from tests.conftest import cantilever_bar

    mesh, fixed, f, e = cantilever_bar(24, 12, 12, seed=9)
    g = GridElasticity(mesh, 0.3, fixed, precision=precision, rtol=1e-10, coarse_cells=100, maxiter=2000)
    g.set_moduli(e)
    fs = [f, np.roll(f, 3)]
    ug = g.solve_many(fs)
    it_g = g.last_iters
    ug2 = g.solve_many(fs)        # replays the same graph again
    assert _lib.lib().lp_tune(g._h, 3, 0) == 1
    ul = g.solve_many(fs)
    assert g.last_iters == it_g
    for a, b, c in zip(ug, ug2, ul):
        assert np.array_equal(a, b) and np.array_equal(a, c)


def test_concurrent_studies_match_sequential(gpu):
    from loadpath_amd.batch import run_studies
    from loadpath_amd.problems import batched_studies

    ps = batched_studies(count=4, seed=5)
    a = run_studies(ps, precision="mixed", max_iterations=3, concurrent=True, autotune=False)
    b = run_studies(ps, precision="mixed", max_iterations=3, concurrent=False, autotune=False)
    for sa, sb in zip(a["studies"], b["studies"]):
        info = {k: (sa[k], sb[k]) for k in ("name", "strict", "fp32_repeats", "strict_switches")}
        info["cg"] = ([h["cg_iters"] for h in sa["history"]], [h["cg_iters"] for h in sb["history"]])
        info["oc"] = ([h["oc_evals"] for h in sa["history"]], [h["oc_evals"] for h in sb["history"]])
        assert sa["iterations"] == sb["iterations"], info
        assert [h["compliance"] for h in sa["history"]] == [h["compliance"] for h in sb["history"]], info
        assert np.array_equal(sa["xphys"], sb["xphys"]), info


def test_concurrent_production_loops_match_sequential(gpu):
    """Four production-loop studies at once (one thread and stream each)
    give what they give one after another, bit for bit."""
    from loadpath_amd.batch import run_studies
    from loadpath_amd.problems import batched_studies

    ps = batched_studies(count=4, seed=11)
    a = run_studies(ps, precision="mixed", max_iterations=6, concurrent=True, autotune=False, loop=True)
    b = run_studies(ps, precision="mixed", max_iterations=6, concurrent=False, autotune=False, loop=True)
    for sa, sb in zip(a["studies"], b["studies"]):
        assert sa["iterations"] == sb["iterations"] == 6
        assert [h["compliance"] for h in sa["history"]] == [h["compliance"] for h in sb["history"]]
        assert [h["beta"] for h in sa["history"]] == [h["beta"] for h in sb["history"]]
        assert np.array_equal(sa["xphys"], sb["xphys"])


def test_autotune_keeps_results(gpu):
    from loadpath_amd.mg import GridElasticity
    from tests.conftest import cantilever_bar

    mesh, fixed, f, e = cantilever_bar(32, 16, 16, seed=10)
    g = GridElasticity(mesh, 0.3, fixed, precision="mixed", rtol=1e-10, coarse_cells=100, maxiter=2000)
    g.set_moduli(e)
    u0 = g.solve_many([f])[0]
    xc = g.autotune(1)
    assert xc in (4, 8, 16, 32)
    u1 = g.solve_many([f])[0]
    assert np.linalg.norm(u1 - u0) <= 1e-8 * np.linalg.norm(u0)
