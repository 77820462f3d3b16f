"""The production loop on the device (lp_simp_set_loop, loadpath_amd.simp.
optimize) against the oracle's restatement of simp._optimize_run
(oracle/loop.py), and the properties the reference's tests assert of
simp.optimize (tests/test_loadpath.py:106-118, 1190-1223): volume on
target to rounding, the step limit never above opts.move, bit-identical
reruns, compliance falling at the sharpest stage."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem():
    from loadpath_amd.problems import BoxProblem, nodes

    nx, ny, nz = 16, 4, 6
    return BoxProblem("cantilever-2case", (nx, ny, nz), 0.3, 1.5,
                      [(nodes(0, range(ny + 1), range(nz + 1)), "xyz")],
                      [[(nodes(nx, range(ny + 1), 0), "z", -1.0)],
                       [(nodes(nx, range(ny + 1), nz), "y", 0.4)]])


def _preserve(dims):
    pres = np.zeros(dims, bool)
    pres[-1, :, 0] = True            # a pad under the first load
    return pres


def _oracle_problem(p):
    from oracle import simp as os_

    return os_.Problem(p.name, p.dims, p.volfrac, p.rmin, list(p.fixed), [list(c) for c in p.cases])


@pytest.mark.parametrize("variant", ["measured", "fixed-step"])
def test_loop_matches_oracle(gpu, variant):
    """FP64 device loop (PCG to 1e-11) against the oracle's direct solves.
    measured: 25 iterations, two beta doublings, the measured step, weighted
    cases and a preserved pad. fixed-step: the fixed limits (move, move / 2
    past beta 4), a per-level schedule, a thinner erosion and a void block."""
    from loadpath_amd.simp import SimpOptions, optimize
    from oracle.loop import LoopOptions, optimize_loop

    p = _problem()
    pres = _preserve(p.dims)
    w = [1.0, 0.5]
    void = None
    kw = {}
    if variant == "fixed-step":
        void = np.zeros(p.dims, bool)
        void[6:10, :, 2:4] = True
        kw = dict(adaptive_move=False, beta_schedule=(3, 4), erode=0.1)
    n_it = 25
    opts = SimpOptions(volume_fraction=0.3, filter_radius=1.5, max_iterations=n_it, solve_rtol=1e-11, **kw)
    r = optimize(p, opts, preserve=pres, void=void, case_weights=w, precision="fp64", element="im", maxiter=3000,
                 coarse_cells=50, delivered=False)
    ro = optimize_loop(_oracle_problem(p), LoopOptions(volume_fraction=0.3, filter_radius=1.5, max_iterations=n_it,
                                                       **kw),
                       preserve=pres.ravel(), weights=w, element="im", void=None if void is None else void.ravel())
    assert len(r.history) == len(ro["history"])
    for a, b in zip(r.history, ro["history"]):
        print(a["iteration"], a["compliance"], b["compliance"], a["beta"], a.get("move"), b.get("move"),
              a["change"], b["change"], a["volumeSolves"], b["volumeSolves"])
        assert a["beta"] == b["beta"]
        assert a["compliance"] == pytest.approx(b["compliance"], rel=1e-7)
        assert a["volume"] == pytest.approx(b["volume"], abs=1e-9)
        assert a.get("move", 0.0) == pytest.approx(b.get("move", 0.0), abs=1e-7)
        assert a["change"] == pytest.approx(b["change"], abs=1e-6)
        assert a["volHeld"] == b["volHeld"]
        assert a["lambda"] == pytest.approx(b["lambda"], rel=1e-5)
    assert r.converged == ro["converged"]
    assert np.abs(r.x - ro["x"]).max() < 1e-6
    assert np.abs(r.xphys - ro["xphys"]).max() < 1e-5
    if void is not None:
        assert (r.x[void] == 0).all() and (r.xphys[void] == 0).all()


def test_loop_properties_mixed_precision(gpu):
    """The mixed-precision path on a larger cantilever to its stop: volume on
    target where the step held it, the limit at most opts.move, the stage
    at beta_max settling, and a rerun bit-identical."""
    from loadpath_amd.problems import cantilever_3d
    from loadpath_amd.simp import SimpOptions, optimize

    p = cantilever_3d(48, 12, 16, volfrac=0.2, rmin=2.0)
    opts = SimpOptions(volume_fraction=0.2, filter_radius=2.0, max_iterations=90)
    runs = [optimize(p, opts, precision="mixed", coarse_cells=200, delivered=False) for _ in range(2)]
    r = runs[0]
    h = r.history
    print(len(h), r.converged, [round(x["compliance"], 3) for x in h[-6:]])
    assert r.converged
    assert all(x["move"] <= 0.2 + 1e-12 for x in h)
    assert [x["beta"] for x in h] == sorted(x["beta"] for x in h)
    pairs = [(a, b) for a, b in zip(h, h[1:]) if a["volHeld"] and a["beta"] == b["beta"]]
    assert len(pairs) > 20
    assert max(abs(b["volume"] - 0.2) for a, b in pairs) < 1e-10
    last = [x for x in h if x["beta"] >= 32][3:]
    assert len(last) >= 3
    rise = max(b["compliance"] / a["compliance"] - 1 for a, b in zip(last, last[1:]))
    assert rise < 1e-3, rise
    assert last[-1]["compliance"] < last[0]["compliance"]
    grey = ((r.xphys > 0.1) & (r.xphys < 0.9)).mean()
    assert grey < 0.1, grey
    assert np.array_equal(runs[0].x, runs[1].x)
    assert [x["compliance"] for x in runs[0].history] == [x["compliance"] for x in runs[1].history]


def test_optimize_delivers(gpu):
    """optimize() hands over the delivered design by default: the opening
    of the eroded design, solid where preserved, its compliance solved."""
    from loadpath_amd.simp import SimpOptions, optimize

    p = _problem()
    pres = _preserve(p.dims)
    opts = SimpOptions(volume_fraction=0.3, filter_radius=1.5, max_iterations=40)
    r = optimize(p, opts, preserve=pres, precision="mixed", coarse_cells=50)
    assert r.xphys.shape == p.dims and (r.xphys[pres] == 1.0).all()
    assert len(r.case_compliance) == 2 and r.compliance == pytest.approx(sum(r.case_compliance), rel=1e-12)
    assert 0.0 < r.volume_fraction < 1.0


def test_loop_rejects_bad_options(gpu):
    from loadpath_amd.problems import cantilever_3d
    from loadpath_amd.simp import GridSimp, SimpOptions

    s = GridSimp(cantilever_3d(8, 2, 4), coarse_cells=20)
    with pytest.raises(Exception, match="bad options"):
        s.set_loop(SimpOptions(volume_fraction=0.3, filter_radius=1.5, eta=1.2))
    with pytest.raises(Exception, match="every cell is preserved"):
        s.set_loop(SimpOptions(volume_fraction=0.3, filter_radius=1.5), preserve=np.ones((8, 2, 4), bool))
    with pytest.raises(ValueError):
        s.set_loop(SimpOptions(volume_fraction=0.3, filter_radius=1.5), case_weights=[1.0, 2.0])


def test_delivery_matches_oracle(gpu):
    """The delivered design (opening with cores kept, solve at 1e-6, dead
    members pruned: simp.py:2452-2548) from the same final design on the
    device and in the oracle."""
    from loadpath_amd.simp import GridSimp, SimpOptions, deliver
    from oracle.loop import Loop, LoopOptions, deliver as o_deliver, optimize_loop

    p = _problem()
    pres = _preserve(p.dims)
    w = [1.0, 0.5]
    lo = LoopOptions(volume_fraction=0.3, filter_radius=1.5, max_iterations=70)
    ro = optimize_loop(_oracle_problem(p), lo, preserve=pres.ravel(), weights=w, element="im")
    want = o_deliver(Loop(_oracle_problem(p), lo, preserve=pres.ravel(), weights=w, element="im"),
                     ro["x"].ravel(), ro["beta"])
    s = GridSimp(p, precision="fp64", element="im", rtol=1e-10, maxiter=3000, coarse_cells=50, filter="padded",
                 volfrac=0.3, rmin=1.5)
    opts = SimpOptions(volume_fraction=0.3, filter_radius=1.5)
    s.set_loop(opts, preserve=pres, case_weights=w)
    got = deliver(s, ro["x"], ro["beta"], opts, preserve=pres)
    print(ro["iterations"], ro["beta"], want["compliance"], got["compliance"],
          int((want["xphys"] > 0.5).sum()), int((got["xphys"] > 0.5).sum()))
    assert np.abs(got["xphys"] - want["xphys"].reshape(p.dims)).max() < 1e-6
    assert got["compliance"] == pytest.approx(want["compliance"], rel=1e-6)
    assert got["volume_fraction"] == pytest.approx(want["volume_fraction"], abs=1e-9)


def test_loop_ties_match_oracle(gpu):
    """Mirror planes and extrusion columns in the production loop (simp.py:
    2223-2226, 2288-2304): a y mirror (4 cells) and z columns (6 cells) on a
    cantilever loaded at one y edge, so neither tie holds by itself. The
    device loop's iterates equal the oracle's, and every design satisfies
    the ties."""
    from loadpath_amd.problems import BoxProblem, nodes
    from loadpath_amd.simp import SimpOptions, optimize
    from oracle.loop import LoopOptions, optimize_loop

    nx, ny, nz = 16, 4, 6
    # loads on one y edge only: the untied optimum is not symmetric in y
    p = BoxProblem("cantilever-edge", (nx, ny, nz), 0.3, 1.5, [(nodes(0, range(ny + 1), range(nz + 1)), "xyz")],
                   [[(nodes(nx, 0, range(nz + 1)), "z", -1.0 / (nz + 1))], [(nodes(nx, 1, 0), "y", 0.4)]])
    n_it = 16
    opts = SimpOptions(volume_fraction=0.3, filter_radius=1.5, max_iterations=n_it, solve_rtol=1e-11, extrusion=2)
    r = optimize(p, opts, mirror="y", precision="fp64", element="im", maxiter=3000, coarse_cells=50,
                 delivered=False)
    ro = optimize_loop(_oracle_problem(p), LoopOptions(volume_fraction=0.3, filter_radius=1.5, max_iterations=n_it,
                                                       mirror="y", extrusion=2), element="im")
    assert len(r.history) == len(ro["history"])
    for a, b in zip(r.history, ro["history"]):
        assert a["beta"] == b["beta"]
        assert a["compliance"] == pytest.approx(b["compliance"], rel=1e-7)
        assert a["volume"] == pytest.approx(b["volume"], abs=1e-9)
    x = r.x.reshape(p.dims)
    assert np.array_equal(x, x[:, ::-1, :])
    assert np.array_equal(x, np.broadcast_to(x[:, :, :1], x.shape))
    assert np.abs(r.x - ro["x"]).max() < 1e-6
    # untied, the same problem is not symmetric in y
    r0 = optimize(p, SimpOptions(volume_fraction=0.3, filter_radius=1.5, max_iterations=4, solve_rtol=1e-11),
                  precision="fp64", element="im", maxiter=3000, coarse_cells=50, delivered=False)
    assert np.abs(r0.x - r0.x[:, ::-1, :]).max() > 1e-4


def _island():
    """Preserved 2 x 2 x 1 block ringed by void cells: alive (preserved) but
    joined to no support once the loop sharpens."""
    dims = (16, 4, 6)
    pres = _preserve(dims)
    pres[12:14, 1:3, 4] = True
    void = np.zeros(dims, bool)
    void[11:15, :, 3:6] = True
    void[12:14, 1:3, 4] = False
    return pres, void


@pytest.mark.parametrize("loaded_island", [False, True])
def test_loop_reduction_matches_oracle(gpu, loaded_island):
    """The multigrid path's live-cell reduction (simp.py:795-826): past beta
    8 fewer than 80 % of the cells are alive and the island is dropped
    (modulus 0), so compliance and designs follow the oracle's reduced
    direct solves. With a load on the island the reduced system cannot hold
    it and every such iteration solves the whole grid (both sides)."""
    from loadpath_amd.problems import BoxProblem, nodes
    from loadpath_amd.simp import GridSimp, SimpOptions, optimize
    from oracle.loop import Loop, LoopOptions, optimize_loop

    p = _problem()
    pres, void = _island()
    if loaded_island:
        cases = [list(p.cases[0]), list(p.cases[1]) + [(nodes(13, 2, 5), "x", 1e-3)]]
        p = BoxProblem(p.name, p.dims, p.volfrac, p.rmin, list(p.fixed), cases)
    n_it = 24
    sched = (3, 3, 3, 3)
    opts = SimpOptions(volume_fraction=0.3, filter_radius=1.5, max_iterations=n_it, solve_rtol=1e-11,
                       beta_schedule=sched)
    r = optimize(p, opts, preserve=pres, void=void, precision="fp64", element="im", maxiter=3000, coarse_cells=50,
                 delivered=False)
    lo = LoopOptions(volume_fraction=0.3, filter_radius=1.5, max_iterations=n_it, beta_schedule=sched)
    op = _oracle_problem(p)
    ro = optimize_loop(op, lo, preserve=pres.ravel(), element="im", void=void.ravel())
    assert len(r.history) == len(ro["history"])
    for a, b in zip(r.history, ro["history"]):
        assert a["beta"] == b["beta"]
        assert a["compliance"] == pytest.approx(b["compliance"], rel=1e-7), a["iteration"]
        assert a["volume"] == pytest.approx(b["volume"], abs=1e-9)
    assert np.abs(r.x - ro["x"]).max() < 1e-6
    # the reduction did act (counted on a rerun's device loop) and, with the
    # loaded island, sent those iterations to the whole grid
    s = GridSimp(p, precision="fp64", filter="padded", volfrac=0.3, rmin=1.5, rtol=1e-11, maxiter=3000,
                 coarse_cells=50, element="im", active=~void)
    s.set_loop(opts, preserve=pres)
    s.iterate(n_it)
    info = s.loop_info()
    if loaded_island:
        assert info["reduced"] == 0 and info["full_for_loads"] > 0, info
    else:
        assert info["reduced"] > 0 and info["full_for_loads"] == 0, info
    # the oracle's own count of reduced solves agrees
    L = Loop(op, lo, preserve=pres.ravel(), element="im", void=void.ravel())
    assert L.support_touch.any() and L.anchor[pres.ravel()].all()
