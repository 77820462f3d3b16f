"""The benchmark workloads end to end on the device against the oracle.

Fixtures (tests/golden/make_configs.py writes them from the oracle: top3d
OC loops with FP64 multigrid solves at rtol 1e-12, checked against direct
solves; a SuperLU direct solve for the MBB copy):

- 48x24x24 cantilever, 50 OC iterations, 3-level MG with Chebyshev(2) and
  with V(2,2) damped Jacobi (omega 0.6), rtol 1e-8, against the oracle
  iteration by iteration (benchmarks/reference.py:214-257);
- half MBB beam 192x64x64 at its seeded density in FP64, FP32 and mixed
  (FP16-operand MFMA fine operator) with a 5-level MG and FP64 outer CG,
  rtol 1e-8: displacement and compliance error of each form against FP64,
  and the same three forms at 48x16x16 against the direct solve;
- 8-case multiload at 256x128x128, mixed against FP64 (compliance, the
  sensitivity, xphys, volume), and its 64x32x32 copy against the oracle;
- the 32 seeded uneven studies in one concurrent run for 3 iterations, the
  smallest against the oracle.

Each test prints the errors it measured (pytest -s shows them)."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b) / np.abs(b)


@pytest.mark.parametrize("precision,smoother", [("fp64", "chebyshev"), ("fp64", "jacobi"), ("mixed", "chebyshev")])
def test_cantilever48_fifty_iterations(gpu, precision, smoother):
    from loadpath_amd.problems import cantilever_48x24x24
    from loadpath_amd.simp import GridSimp

    z = np.load(GOLDEN / "config1_cantilever48.npz")
    s = GridSimp(cantilever_48x24x24(), precision=precision, smoother=smoother, omega=0.6, degree=2, rtol=1e-8,
                 maxiter=500, element="im", filter="top3d")
    hist = s.iterate(10)
    assert len(s.grid.levels) == 3
    x10 = s.xphys.ravel().copy()
    hist += s.iterate(40)
    c = np.array([h["compliance"] for h in hist])
    rel = _rel(c, z["compliance"])
    dx10 = float(np.abs(x10 - z["xphys_10"]).max())
    dx50 = float(np.abs(s.xphys.ravel() - z["xphys_50"]).max())
    print(f"{precision}/{smoother}: compliance rel err its 1-10 {rel[:10].max():.2e}, 1-50 {rel.max():.2e}; "
          f"max|dxphys| it 10 {dx10:.2e}, it 50 {dx50:.2e}; cg {[h['cg_iters'] for h in hist[:5]]}")
    early = 1e-8 if precision == "fp64" else 1e-6
    assert rel[:10].max() <= early
    # float32 storage of the fixture's field: 3e-8 of the error budget
    assert dx10 <= (1e-6 if precision == "fp64" else 1e-5)
    assert rel.max() <= 1e-5
    assert dx50 <= 1e-3
    assert np.allclose([h["change"] for h in hist[:10]], z["change"][:10], rtol=1e-5, atol=1e-8)
    assert abs(s.xphys.mean() - 0.3) <= 1e-3


def _mbb_forms(nx, ny, nz, coarse_cells):
    from loadpath_amd.mg import GridElasticity
    from loadpath_amd.problems import mbb_192x64x64

    p = mbb_192x64x64(nx=nx, ny=ny, nz=nz)
    mesh = p.mesh()
    fixed, loads = p.dof_vectors(mesh)
    f = loads[0]
    x = p.density[tuple(mesh.elem_ijk.T)]
    e = 1e-9 + x ** 3 * (1 - 1e-9)
    out = {}
    for prec in ("fp64", "fp32", "mixed"):
        g = GridElasticity(mesh, 0.3, fixed, rtol=1e-8, maxiter=2000, coarse_cells=coarse_cells, max_levels=5,
                           dtype=np.float64 if prec == "fp64" else np.float32, precision=prec)
        g.set_moduli(e)
        u = g.solve_many([f])[0]
        out[prec] = dict(u=u, c=float(f @ u), levels=len(g.levels), cg=g.last_iters)
        del g
    return mesh, f, out


def test_mbb192_three_forms_against_fp64(gpu):
    _, _, out = _mbb_forms(192, 64, 64, coarse_cells=200)
    ref = out["fp64"]
    assert all(o["levels"] == 5 for o in out.values())
    for prec in ("fp32", "mixed"):
        o = out[prec]
        du = float(np.linalg.norm(o["u"] - ref["u"]) / np.linalg.norm(ref["u"]))
        dc = abs(o["c"] - ref["c"]) / abs(ref["c"])
        print(f"mbb 192x64x64 {prec}: displacement rel err {du:.2e}, compliance rel err {dc:.2e}, "
              f"cg {o['cg']} (fp64 {ref['cg']})")
        assert du <= 1e-6
        assert dc <= 1e-8


def test_mbb48_three_forms_against_direct(gpu):
    z = np.load(GOLDEN / "config2_mbb48.npz")
    mesh, f, out = _mbb_forms(48, 16, 16, coarse_cells=20)
    for prec, o in out.items():
        # the fixture is in the box mesh's dof order, as is this mesh's
        du = float(np.linalg.norm(o["u"] - z["u"]) / np.linalg.norm(z["u"]))
        dc = abs(o["c"] - float(z["compliance"])) / float(z["compliance"])
        print(f"mbb 48x16x16 {prec}: displacement rel err vs direct {du:.2e}, compliance {dc:.2e}, "
              f"{o['levels']} levels, cg {o['cg']}")
        assert du <= 1e-6
        assert dc <= 1e-8


def _multiload(nx, ny, nz, precision, iterations, levels=0):
    from loadpath_amd.problems import multiload_256x128x128
    from loadpath_amd.simp import GridSimp

    # levels: the hierarchy depth (0: down to the default coarse size)
    kw = dict(max_levels=levels, coarse_cells=0) if levels else {}
    s = GridSimp(multiload_256x128x128(nx, ny, nz), precision=precision, rtol=1e-8, maxiter=500, element="im",
                 filter="top3d", **kw)
    hist = s.iterate(iterations)
    out = dict(c=np.array([h["compliance"] for h in hist]), change=np.array([h["change"] for h in hist]),
               xphys=s.xphys.ravel().copy(), dc=np.asarray(s.dc).ravel().copy(), levels=len(s.grid.levels),
               cg=[h["cg_iters"] for h in hist])
    del s
    return out


def test_multiload256_mixed_against_fp64(gpu):
    """The headline workload, 3 iterations of each path from rho = 0.15:
    compliance <= 1e-6, sensitivity <= 1e-4 (max abs error over max |dc|),
    xphys <= 1e-6, the volume on target."""
    m = _multiload(256, 128, 128, "mixed", 3, levels=6)
    f = _multiload(256, 128, 128, "fp64", 3, levels=6)
    dc_c = float(_rel(m["c"], f["c"]).max())
    dc_s = float(np.abs(m["dc"] - f["dc"]).max() / np.abs(f["dc"]).max())
    dx = float(np.abs(m["xphys"] - f["xphys"]).max())
    print(f"multiload 256x128x128: compliance {dc_c:.2e}, sensitivity {dc_s:.2e}, xphys {dx:.2e}, "
          f"levels {m['levels']}, cg mixed {m['cg']} fp64 {f['cg']}")
    assert m["levels"] == f["levels"] == 6
    assert dc_c <= 1e-6
    assert dc_s <= 1e-4
    assert dx <= 1e-6
    assert abs(m["xphys"].mean() - 0.15) <= 1e-3 and abs(f["xphys"].mean() - 0.15) <= 1e-3


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multiload64_against_oracle(gpu, precision):
    z = np.load(GOLDEN / "config4_multiload64.npz")
    r = _multiload(64, 32, 32, precision, 5)
    rel = _rel(r["c"], z["compliance"])
    dx = float(np.abs(r["xphys"] - z["xphys_5"]).max())
    print(f"multiload 64x32x32 {precision}: compliance rel err {rel.max():.2e}, max|dxphys| {dx:.2e}, cg {r['cg']}")
    assert rel.max() <= (1e-8 if precision == "fp64" else 1e-6)
    assert dx <= (1e-6 if precision == "fp64" else 1e-5)
    assert np.allclose(r["change"], z["change"], rtol=1e-5, atol=1e-8)


def test_thirty_two_studies(gpu):
    from loadpath_amd.batch import run_studies
    from loadpath_amd.problems import batched_studies

    z = np.load(GOLDEN / "config5_smallest.npz")
    ps = batched_studies(32, seed=5)
    r = run_studies(ps, precision="mixed", max_iterations=3, concurrent=True, autotune=False)
    assert len(r["studies"]) == 32
    assert all(st["iterations"] == 3 for st in r["studies"])
    assert all(np.isfinite(st["compliance"]) and st["compliance"] > 0 for st in r["studies"])
    k = int(z["index"])
    st = r["studies"][k]
    assert st["name"] == str(z["name"])
    rel = _rel([h["compliance"] for h in st["history"]], z["compliance"])
    dx = float(np.abs(np.asarray(st["xphys"]).ravel() - z["xphys_3"]).max())
    print(f"32 studies: {r['studies_per_s']:.2f} studies/s (3 iterations each), p50 {r['p50_s']:.3f} s, "
          f"p99 {r['p99_s']:.3f} s; smallest {st['name']}: compliance rel err {rel.max():.2e}, "
          f"max|dxphys| {dx:.2e}")
    assert rel.max() <= 1e-6
    assert dx <= 1e-5
