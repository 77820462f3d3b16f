"""Writes the oracle's runs of the benchmark workloads as fixtures for the
GPU tests (tests/test_gpu_configs.py). Inputs are the seeded problems of
loadpath_amd.problems; every output comes from the oracle (FP64, CPU).

- config1_cantilever48.npz: the 48x24x24 cantilever (x = 0 clamped, unit
  load down at node (48, 0, 12), vf 0.3, rmin 1.5), 50 top3d OC iterations
  from rho = 0.3 (benchmarks/reference.py:214-257): compliance and change
  per iteration, xphys after iterations 10 and 50 (float32).
- config2_mbb48.npz: the half MBB beam scaled to 48x16x16 with its seeded
  density U(0.05, 1): the assembled direct solve (SuperLU, MMD ordering on
  K^T + K) of the free dofs, compliance and u (float64, HexMesh dof order).
- config4_multiload64.npz: the 8-case multiload problem at 64x32x32 (vf
  0.15, rmin 2.5), 5 top3d OC iterations, objective the sum of the eight
  compliances: per-iteration compliance and change, xphys after 5 (float32).
- config5_smallest.npz: the smallest of the 32 seeded studies
  (batched_studies(32, seed 5)), 3 top3d OC iterations: compliance, change,
  xphys (float32).

The OC loops solve with the oracle's multigrid PCG (oracle.mg.GridOracle,
V-cycle in FP64) at rtol 1e-12, warm-started from the previous iteration:
the 48x24x24 direct solve takes minutes per iteration here. The script
checks each loop's first compliance against a direct solve where that is
affordable (configs 1 and 5), to 1e-10.

    python tests/golden/make_configs.py [config1 config2 config4 config5]
"""
import sys
import time
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from loadpath_amd import problems as P  # noqa: E402
from oracle import simp as os_  # noqa: E402
from oracle.mg import GridOracle  # noqa: E402

OUT = Path(__file__).resolve().parent


def oracle_problem(p):
    """BoxProblem -> the oracle's Problem (vector loads split by component)."""
    cases = []
    for case in p.cases:
        terms = []
        for t in case:
            if len(t) == 3:
                terms.append(t)
            else:
                nd, vec = t
                terms.extend((nd, c, float(vec[i])) for i, c in enumerate("xyz") if vec[i] != 0.0)
        cases.append(terms)
    return os_.Problem(p.name, p.dims, p.volfrac, p.rmin, list(p.fixed), cases)


def direct(ev, e, f):
    data = (np.asarray(e, float)[:, None, None] * ev.ke[None]).ravel()
    n = ev.mesh.n_dof
    k = sp.csr_matrix((data, (ev._rows, ev._cols)), shape=(n, n))[ev.free][:, ev.free].tocsc()
    u = np.zeros(n)
    u[ev.free] = spla.spsolve(k, f[ev.free], permc_spec="MMD_AT_PLUS_A")
    return u


def mg_loop(p, iterations, keep=(), check_direct=False):
    """oracle.simp.optimize_top3d with GridOracle solves; one run per kept
    snapshot (the loop is deterministic, so a shorter run is a prefix)."""
    q = oracle_problem(p)
    ev = os_.Evaluator(q, "im")
    checked = []

    def run(n):
        g = GridOracle(ev.mesh, 0.3, ev.fixed, rtol=1e-12, maxiter=4000, vdtype=np.float64)
        prev = [None]

        def solve(e):
            g.set_moduli(e)
            us = g.solve_many(ev.fs, prev[0])
            prev[0] = us
            if check_direct and not checked:
                cm = sum(float(f @ u) for f, u in zip(ev.fs, us))
                cd = sum(float(f @ direct(ev, e, f)) for f in ev.fs)
                assert abs(cm - cd) <= 1e-10 * abs(cd), (cm, cd)
                checked.append(cd)
            return us

        return os_.optimize_top3d(q, solve=solve, iterations=n, element="im")

    snaps = {f"xphys_{k}": run(k)["x"].ravel().astype(np.float32) for k in keep}
    r = run(iterations)
    snaps[f"xphys_{iterations}"] = r["x"].ravel().astype(np.float32)
    hist = r["history"]
    return np.array([c for _, c, _ in hist]), np.array([ch for _, _, ch in hist]), snaps


def config1():
    c, ch, snaps = mg_loop(P.cantilever_48x24x24(), 50, keep=(10,), check_direct=True)
    np.savez_compressed(OUT / "config1_cantilever48.npz", compliance=c, change=ch, **snaps)


def config2():
    p = P.mbb_192x64x64(nx=48, ny=16, nz=16)
    q = oracle_problem(p)
    ev = os_.Evaluator(q, "im")
    e = os_.simp_modulus(p.density.ravel(), 3.0)
    u = direct(ev, e, ev.f)
    np.savez_compressed(OUT / "config2_mbb48.npz", compliance=float(ev.f @ u), u=u)


def config4():
    c, ch, snaps = mg_loop(P.multiload_256x128x128(64, 32, 32), 5)
    np.savez_compressed(OUT / "config4_multiload64.npz", compliance=c, change=ch, **snaps)


def config5():
    ps = P.batched_studies(32, seed=5)
    k = int(np.argmin([np.prod(p.dims) for p in ps]))
    c, ch, snaps = mg_loop(ps[k], 3, check_direct=True)
    np.savez_compressed(OUT / "config5_smallest.npz", index=k, name=ps[k].name, compliance=c, change=ch, **snaps)


if __name__ == "__main__":
    todo = sys.argv[1:] or ["config1", "config2", "config4", "config5"]
    for name in todo:
        t0 = time.time()
        globals()[name]()
        print(name, f"{time.time() - t0:.1f} s", flush=True)
