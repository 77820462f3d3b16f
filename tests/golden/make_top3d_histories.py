"""Writes the oracle's iteration histories of the reference's two golden
top3d runs (benchmarks/reference/{cantilever-3d,bridge-3d}.top3d.npz, kept
here as tests/golden/*.top3d.npz) as fixtures for tests/test_gpu_simp.py.

For each problem (plain element, the reference's evaluator):

1. the oracle's top3d loop with the reference evaluator's own direct solve
   (scipy spsolve, default ordering) — the script asserts that it lands on
   the golden (same stop iteration, compliance to 1e-12, the stored float32
   design to 1e-6) and keeps its per-iteration compliance and change and
   the design after iterations 10 and 50 (float32);
2. the same loop with an MMD-ordered LU (a different rounding of the same
   solves) for the golden's iteration count: its per-iteration compliance
   and its final design's distance from the golden (max and mean |dx|,
   share of cells on the other side of 0.5). That is how far rounding alone
   moves this loop;
3. the same loop with the oracle's multigrid PCG (oracle.mg.GridOracle,
   FP64 V-cycle, rtol 1e-12, warm-started): the same distances (mg_*). A
   solve converged to 1e-12 perturbs every iteration a thousand times more
   than a change of LU ordering does, and the loop amplifies it; this is
   the yardstick for the device loop, which solves the same way.

    python tests/golden/make_top3d_histories.py [cantilever-3d bridge-3d]
"""
import sys
import time
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import simp as S  # noqa: E402

HERE = Path(__file__).resolve().parent
PROBLEMS = {"cantilever-3d": S.cantilever_3d, "bridge-3d": S.bridge_3d}


def mmd_solve(ev):
    def solve(e):
        data = (np.asarray(e, float)[:, None, None] * ev.ke[None]).ravel()
        n = ev.mesh.n_dof
        k = sp.csr_matrix((data, (ev._rows, ev._cols)), shape=(n, n))[ev.free][:, ev.free].tocsc()
        out = []
        for f in ev.fs:
            u = np.zeros(n)
            u[ev.free] = spla.spsolve(k, f[ev.free], permc_spec="MMD_AT_PLUS_A")
            out.append(u)
        return out
    return solve


def run(name):
    z = np.load(HERE / f"{name}.top3d.npz")
    p = PROBLEMS[name]()
    n_gold = int(z["iterations"])
    snaps = {}
    hist = []
    def keep(it, xphys):
        if it in (10, 50):
            snaps[f"xphys_{it}"] = xphys.reshape(p.dims).astype(np.float32)

    r = S.optimize_top3d(p, element="plain", on_iter=lambda it, c, ch: hist.append((c, ch)), on_design=keep)
    assert r["iterations"] == n_gold, (r["iterations"], n_gold)
    assert abs(r["compliance_reported"] - float(z["compliance_reported"])) <= 1e-12 * float(z["compliance_reported"])
    assert np.abs(r["x"] - z["x"]).max() < 1e-6

    hm = []
    ev = S.Evaluator(p, "plain")
    rm = S.optimize_top3d(p, solve=mmd_solve(ev), iterations=n_gold, element="plain",
                          on_iter=lambda it, c, ch: hm.append(c))
    dx = np.abs(rm["x"] - z["x"])

    from oracle.mg import GridOracle

    g = GridOracle(ev.mesh, 0.3, ev.fixed, rtol=1e-12, maxiter=4000, coarse_cells=500, vdtype=np.float64,
                   element="plain")
    prev = [None]

    def mg_solve(e):
        g.set_moduli(e)
        us = g.solve_many(ev.fs, prev[0])
        prev[0] = us
        return us

    rg = S.optimize_top3d(p, solve=mg_solve, iterations=n_gold, element="plain")
    dg = np.abs(rg["x"] - z["x"])
    mg = dict(mg_max_dx=float(dg.max()), mg_mean_dx=float(dg.mean()),
              mg_flipped=float(((rg["x"] > 0.5) != (z["x"] > 0.5)).mean()),
              mg_compliance_reported=float(rg["compliance_reported"]))
    np.savez_compressed(HERE / f"{name}.history.npz", compliance=np.array([c for c, _ in hist]), **mg,
                        change=np.array([ch for _, ch in hist]), mmd_compliance=np.array(hm),
                        mmd_max_dx=float(dx.max()), mmd_mean_dx=float(dx.mean()),
                        mmd_flipped=float(((rm["x"] > 0.5) != (z["x"] > 0.5)).mean()),
                        mmd_compliance_reported=float(rm["compliance_reported"]), **snaps)
    c = np.array([c for c, _ in hist])
    rel = np.abs(np.array(hm) - c[:len(hm)]) / c[:len(hm)]
    print(name, "iterations", n_gold, "mmd vs default compliance rel diff: its 1-50", f"{rel[:50].max():.2e}",
          "all", f"{rel.max():.2e}", "final design max|dx|", f"{dx.max():.3e}", "mean", f"{dx.mean():.3e}",
          "| multigrid: max|dx|", f"{dg.max():.3e}", "mean", f"{dg.mean():.3e}", "flipped", f"{mg['mg_flipped']:.4f}",
          "compliance", f"{abs(rg['compliance_reported'] / float(z['compliance_reported']) - 1):.2e}")


if __name__ == "__main__":
    for name in sys.argv[1:] or list(PROBLEMS):
        t0 = time.time()
        run(name)
        print(name, f"{time.time() - t0:.1f} s", flush=True)
