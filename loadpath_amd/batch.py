"""Many independent SIMP studies on one GPU at once.

Each study is a device GridSimp with its own HIP stream; one host thread per
study drives its loop (the library releases the GIL inside every call and a
handle is used by one thread), so the studies' kernels interleave on the
GPU. Before the run each distinct grid shape is autotuned once (the x-chunk
of the V-cycle's fine operator) and every PCG replays its V-cycle from a
captured hipGraph, which is what makes small grids cheap: their kernels are
short and launch latency would dominate.

    res = run_studies(problems, precision="mixed", max_iterations=40)
    res["studies_per_s"], res["p50_s"], res["p99_s"]

A study ends at the top3d stop (change <= 0.01) or at max_iterations; with
``loop=True`` each study runs the production loop instead (simp.optimize's:
eroded analysis, projection continuation, measured step, multiplier root;
loadpath_amd.simp.optimize) and ends at its stop test.
"""
from __future__ import annotations

import threading
import time

import numpy as np

__all__ = ["run_studies"]


def _one(problem, precision, max_iterations, tol, graph, autotune, out, k, start_barrier, loop=False, progress=None):
    try:
        _run_one(problem, precision, max_iterations, tol, graph, autotune, out, k, start_barrier, loop)
    except Exception as e:  # reported by run_studies
        if start_barrier is not None:
            start_barrier.abort()
        out[k] = dict(name=problem.name, error=f"{type(e).__name__}: {e}")
    if progress is not None:
        progress(k, out[k])


def _run_one(problem, precision, max_iterations, tol, graph, autotune, out, k, start_barrier, loop=False):
    from . import _lib
    from .simp import GridSimp, SimpOptions

    if loop:
        s = GridSimp(problem, precision=precision, rtol=1e-8, maxiter=500, coarse_cells=500, filter="padded")
        s.set_loop(SimpOptions(volume_fraction=problem.volfrac, filter_radius=problem.rmin,
                               max_iterations=max_iterations))
    else:
        s = GridSimp(problem, precision=precision, rtol=1e-8, maxiter=500, coarse_cells=500)
    lib = _lib.lib()
    lib.lp_tune(s.grid._h, 3, 1 if graph else 0)
    if start_barrier is not None:
        start_barrier.wait()
    t0 = time.perf_counter()
    hist = []
    for _ in range(max_iterations):
        st = s.iterate(1)[0]
        hist.append(st)
        if autotune and len(hist) == 1:
            s.grid.autotune(s.nrhs)
        if (st["converged"] if loop else st["change"] <= tol):
            break
    out[k] = dict(name=problem.name, seconds=time.perf_counter() - t0, iterations=len(hist),
                  compliance=hist[-1]["compliance"], history=hist, xphys=s.xphys, strict=bool(s.grid.strict),
                  fp32_repeats=int(lib.lp_tune(s.grid._h, 6, 0)), strict_switches=int(lib.lp_tune(s.grid._h, 11, 0)))


def run_studies(problems, precision="mixed", max_iterations=40, tol=0.01, concurrent=True, graph=True,
                autotune=True, loop=False, progress=None):
    """Run the studies concurrently (one host thread and stream each) or one
    after another. Returns per-study results and the batch's studies/s and
    p50 / p99 time to converge (seconds per study, from its first
    iteration). progress(k, result) is called as each study ends."""
    n = len(problems)
    out = [None] * n
    t0 = time.perf_counter()
    if concurrent:
        barrier = threading.Barrier(n)
        threads = [threading.Thread(target=_one, args=(p, precision, max_iterations, tol, graph, autotune, out, k,
                                                       barrier, loop, progress)) for k, p in enumerate(problems)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    else:
        for k, p in enumerate(problems):
            _one(p, precision, max_iterations, tol, graph, autotune, out, k, None, loop, progress)
    wall = time.perf_counter() - t0
    bad = [o for o in out if o is None or "error" in o]
    if bad:
        raise RuntimeError(f"{len(bad)} studies failed: {bad[0]}")
    secs = np.array([o["seconds"] for o in out])
    return dict(studies=out, wall_s=wall, studies_per_s=n / wall, p50_s=float(np.percentile(secs, 50)),
                p99_s=float(np.percentile(secs, 99)))
