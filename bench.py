#!/usr/bin/env python3
"""Headline benchmark: SIMP iterations per second on the 256x128x128 hex8
multi-load workload (8 load cases, bottom face clamped, vf 0.15, rmin 2.5),
mixed-precision MG-PCG vs plain FP64 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

One step = one SIMP iteration on the device: moduli, hierarchy rebuild,
8 right-hand sides solved in one batched PCG (rtol 1e-8), sensitivities,
density filter, OC update with its volume bisection. Prints ONE JSON line.
`value` is the mixed path's iterations/s; the same iterations are run in
plain FP64 from the same start, and the largest relative difference in
compliance between the two is reported (`compliance_rel_err`).

`roofline` is the dominant kernel's (the V-cycle's fine-level operator)
algorithmic bytes / device time, summed over every launch in the timed
region (HIP events on the grid's stream). `cpu_baseline` times the FP64
numpy oracle on a bounded sample (see its "sample" text).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

_CORES = len(os.sched_getaffinity(0))
for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = str(_CORES)

HBM_PEAK_GBS = 8000.0
FP16_MFMA_PEAK_TFLOPS = 2500.0   # dense FP16 MFMA, MI355X
# HBM bytes per launch of the fine-level FP16-MFMA operator (8 right-hand
# sides, mean over the 52 V-cycle launches of one mixed SIMP iteration, 13 of
# each of its four epilogue kinds): FETCH_SIZE x 2 + WRITE_SIZE x 1, the
# factors from a 1 GiB copy in the same run; from
# profiles/counters_multiload_256x128x128.txt
FINE_MFMA_TRAFFIC_BYTES = 2.0775e9
WORKLOAD = dict(nx=256, ny=128, nz=128, cases=8)
FINE_MFMA_KERNEL = ("k_apply_wave<K, false> (V-cycle fine-level operator: FP16 MFMA contraction, node sums "
                    "across lanes in registers, fused Chebyshev / residual epilogues in FP32)")


def run_gpu(precision, steps, warmup, nx, ny, nz, profile=False, log=None, keep_fields=False, cg32=None):
    import ctypes as C

    import numpy as np

    from loadpath_amd import _lib
    from loadpath_amd.problems import multiload_256x128x128
    from loadpath_amd.simp import GridSimp

    p = multiload_256x128x128(nx, ny, nz)
    s = GridSimp(p, precision=precision, smoother="chebyshev", rtol=1e-8, maxiter=500, max_levels=6,
                 coarse_cells=0, element="im", filter="top3d")
    lib = _lib.lib()
    if cg32 is not None and precision == "mixed":
        _lib.check(lib.lp_tune(s.grid._h, 13, int(cg32)), "tune")
        if os.environ.get("LP_CG32_DELTA"):
            lib.lp_tune(s.grid._h, 15, int(os.environ["LP_CG32_DELTA"]))
        if os.environ.get("LP_CG32_RESTART"):
            lib.lp_tune(s.grid._h, 19, int(os.environ["LP_CG32_RESTART"]))
    if precision == "mixed" and os.environ.get("LP_BF16"):
        _lib.check(lib.lp_tune(s.grid._h, 16, int(os.environ["LP_BF16"])), "tune")
    if precision != "fp64" and os.environ.get("LP_L1_EXPLICIT32"):
        _lib.check(lib.lp_tune(s.grid._h, 18, int(os.environ["LP_L1_EXPLICIT32"])), "tune")
    # LP_TUNE="param:value,...": lp_tune settings of the mixed path's grid
    # (kernel-variant comparisons; the defaults are the measured best)
    if precision == "mixed" and os.environ.get("LP_TUNE"):
        for item in os.environ["LP_TUNE"].split(","):
            pv = item.split(":")
            if lib.lp_tune(s.grid._h, int(pv[0]), int(pv[1])) < 0:
                raise SystemExit(f"LP_TUNE: lp_tune({pv[0]}, {pv[1]}) refused")
    hist = []
    for k in range(warmup):
        hist += s.iterate(1)
        if log:
            log(f"  {precision} warmup {k}: {hist[-1]}")
    lib = _lib.lib()
    if warmup > 0:
        # launch shape of the V-cycle's fine-level operator, timed on this
        # grid (set-up, outside the timed region)
        xc = s.grid.autotune(s.nrhs)
        if log:
            log(f"  {precision} autotuned x-chunk: {xc}")
    eb = 8 if precision == "fp64" else 4
    if profile:
        _lib.check(lib.lp_profile(s.grid._h, 0, eb, 1), "profile")
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timed = s.iterate(steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if log:
        for k, st in enumerate(timed):
            log(f"  {precision} step {k}: {st}")
    prof = None
    if profile:
        n = C.c_longlong()
        ms = C.c_double()
        by = C.c_double()
        rc = lib.lp_profile_stats(s.grid._h, C.byref(n), C.byref(ms), C.byref(by))
        prof = dict(launches=n.value, ms=ms.value, bytes=by.value, complete=(rc == 0))
        lib.lp_profile(s.grid._h, 0, eb, 0)
    levels = [lv.dims for lv in s.grid.levels]
    out = dict(seconds=dt, iters_per_s=steps / dt, hist=hist + timed, timed=timed, prof=prof, levels=levels,
               fp32_repeats=int(lib.lp_tune(s.grid._h, 6, 0)), strict_switches=int(lib.lp_tune(s.grid._h, 11, 0)),
               reliable_updates=int(lib.lp_tune(s.grid._h, 14, 0)),
               strict=bool(s.grid.strict))
    if keep_fields:
        out.update(xphys=s.xphys, dc=s.dc, u=s.displacements())
    del s
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(gpu_timed, nx, ny, nz, log=None):
    """The FP64 numpy oracle on the headline grid itself (same supports,
    loads and 6-level hierarchy): one hierarchy build and one PCG iteration
    of one load case, timed; a SIMP iteration is estimated as the build +
    that PCG iteration's time x the CG iterations the 8 cases took per SIMP
    iteration in the GPU run's timed iterations (the oracle runs the same
    algorithm, so its counts match to a few iterations)."""
    import numpy as np

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from loadpath_amd.problems import multiload_256x128x128
    from oracle.fem import box_mesh
    from oracle.mg import GridOracle

    p = multiload_256x128x128(nx, ny, nz)
    mesh = box_mesh(nx, ny, nz)
    fg = p.fixed_grid()
    fdof = np.empty(mesh.n_dof, bool)
    for i in range(3):
        fdof[i::3] = fg[i].ravel()
    cc = max(nx * ny * nz // 2 ** 15, 8)       # 6 levels at 256x128x128 (coarsest 8x4x4), as the GPU runs
    g = GridOracle(mesh, 0.3, fdof, rtol=1e-8, maxiter=1, coarse_cells=cc)
    e = np.full(mesh.n_elem, 1e-9 + 0.15 ** 3 * (1 - 1e-9))
    t0 = time.perf_counter()
    g.set_moduli(e)
    t_build = time.perf_counter() - t0
    b = p.load_grids()[0] * g.levels[0].mask
    t0 = time.perf_counter()
    _, it, _ = g._pcg(b, None)
    t_it = (time.perf_counter() - t0) / max(it, 1)
    cg_total = float(np.mean([st["cg_iters_total"] for st in gpu_timed]))
    t_simp = t_build + t_it * cg_total
    if log:
        log(f"  cpu: build {t_build:.2f}s ({len(g.levels)} levels), {t_it:.3f}s per PCG iteration, "
            f"{cg_total:.0f} CG its/SIMP it")
    return dict(value=1.0 / t_simp, unit="SIMP iter/s", cores=_CORES, kind="port", cpu_model=_cpu_model(),
                sample=f"oracle (numpy FP64 CG + FP32 V-cycle) on the {nx}x{ny}x{nz} grid, {len(g.levels)} levels: "
                       f"hierarchy build {t_build:.1f}s + {it} PCG iteration ({t_it:.2f}s) of one load case, timed; "
                       f"a SIMP iteration = build + {cg_total:.0f} PCG iterations (the 8 cases' count per SIMP "
                       f"iteration in the GPU run), extrapolated by that count only")


def config1_bench(iterations, cpu_iterations, log=None):
    """The 48x24x24 cantilever (27,648 cells, x = 0 clamped, unit load at
    node (48, 0, 12), vf 0.3, rmin 1.5, rtol 1e-8, 3-level MG) for
    `iterations` top3d OC iterations on the device (mixed and FP64) and the
    oracle's loop with its multigrid solves on the host cores for
    `cpu_iterations`, each end to end from rho = 0.3."""
    import numpy as np
    import torch

    from loadpath_amd.problems import cantilever_48x24x24
    from loadpath_amd.simp import GridSimp

    p = cantilever_48x24x24()
    res = {}
    for prec in ("mixed", "fp64"):
        s = GridSimp(p, precision=prec, smoother="chebyshev", rtol=1e-8, maxiter=500, element="im",
                     filter="top3d")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h = s.iterate(iterations)
        torch.cuda.synchronize()
        res[prec] = dict(seconds=time.perf_counter() - t0, c=[q["compliance"] for q in h],
                         cg=[q["cg_iters"] for q in h])
        del s
    out = dict(workload="cantilever_48x24x24, top3d OC, 3-level MG, Chebyshev(2), rtol 1e-8",
               iterations=iterations,
               gpu_mixed_iter_s=round(iterations / res["mixed"]["seconds"], 3),
               gpu_fp64_iter_s=round(iterations / res["fp64"]["seconds"], 3),
               compliance_last=res["fp64"]["c"][-1],
               compliance_rel_err=max(abs(a - b) / abs(b) for a, b in zip(res["mixed"]["c"], res["fp64"]["c"])),
               cg_iters_mixed=res["mixed"]["cg"], cg_iters_fp64=res["fp64"]["cg"])
    if cpu_iterations > 0:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle import simp as osimp
        from oracle.mg import GridOracle

        q = osimp.Problem(p.name, p.dims, p.volfrac, p.rmin, p.fixed, p.cases)
        ev = osimp.Evaluator(q, "im")
        g = GridOracle(ev.mesh, 0.3, ev.fixed, rtol=1e-8, maxiter=500)
        prev = [None]

        def solve(e):
            g.set_moduli(e)
            us = g.solve_many(ev.fs, prev[0])
            prev[0] = us
            return us

        t0 = time.perf_counter()
        r = osimp.optimize_top3d(q, solve=solve, iterations=cpu_iterations, element="im")
        dt = time.perf_counter() - t0
        n = min(cpu_iterations, iterations)
        out.update(cpu_iter_s=round(cpu_iterations / dt, 4), cpu_iterations=cpu_iterations, cpu_cores=_CORES,
                   cpu_compliance_rel_err=abs(r["history"][n - 1][1] - res["fp64"]["c"][n - 1])
                   / abs(res["fp64"]["c"][n - 1]))
    if log:
        log(f"  config 1: {out}")
    return out


# The reference's own published timings on this path (its author's 4-core
# CPU): the MG-PCG solve of a solid 128 x 64 x 64 cantilever (524,288 cells,
# rtol 1e-6, 11 CG iterations; docs/BENCHMARKS.md:788) and the two golden
# top3d runs (sparse direct solves; the 'seconds' and 'iterations' of
# benchmarks/reference/{cantilever-3d,bridge-3d}.top3d.npz).
PUBLISHED = {
    "mg_solve_solid_cantilever_128x64x64": {"seconds": 1.46, "cg_iters": 11, "source": "docs/BENCHMARKS.md:788"},
    "top3d_cantilever-3d": {"seconds": 142.0, "iterations": 133,
                            "source": "benchmarks/reference/cantilever-3d.top3d.npz"},
    "top3d_bridge-3d": {"seconds": 761.0, "iterations": 172, "source": "benchmarks/reference/bridge-3d.top3d.npz"},
}


def published_bench(log=None):
    """The runs behind the reference's published CPU timings, here:
    - the MG-PCG solve of a solid cantilever 128 x 64 x 64 (x = 0 clamped,
      the x = 128 bottom edge loaded down; unit moduli) at rtol 1e-6 through
      GridElasticity's host API (set_moduli + solve_many, host arrays in and
      out), the V-cycle in FP32 (the reference's default) and mixed, after
      one untimed solve;
    - the golden top3d runs (plain element, vf and rmin of each, FP64 device
      loop with multigrid PCG to 1e-12, warm-started) for the golden's
      iteration count, with the final compliance beside the golden's."""
    import numpy as np
    import torch

    from loadpath_amd.mg import GridElasticity
    from loadpath_amd.problems import bridge_3d, cantilever_3d
    from loadpath_amd.simp import GridSimp

    out = {}
    p = cantilever_3d(128, 64, 64)
    mesh = p.mesh()
    fixed, loads = p.dof_vectors(mesh)
    e = np.ones(mesh.n_elem)
    for prec in ("fp32", "mixed"):
        g = GridElasticity(mesh, 0.3, fixed, rtol=1e-6, maxiter=500, precision=prec)
        g.set_moduli(e)
        g.solve_many([loads[0]])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.set_moduli(e)
        u = g.solve_many([loads[0]])[0]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[f"mg_solve_solid_cantilever_128x64x64_{prec}"] = dict(
            seconds=round(dt, 4), cg_iters=int(g.last_iters), compliance=float(loads[0] @ u),
            published_cpu_seconds=PUBLISHED["mg_solve_solid_cantilever_128x64x64"]["seconds"],
            published_cg_iters=PUBLISHED["mg_solve_solid_cantilever_128x64x64"]["cg_iters"],
            speedup=round(PUBLISHED["mg_solve_solid_cantilever_128x64x64"]["seconds"] / dt, 1))
        del g
    for name, make in (("cantilever-3d", cantilever_3d), ("bridge-3d", bridge_3d)):
        pub = PUBLISHED[f"top3d_{name}"]
        s = GridSimp(make(), element="plain", rtol=1e-12, maxiter=4000, coarse_cells=500, precision="fp64")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h = s.iterate(pub["iterations"])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[f"top3d_{name}"] = dict(iterations=pub["iterations"], seconds=round(dt, 3),
                                    iter_s=round(pub["iterations"] / dt, 2), compliance_last=h[-1]["compliance"],
                                    published_cpu_seconds=pub["seconds"],
                                    published_iter_s=round(pub["iterations"] / pub["seconds"], 3),
                                    speedup=round(pub["seconds"] / dt, 1))
        del s
    if log:
        log(f"  published runs: {out}")
    return out


def run_studies_bench(log=None, max_iterations=40, loop=False):
    """The 32 uneven studies (problems.batched_studies, default_rng(5)):
    all at once on their own streams (mixed path, graph-replayed V-cycles,
    autotuned fine operator) against the same 32 run one after another on
    the plain FP64 path (no graphs, default launch shape). Each study runs
    to change <= 0.01 or max_iterations SIMP iterations; with loop, each
    runs the production loop (simp.optimize's) to its stop or
    max_iterations."""
    import numpy as np

    from loadpath_amd.batch import run_studies
    from loadpath_amd.problems import batched_studies

    ps = batched_studies()
    run_studies(ps[:2], precision="mixed", max_iterations=1, loop=loop)   # first-use costs out of the timing
    def done(k, st):
        if log:
            log(f"  study {k} ({st.get('name')}): {st.get('iterations')} iterations, {st.get('seconds', 0.0):.2f} s")

    b = run_studies(ps, precision="mixed", max_iterations=max_iterations, concurrent=True, loop=loop, progress=done)
    if log:
        log(f"  batch: {b['wall_s']:.2f} s, p50 {b['p50_s']:.2f} s, p99 {b['p99_s']:.2f} s")
    q = run_studies(ps, precision="fp64", max_iterations=max_iterations, concurrent=False, graph=False,
                    autotune=False, loop=loop, progress=done)
    if log:
        log(f"  sequential fp64: {q['wall_s']:.2f} s")

    def stopped(st):
        # reached its own stop test before the iteration cap
        last = st["history"][-1]
        return bool(last["converged"]) if loop else last["change"] <= 0.01

    conv = [st for st in b["studies"] if stopped(st)]
    conv_s = sorted(st["seconds"] for st in conv)
    return {
        "metric": "studies/s, 32 uneven SIMP studies (nx 32-96, ny/nz 16-48) in one batch"
                  + (", production loop" if loop else ""),
        "value": round(b["studies_per_s"], 4), "unit": "studies/s", "n_gpus": 1, "higher_is_better": True,
        "p50_s": round(b["p50_s"], 3), "p99_s": round(b["p99_s"], 3), "wall_s": round(b["wall_s"], 3),
        "sequential_fp64_studies_per_s": round(q["studies_per_s"], 4),
        "sequential_fp64_p50_s": round(q["p50_s"], 3), "sequential_fp64_p99_s": round(q["p99_s"], 3),
        "speedup_vs_sequential_fp64": round(q["wall_s"] / b["wall_s"], 3),
        "iterations": [s["iterations"] for s in b["studies"]],
        "converged": len(conv),
        "converged_p50_s": round(float(np.percentile(conv_s, 50)), 3) if conv else None,
        "converged_p99_s": round(float(np.percentile(conv_s, 99)), 3) if conv else None,
        "sequential_fp64_converged": sum(stopped(st) for st in q["studies"]),
        "config": {"workload": "batched_studies(count=32, seed=5): cantilever / half-MBB alternating, point or "
                               "4x4-node patch load in a seeded direction, vf U(0.1, 0.4), rmin 1.5 / 2.0, "
                               + ("the production loop (padded filter, eroded analysis, continuation, measured "
                                  "step), stop at its test or %d iterations" if loop else
                                  "top3d filter + OC, stop at change <= 0.01 or %d iterations") % max_iterations},
        "data": "synthetic",
    }


def run_loop_bench(steps, warmup, nx, ny, nz, log=None):
    """The production loop (simp.optimize's loop: eroded analysis, projection
    continuation, measured step, multiplier root to 1e-13) on the same
    256x128x128 multi-load workload: mixed and FP64 iterations/s from the
    same start, with the loop's own padded filter and E_MIN."""
    from loadpath_amd.problems import multiload_256x128x128
    from loadpath_amd.simp import GridSimp, SimpOptions
    import torch

    out = {}
    for precision in ("mixed", "fp64"):
        p = multiload_256x128x128(nx, ny, nz)
        opts = SimpOptions(volume_fraction=p.volfrac, filter_radius=p.rmin, solve_rtol=1e-8)
        s = GridSimp(p, precision=precision, smoother="chebyshev", rtol=1e-8, maxiter=500, max_levels=6,
                     coarse_cells=0, element="im", filter="padded", volfrac=opts.volume_fraction,
                     rmin=opts.filter_radius)
        s.set_loop(opts)
        hist = s.iterate(warmup) if warmup > 0 else []
        if warmup > 0:
            s.grid.autotune(s.nrhs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        timed = s.iterate(steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if log:
            for st in hist + timed:
                log(f"  loop {precision}: {st}")
        out[precision] = dict(value=steps / dt, timed=timed)
        del s
    m, f = out["mixed"], out["fp64"]
    return {
        "metric": "production-loop SIMP iter/s at 256x128x128 hex8 (eroded analysis, projection continuation, "
                  "measured step, volume root to 1e-13)",
        "value": round(m["value"], 4), "unit": "SIMP iter/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 / m["value"], 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 CG / f32 V-cycle / f16 MFMA fine operator", "data": "synthetic (seeded load cases)",
        "config": {"workload": "multiload_256x128x128 as the default line, the production loop's options at their "
                               "defaults (beta 1 -> 32, eta 0.5, erode 0.15, move 0.2 measured), padded filter "
                               "r 2.5, E_MIN 1e-6, rtol 1e-8", "cells": nx * ny * nz},
        "fp64_value": round(f["value"], 4), "speedup_vs_fp64": round(m["value"] / f["value"], 3),
        "compliance_rel_err": max(abs(a["compliance"] - b["compliance"]) / abs(b["compliance"])
                                  for a, b in zip(m["timed"], f["timed"])),
        "volume_evaluations_per_step": [st["oc_evals"] for st in m["timed"]],
        "cg_iters_per_step": [st["cg_iters_total"] for st in m["timed"]],
    }


def run_constrained_bench(steps, warmup, no_cpu=False, log=None):
    """The constrained minimum-mass loop (constrained._constrained_run on a
    box: MMA, deflection and stress limits, beta continuation) on the
    48x24x24 cantilever: iterations/s with the corner-sampled stress and
    with the material-averaged stress, each from the same start, timed from
    the loop's own per-iteration clock after the warmup iterations; the
    oracle's iteration (direct solves) timed on one iteration beside it."""
    import numpy as np

    from loadpath_amd.constrained import Limits, optimize_constrained
    from loadpath_amd.problems import cantilever_48x24x24
    from loadpath_amd.simp import SimpOptions

    p = cantilever_48x24x24()
    pres = np.zeros(p.dims, bool)
    pres[-2:, :2, 10:15] = True          # a pad under the load
    kw = dict(volume_fraction=0.3, filter_radius=1.5, max_iterations=warmup + steps, beta_every=10, beta_max=8.0)
    limits = dict(stress=1.5, deflection=0.2, modulus=1000.0)
    out = {}
    for name, radius in (("corner", None), ("averaged", 1.3)):
        r = optimize_constrained(p, 0.3, SimpOptions(solve_rtol=1e-6, **kw), Limits(radius=radius, **limits),
                                 preserve=pres, precision="mixed", autotune=True, max_levels=4, coarse_cells=0)
        h = r.history
        t_w = h[warmup - 1]["seconds"] if warmup > 0 else 0.0
        dt = h[warmup + steps - 1]["seconds"] - t_w
        if log:
            for q in h:
                log(f"  constrained {name}: {q}")
        out[name] = dict(value=steps / dt, hist=h[warmup:warmup + steps])
    cpu = None
    if not no_cpu:
        from oracle.constrained import Constrained, Limits as OLimits
        from oracle.loop import LoopOptions
        from oracle.simp import Problem

        # a half-scale cantilever (24x12x12, the same load and pad): the
        # oracle's direct solve of the full grid takes minutes per iteration
        from oracle.simp import nodes as onodes

        op = Problem("cantilever-24x12x12", (24, 12, 12), p.volfrac, p.rmin,
                     [(onodes(0, range(13), range(13)), "xyz")], [[(onodes(24, 0, 6), "z", -1.0)]])
        hp = np.zeros((24, 12, 12), bool)
        hp[-1:, :1, 5:8] = True
        L = Constrained(op, LoopOptions(**kw), OLimits(**limits), preserve=hp.ravel())
        x = np.full(L.n, 0.3)
        x[L.pres] = 1.0
        reps, t0 = 0, time.perf_counter()
        while reps < 10 and (reps == 0 or time.perf_counter() - t0 < 10.0):
            L.constraints(x, 1.0, [1.0], 1)
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        cpu = {"value": round(1.0 / dt, 5), "unit": "constrained iter/s", "cores": 1, "kind": "port",
               "sample": "one iteration of the oracle's loop (filter, projections, direct solve of the load case "
                         "and the adjoint, deflection and corner-stress partials) at the first iterate, on the "
                         "half-scale 24x12x12 cantilever (3456 cells; the full grid's direct solve takes minutes)"}
    c, a = out["corner"], out["averaged"]
    return {
        "metric": "constrained min-mass iter/s at 48x24x24 hex8 (MMA, deflection + corner-stress limits)",
        "value": round(c["value"], 4), "unit": "constrained iter/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(1e3 / c["value"], 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 CG / f32 V-cycle / f16 MFMA fine operator, f64 stress and MMA",
        "data": "synthetic (tip load on a clamped cantilever)",
        "config": {"workload": "cantilever_48x24x24, preserved pad under the load, stress limit 1.5, deflection 0.2, "
                               "modulus 1000, beta 1 -> 8, solve rtol 1e-6", "cells": 48 * 24 * 24},
        "averaged_value": round(a["value"], 4),
        "constraints_last": c["hist"][-1]["constraints"] if c["hist"] else None,
        "cg_iters_per_step": [q.get("cgIterations") for q in c["hist"]],
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--cg32", type=int, default=None, choices=(0, 1),
                    help="the mixed path's CG in FP32 with reliable FP64 updates (1) or in FP64 (0); default: the "
                         "library's")
    ap.add_argument("--config1-cpu-iters", type=int, default=50,
                    help="oracle iterations of the 48x24x24 cantilever timed on the host cores (0: none)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--config1-only", action="store_true",
                    help="instead: only the 48x24x24 cantilever leg (config1 in the default line), as its own JSON")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the runs behind the reference's published CPU timings (the 128x64x64 MG solve, the "
                         "golden top3d runs)")
    ap.add_argument("--size", type=int, nargs=3, default=None, help="override the grid (testing)")
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--studies", action="store_true",
                    help="instead: the 32 uneven batched studies (concurrent, mixed, graphs, autotuned) against "
                         "the same studies one after another on the plain FP64 path")
    ap.add_argument("--studies-iters", type=int, default=40,
                    help="with --studies: each study's iteration cap (its own stop test ends it earlier)")
    ap.add_argument("--loop", action="store_true",
                    help="instead: the production loop (simp.optimize's) on the same workload, mixed against FP64")
    ap.add_argument("--constrained", action="store_true",
                    help="instead: the constrained minimum-mass loop (MMA with stress and deflection limits) on the "
                         "48x24x24 cantilever")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB elementwise copy (3 times): the known-size kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the multi-load solve shards by load case with no exchange during the "
              "solve; runs on more than one GPU are not measured by this script.", file=sys.stderr)
        sys.exit(2)
    log = (lambda m: print(m, file=sys.stderr, flush=True)) if a.verbose else None
    if a.studies:
        print(json.dumps(run_studies_bench(log, max_iterations=a.studies_iters, loop=a.loop)))
        return
    if a.constrained:
        print(json.dumps(run_constrained_bench(a.steps, a.warmup, a.no_cpu, log)))
        return
    if a.config1_only:
        print(json.dumps(config1_bench(50, 0 if a.no_cpu else a.config1_cpu_iters, log)))
        return
    nx, ny, nz = a.size or (WORKLOAD["nx"], WORKLOAD["ny"], WORKLOAD["nz"])
    if a.loop:
        print(json.dumps(run_loop_bench(a.steps, a.warmup, nx, ny, nz, log)))
        return
    if a.copy_calibration:
        import torch

        # an elementwise kernel (grid over the whole array, so it runs on
        # every XCD like the kernels it calibrates): 1 GiB read, 1 GiB written
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    mixed = run_gpu("mixed", a.steps, a.warmup, nx, ny, nz, profile=True, log=log, keep_fields=not a.no_fp64,
                    cg32=a.cg32)
    fp64 = None if a.no_fp64 else run_gpu("fp64", a.steps, a.warmup, nx, ny, nz, log=log, keep_fields=True)
    errs = {}
    if fp64 is not None:
        import numpy as np

        errs["compliance_rel_err"] = max(abs(m["compliance"] - f["compliance"]) / abs(f["compliance"])
                                         for m, f in zip(mixed["hist"], fp64["hist"]))
        # the fields after the last iteration, both paths from the same start
        errs["displacement_rel_err"] = float(np.linalg.norm(mixed["u"] - fp64["u"]) / np.linalg.norm(fp64["u"]))
        errs["sensitivity_rel_err"] = float(np.abs(mixed["dc"] - fp64["dc"]).max() / np.abs(fp64["dc"]).max())
        errs["density_max_abs_err"] = float(np.abs(mixed["xphys"] - fp64["xphys"]).max())
        errs["volume_fraction"] = [round(float(mixed["xphys"].mean()), 9), round(float(fp64["xphys"].mean()), 9)]
    prof = mixed["prof"]
    roof = None
    if prof and prof["launches"]:
        # the byte model counts every right-hand side of a launch, but the
        # kernel skips the blocks of cases whose CG has converged: only the
        # active share of the per-case bytes moves. Active share = CG
        # iterations summed over the cases / (cases x the batch's
        # iterations), over the timed steps; the shared part of a launch
        # (the inverse diagonal in 3 of the 4 kinds, moduli, mask bits)
        # moves whatever the count.
        nc = WORKLOAD["cases"]
        tot = sum(st["cg_iters_total"] for st in mixed["timed"])
        mx = sum(st["cg_iters"] for st in mixed["timed"])
        f_act = tot / (nc * mx) if mx else 1.0
        nn = (nx + 1) * (ny + 1) * (nz + 1)
        shared = 0.75 * 3 * nn * 4 + nx * ny * nz * 4 + 3 * nn / 8
        L = prof["launches"]
        active_bytes = L * shared + (prof["bytes"] - L * shared) * f_act
        gbs = active_bytes / (prof["ms"] * 1e-3) / 1e9
        flops = 1176.0 * nx * ny * nz * nc * f_act * L
        tfs = flops / (prof["ms"] * 1e-3) / 1e12
        roof = dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 4),
                    traffic=FINE_MFMA_TRAFFIC_BYTES if (nx, ny, nz) == (WORKLOAD["nx"], WORKLOAD["ny"], WORKLOAD["nz"])
                    else None, kernel=FINE_MFMA_KERNEL,
                    launches=L, avg_us=round(prof["ms"] * 1e3 / L, 2), avg_bytes=round(active_bytes / L),
                    avg_bytes_all_cases=round(prof["bytes"] / L), active_case_fraction=round(f_act, 4),
                    compute=dict(achieved=round(tfs, 2), peak=FP16_MFMA_PEAK_TFLOPS, unit="TFLOP/s",
                                 frac=round(tfs / FP16_MFMA_PEAK_TFLOPS, 4), flop_per_cell=1176))
    cfg1 = None if a.no_config1 else config1_bench(50, 0 if a.no_cpu else a.config1_cpu_iters, log)
    pubr = None if a.no_published else published_bench(log)
    cpu = None if a.no_cpu else cpu_baseline(mixed["timed"], nx, ny, nz, log)
    ms = 1e3 * mixed["seconds"] / a.steps
    out = {
        "metric": "SIMP iter/s at 256x128x128 hex8, mixed vs FP64 GPU, <=1e-6 rel compliance err",
        "value": round(mixed["iters_per_s"], 4),
        "unit": "SIMP iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 CG / f32 V-cycle / f16 MFMA fine operator",
        "data": "synthetic (seeded load cases, uniform start density)",
        "config": {"workload": f"multiload_{nx}x{ny}x{nz}: 8 load cases (unit force, seeded direction, 8x8-node top "
                               "patch), bottom face clamped, vf 0.15, rmin 2.5, SIMP p=3 Emin 1e-9, top3d density "
                               "filter + OC bisection (move 0.2), batched MG-PCG (6 levels, Chebyshev(2), rtol 1e-8)",
                   "cells": nx * ny * nz, "load_cases": 8, "levels": mixed["levels"]},
        "fp64_value": round(fp64["iters_per_s"], 4) if fp64 else None,
        "speedup_vs_fp64": round(mixed["iters_per_s"] / fp64["iters_per_s"], 3) if fp64 else None,
        **errs,
        "cg_iters_per_step": [st["cg_iters_total"] for st in mixed["timed"]],
        "fp64_cg_iters_per_step": [st["cg_iters_total"] for st in fp64["timed"]] if fp64 else None,
        "fp32_repeats": mixed["fp32_repeats"],
        "reliable_updates": mixed["reliable_updates"],
        "strict_switches": [mixed["strict_switches"], fp64["strict_switches"] if fp64 else None],
        "roofline": roof,
        "cpu_baseline": cpu,
        "config1": cfg1,
        "published_runs": pubr,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
