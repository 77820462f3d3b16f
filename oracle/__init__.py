"""FP64 CPU oracle for loadpath's voxel SIMP hot path.

TEST INFRASTRUCTURE ONLY. This package is the checker the HIP path is
compared against: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it. The product package
``loadpath_amd`` never imports, links or runs anything from here, and has no
CPU fallback onto it.

Each module restates the reference algorithm in plain numpy (FP64), citing
the reference file:line it follows:

- ``oracle.fem``     element stiffness, mesh numbering (loadpath/fem.py)
- ``oracle.mg``      matrix-free geometric multigrid + PCG (loadpath/mg.py,
                     loadpath/mgops.py)
- ``oracle.simp``    filters, sensitivities, OC loops, evaluator
                     (loadpath/simp.py, loadpath/benchmarks/*.py)
- ``oracle.mma``     Svanberg MMA (loadpath/mma.py)
- ``oracle.stress``  p-norm von Mises + adjoint (loadpath/constrained.py)

How each function is pinned (golden vectors, closed forms, invariants, brute
force) is listed in each module's header; tests/test_oracle_*.py hold the
checks.
"""
