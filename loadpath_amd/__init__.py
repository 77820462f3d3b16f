"""loadpath_amd: loadpath's voxel SIMP hot path on AMD MI355X (gfx950).

Matrix-free hex8 Ke.u operator, geometric-multigrid-preconditioned CG,
compliance sensitivities, density filter and OC update, as hand-written HIP
kernels in liblpmi.so (built in-tree: ``python -m loadpath_amd.build``),
behind loadpath's own Python call signatures:

- ``loadpath_amd.mg.GridElasticity`` for ``loadpath.mg.GridElasticity``
- ``loadpath_amd.mg.prolong`` / ``restrict``
- ``loadpath_amd.simp.density_filter`` / ``GridFilter`` / ``optimize_top3d``
- ``loadpath_amd.gpu.resolve_device`` / ``HipBackend`` (the backend slot)
"""

__version__ = "0.1.0"
