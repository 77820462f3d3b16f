"""The compliance autograd function on CUDA tensors (moduli, solve, strain
energy and sensitivity in device memory: lp_simp_moduli, lp_mask_rhs,
lp_solve and lp_strain_energy on device pointers) against the host path on
the same grid and against the oracle's direct solve and analytic
sensitivity (simp.py:1897-1904)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PENAL, EMIN = 3.0, 1e-9


def test_device_compliance_matches_host_and_oracle(gpu):
    import torch

    from loadpath_amd.autograd import compliance
    from loadpath_amd.mg import GridElasticity
    from loadpath_amd.problems import cantilever_3d
    from oracle.simp import Evaluator, compliance_sensitivity, element_energy
    from oracle.simp import cantilever_3d as oracle_cantilever

    p = cantilever_3d(16, 8, 8)
    mesh = p.mesh()
    fixed, loads = p.dof_vectors(mesh)
    f = loads[0]
    x = np.random.default_rng(11).uniform(0.2, 1.0, len(mesh.elem_ijk))
    g = GridElasticity(mesh, 0.3, fixed, rtol=1e-12, maxiter=4000, coarse_cells=200, dtype=np.float64)
    assert g.cell_order_is_grid
    xh = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    ch = compliance(xh, g, f, PENAL, EMIN)
    ch.backward()
    for _ in range(2):  # the second call reuses the cached load grid
        xd = torch.tensor(x, dtype=torch.float64, device="cuda", requires_grad=True)
        cd = compliance(xd, g, f, PENAL, EMIN)
        cd.backward()
        assert cd.is_cuda and xd.grad.is_cuda
        assert cd.item() == pytest.approx(ch.item(), rel=1e-10)
        gd = xd.grad.cpu().numpy()
        assert np.abs(gd - xh.grad.numpy()).max() <= 1e-9 * np.abs(gd).max()

    # the oracle's assembled FP64 direct solve of the same box (cells in
    # the same C order)
    ev = Evaluator(oracle_cantilever(16, 8, 8))
    u = ev.solve(EMIN + x ** PENAL * (1 - EMIN))
    assert cd.item() == pytest.approx(float(ev.f @ u), rel=1e-9)
    dc = compliance_sensitivity(x, element_energy(u, ev.edof, ev.ke), PENAL, EMIN)
    assert np.abs(gd - dc).max() <= 1e-8 * np.abs(dc).max()


def test_device_path_needs_grid_order(gpu):
    """An L-shaped mesh (cells not the full box) takes the host path and
    still returns a gradient on x's device."""
    import torch

    from loadpath_amd.autograd import compliance
    from loadpath_amd.mg import GridElasticity
    from loadpath_amd.problems import lbracket_128x64x64

    p = lbracket_128x64x64(nx=16, ny=8, nz=8)
    mesh = p.mesh()
    fixed, loads = p.dof_vectors(mesh)
    g = GridElasticity(mesh, 0.3, fixed, rtol=1e-10, maxiter=4000, coarse_cells=200, dtype=np.float64)
    assert not g.cell_order_is_grid
    x = p.density[tuple(mesh.elem_ijk.T)]
    xd = torch.tensor(x, dtype=torch.float64, device="cuda", requires_grad=True)
    c = compliance(xd, g, loads[0], PENAL, EMIN)
    c.backward()
    xh = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    ch = compliance(xh, g, loads[0], PENAL, EMIN)
    ch.backward()
    assert xd.grad.is_cuda
    assert c.item() == pytest.approx(ch.item(), rel=1e-10)
    assert np.allclose(xd.grad.cpu().numpy(), xh.grad.numpy(), rtol=1e-9, atol=0)
