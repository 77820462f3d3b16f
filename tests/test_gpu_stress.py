"""The p-mean stress aggregate and the adjoint gradients on the GPU
(constrained.py:607-664) against the FP64 oracle (direct solves), and the
torch.autograd bindings against the oracle's analytic sensitivities and
FP64 central finite differences."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PENAL, EMIN = 3.0, 1e-9


def _lbracket(nx, ny, nz, rtol=1e-12, **kw):
    from loadpath_amd.mg import GridElasticity
    from loadpath_amd.problems import lbracket_128x64x64

    p = lbracket_128x64x64(nx=nx, ny=ny, nz=nz)
    mesh = p.mesh()
    fixed, loads = p.dof_vectors(mesh)
    x = p.density[tuple(mesh.elem_ijk.T)]
    g = GridElasticity(mesh, 0.3, fixed, rtol=rtol, maxiter=4000, coarse_cells=200, dtype=np.float64, **kw)
    return p, mesh, fixed, loads[0], x, g


def _oracle_mesh(mesh):
    from oracle.fem import Mesh

    return Mesh(mesh.elem_ijk, mesh.pitch)


@pytest.mark.parametrize("points", ["center", "corners"])
def test_pnorm_parts_match_oracle(gpu, points):
    from loadpath_amd.stress import StressPNorm
    from oracle import stress as S

    _, mesh, fixed, f, x, g = _lbracket(16, 8, 8)
    g.set_moduli(EMIN + x ** PENAL * (1 - EMIN))
    u = g.solve_many([f])[0]
    agg = StressPNorm(g, points)
    ev = agg.evaluate([u], x, kappa=1.3)
    om = _oracle_mesh(mesh)
    cbs = S.sample_cb(0.3, points)
    gg, pm, smax, gx, rhs_e = S.pnorm_partials(u[om.edof], x, cbs, kappa=1.3)
    rhs = S.assemble_rhs(om.edof, rhs_e, om.n_dof, fixed)
    assert ev["g"][0] == pytest.approx(gg, rel=1e-12)
    assert ev["smax"][0] == pytest.approx(smax, rel=1e-12)
    assert np.abs(ev["gx"][0] - gx).max() <= 1e-12 * np.abs(gx).max()
    assert np.abs(ev["rhs"][0] - rhs).max() <= 1e-12 * np.abs(rhs).max()


def test_element_product_matches_oracle(gpu):
    from oracle import stress as S
    from oracle.fem import hex8_ke

    _, mesh, fixed, f, x, g = _lbracket(16, 8, 8)
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal((2, mesh.n_dof))
    got = g.element_product(a, b)
    want = S.element_product(a, b, _oracle_mesh(mesh).edof, hex8_ke(0.3))
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("points", ["center", "corners"])
def test_pnorm_gradient_matches_oracle_direct(gpu, points):
    from loadpath_amd.stress import StressPNorm
    from oracle import stress as S
    from oracle.fem import hex8_ke

    _, mesh, fixed, f, x, g = _lbracket(16, 8, 8)
    g.set_moduli(EMIN + x ** PENAL * (1 - EMIN))
    u = g.solve_many([f])[0]
    agg = StressPNorm(g, points)
    de = PENAL * x ** (PENAL - 1) * (1 - EMIN)
    val, grads = agg.gradient([u], x, de)
    want_g, want, _, _ = S.pnorm_gradient(_oracle_mesh(mesh), hex8_ke(0.3), fixed, f, x, S.sample_cb(0.3, points),
                                          PENAL, EMIN)
    assert val[0] == pytest.approx(want_g, rel=1e-9)
    assert np.abs(grads[0] - want).max() <= 1e-7 * np.abs(want).max()


def test_autograd_matches_oracle(gpu):
    import torch

    from loadpath_amd.autograd import compliance, stress_pnorm
    from loadpath_amd.stress import StressPNorm
    from oracle import stress as S
    from oracle.fem import hex8_ke
    from oracle.simp import compliance_sensitivity, element_energy

    _, mesh, fixed, f, x, g = _lbracket(16, 8, 8)
    om = _oracle_mesh(mesh)
    ke = hex8_ke(0.3)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    c = compliance(xt, g, f, PENAL, EMIN)
    c.backward()
    _, _, u, _ = S.pnorm_gradient(om, ke, fixed, f, x, S.sample_cb(0.3), PENAL, EMIN)
    assert c.item() == pytest.approx(float(f @ u), rel=1e-9)
    dc = compliance_sensitivity(x, element_energy(u, om.edof, ke), PENAL, EMIN)
    assert np.abs(xt.grad.numpy() - dc).max() <= 1e-8 * np.abs(dc).max()

    xt2 = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    agg = StressPNorm(g, "center")
    # compose with a torch operation on x: d/dx of s(2 x / 2) is s'(x)
    s = stress_pnorm(xt2 * 2.0 / 2.0, agg, f, PENAL, EMIN)
    s.backward()
    want_g, want, _, _ = S.pnorm_gradient(om, ke, fixed, f, x, S.sample_cb(0.3), PENAL, EMIN)
    assert s.item() == pytest.approx(want_g, rel=1e-9)
    assert np.abs(xt2.grad.numpy() - want).max() <= 1e-7 * np.abs(want).max()


def test_lbracket_gradients_vs_central_differences(gpu):
    """The L-bracket at 128 x 64 x 64 (393,216 cells, 1.24M dofs), density
    U(0.1, 1) seeded, p = 8 von Mises at the element centroids: the
    adjoint gradient and the compliance sensitivity against FP64 central
    differences on 32 seeded random elements, at h = 1e-6 (the workload's
    step) and h = 1e-4. Each perturbed state is solved to rtol 1e-13 from
    the unperturbed state. One element moves a global aggregate by ~1e-4
    of its value, so at h = 1e-6 the solves' 1e-13 relative error alone
    puts ~1e-3 of noise into the difference quotient; at h = 1e-4 that noise
    is 1e-5 and the truncation error (h^2 times smooth third derivatives)
    smaller still, so there the adjoint must agree to 1e-4."""
    from loadpath_amd.stress import StressPNorm

    t0 = time.time()
    _, mesh, fixed, f, x, g = _lbracket(128, 64, 64, rtol=1e-13, max_rhs=1)
    agg = StressPNorm(g, "center")
    e0 = EMIN + x ** PENAL * (1 - EMIN)
    g.set_moduli(e0)
    u0 = g.solve_many([f])[0]
    de = PENAL * x ** (PENAL - 1) * (1 - EMIN)
    val, grads = agg.gradient([u0], x, de)
    dc = -de * g.strain_energy(u0)
    idx = np.random.default_rng(7).choice(mesh.n_elem, 32, replace=False)
    gs, gc = grads[0][idx], dc[idx]
    for h, tol in ((1e-6, 1e-3), (1e-4, 1e-4)):
        fd_s, fd_c = [], []
        for e in idx:
            vals = []
            for sgn in (1.0, -1.0):
                xp = x.copy()
                xp[e] += sgn * h
                g.set_moduli(EMIN + xp ** PENAL * (1 - EMIN))
                up = g.solve_many([f], [u0])[0]
                vals.append((agg.evaluate([up], xp, need_grad=False)["g"][0], float(f @ up)))
            fd_s.append((vals[0][0] - vals[1][0]) / (2 * h))
            fd_c.append((vals[0][1] - vals[1][1]) / (2 * h))
        fd_s, fd_c = np.array(fd_s), np.array(fd_c)
        err_s = np.abs(gs - fd_s).max() / np.abs(fd_s).max()
        err_c = np.abs(gc - fd_c).max() / np.abs(fd_c).max()
        print(f"lbracket 128x64x64: pm {val[0]:.6g}, 32 elements, h {h:g}: max |adjoint - FD| / max |FD| = "
              f"{err_s:.2e} (stress), {err_c:.2e} (compliance); {time.time() - t0:.1f} s", flush=True)
        assert err_s < tol and err_c < tol, (h, err_s, err_c)
