"""The assembled route's direct solve on the device (lp_band_*: the band
Cholesky behind fem.DirectSolver and fem.Solver, in place of
direct.DirectSolver and fem.Solver's CG + AMG, direct.py:158-246,
fem.py:586-792) against scipy's sparse direct solve of the same CSR matrix
(Stiffness.assemble, the reference's layout), and on the 48x24x24 benchmark
cantilever against the oracle's compliance."""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _holed_bar(nx, ny, nz, seed):
    """A bar clamped at x = 0 with a few cells removed (not at the root) and
    a tip load in two directions; random moduli."""
    from loadpath_amd.fem import HexMesh

    rng = np.random.default_rng(seed)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)
    keep = (rng.random(len(cells)) > 0.08) | (cells[:, 0] < 2)
    mesh = HexMesh(cells[keep], 1.0)
    fixed = np.zeros(mesh.n_dof, bool)
    root = np.flatnonzero(mesh.node_ijk[:, 0] == 0)
    for i in range(3):
        fixed[3 * root + i] = True
    tip = np.flatnonzero(mesh.node_ijk[:, 0] == nx)
    f1 = np.zeros(mesh.n_dof)
    f1[3 * tip + 2] = -1.0 / len(tip)
    f2 = np.zeros(mesh.n_dof)
    f2[3 * tip + 1] = 0.5 / len(tip)
    scale = 1e-3 + rng.uniform(0.05, 1.0, mesh.n_elem) ** 3
    return mesh, fixed, [f1, f2], scale


@pytest.mark.parametrize("dims", [(12, 6, 4), (5, 9, 7)])
def test_band_factor_matches_sparse_direct(gpu, dims):
    from loadpath_amd.fem import DirectSolver, Stiffness

    mesh, fixed, fs, scale = _holed_bar(*dims, seed=sum(dims))
    st = Stiffness(mesh, 0.3, fixed, "im")
    k = st.assemble(scale)
    want = [spla.spsolve(k.tocsc(), f) for f in fs]
    d = DirectSolver(st)
    # the band of the longest-axis-slowest order: 3 (Nm Nf + Nf + 1) + 2 over
    # the box's node counts (the removed cells only narrow it)
    n = sorted(d_ + 1 for d_ in dims)
    assert d.band <= 3 * (n[1] * n[0] + n[0] + 1) + 2
    d.factor(scale)
    got = d.solve_many(None, fs)
    print(dims, "band", d.band, "relres", d.last_relres)
    assert d.last_relres <= 1e-11
    for g, w in zip(got, want):
        assert np.linalg.norm(g - w) <= 1e-10 * np.linalg.norm(w)
    # the CSR route: the same matrix handed over as the reference's callers do
    got2 = d.solve_many(k, fs)
    for g, w in zip(got2, want):
        assert np.linalg.norm(g - w) <= 1e-10 * np.linalg.norm(w)


def test_solver_interface_and_mechanism(gpu):
    from loadpath_amd.fem import Solver, Stiffness

    mesh, fixed, fs, scale = _holed_bar(10, 5, 5, seed=3)
    st = Stiffness(mesh, 0.3, fixed, "im")
    k = st.assemble(scale)
    s = Solver(mesh, rtol=1e-10)
    u = s.solve(k, fs[0])
    assert np.linalg.norm(u - spla.spsolve(k.tocsc(), fs[0])) <= 1e-10 * np.linalg.norm(u)
    assert 1 <= s.last_iters <= 2
    us = s.solve_many(k, fs)
    assert len(us) == 2 and np.allclose(us[0], u, rtol=0, atol=1e-12 * np.abs(u).max())
    # a second design on the same mesh is factored afresh
    k2 = st.assemble(scale * 2.0)
    u2 = s.solve(k2, fs[0])
    assert np.allclose(u2, u / 2.0, rtol=1e-9, atol=1e-12 * np.abs(u).max())
    # no supports and a net force: no solution; the solve says so
    free = Stiffness(mesh, 0.3, np.zeros(mesh.n_dof, bool), "im")
    with pytest.raises(RuntimeError):
        Solver(mesh, rtol=1e-8).solve(free.assemble(scale), fs[0])


def test_band_factor_cantilever48(gpu):
    """The 48x24x24 benchmark cantilever (91,875 dofs; band 1,955 in the
    x-slowest order) at the uniform start rho = 0.3: the compliance of the
    device factor equals the oracle's first iteration (its multigrid solve
    to 1e-12, checked against an MMD-ordered LU to 1e-10) to 1e-10."""
    import time

    import torch

    from loadpath_amd.fem import DirectSolver, Stiffness
    from loadpath_amd.problems import cantilever_48x24x24

    z = np.load(GOLDEN / "config1_cantilever48.npz")
    p = cantilever_48x24x24()
    mesh = p.mesh()
    fixed, loads = p.dof_vectors(mesh)
    st = Stiffness(mesh, 0.3, fixed, "im")
    d = DirectSolver(st)
    assert d.band == 3 * (25 * 25 + 25 + 1) + 2
    scale = np.full(mesh.n_elem, 1e-9 + 0.3 ** 3 * (1 - 1e-9))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.factor(scale)
    t1 = time.perf_counter()
    u = d.solve_many(None, [loads[0]])[0]
    t2 = time.perf_counter()
    c = float(loads[0] @ u)
    print(f"48x24x24: band {d.band}, factor {d.factor_bytes / 2**30:.2f} GiB, factor {t1 - t0:.3f} s, "
          f"solve {t2 - t1:.3f} s, relres {d.last_relres:.1e}, compliance rel err "
          f"{abs(c - z['compliance'][0]) / z['compliance'][0]:.1e}")
    assert d.last_relres <= 1e-12
    assert c == pytest.approx(float(z["compliance"][0]), rel=1e-10)
