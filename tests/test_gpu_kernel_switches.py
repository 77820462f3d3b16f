"""The kernel-level switches (mgops.apply_masked / coarse_residual /
add_correction / dot / cast, femops.element_energy) on the device, against
brute-force assembly of the per-cell matrices, the oracle's transfers and
numpy (the reference's own definitions of each, mgops.py:247-261, 557-585,
753-773; femops.py:236-242)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(dims=(5, 4, 3), seed=0):
    rng = np.random.default_rng(seed)
    active = rng.random(dims) < 0.75
    nshape = (3,) + tuple(d + 1 for d in dims)
    return rng, active, nshape


def _assemble_apply(active, mats, u, m):
    """sum over active cells of A_c (u m)_c scattered to the grid."""
    nx, ny, nz = active.shape
    Nx, Ny, Nz = nx + 1, ny + 1, nz + 1
    nn = Nx * Ny * Nz
    uf = (u * m).reshape(-1)
    out = np.zeros(3 * nn)
    for c, (i, j, k) in enumerate(np.argwhere(active)):
        dofs = []
        for q in range(8):
            node = ((i + (q & 1)) * Ny + (j + ((q >> 1) & 1))) * Nz + (k + ((q >> 2) & 1))
            dofs += [d * nn + node for d in range(3)]
        dofs = np.array(dofs)
        out[dofs] += mats[c] @ uf[dofs]
    return out.reshape(u.shape)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind", ["scaled", "weights", "explicit"])
def test_apply_masked_matches_assembly(gpu, kind, dtype):
    from loadpath_amd import mgops
    from oracle.fem import hex8_ke
    from oracle.mg import child_prolongations

    rng, active, nshape = _setup(seed=1)
    cells = mgops.Cells(active)
    ke = hex8_ke(0.3)
    n = cells.n
    if kind == "scaled":
        e = rng.uniform(0.1, 1.0, n)
        a, b = ke, e
        mats = [e[c] * ke for c in range(n)]
    elif kind == "weights":
        pk = child_prolongations()
        mq = np.stack([pk[q].T @ ke @ pk[q] for q in range(8)])
        w = rng.uniform(0.0, 1.0, (8, n))
        a, b = mq, w
        mats = [sum(w[q, c] * mq[q] for q in range(8)) for c in range(n)]
    else:
        kc = rng.standard_normal((24, 24, n))
        a, b = kc, None
        mats = [kc[:, :, c] for c in range(n)]
    u = rng.standard_normal(nshape)
    m = (rng.random(nshape) < 0.8).astype(float)
    tol = 1e-13 if dtype == np.float64 else 2e-6
    for finish in (False, True):
        out = np.empty(nshape, dtype)
        mgops.apply_masked(kind, u.astype(dtype), np.asarray(a, dtype), None if b is None else np.asarray(b, dtype),
                           cells, m.astype(dtype), out, finish)
        want = _assemble_apply(active, mats, u, m)
        if finish:
            want = want * m + (1 - m) * u
        assert np.abs(out - want).max() <= tol * np.abs(want).max(), (kind, finish)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_transfers_match_oracle(gpu, dtype):
    from loadpath_amd import mgops
    from oracle.mg import prolong, restrict

    rng = np.random.default_rng(2)
    cshape = (3, 4, 3, 5)
    pshape = (3,) + tuple(2 * n - 1 for n in cshape[1:])
    fshape = (3, pshape[1] - 1, pshape[2], pshape[3] - 1)      # b's nodes inside the padded grid
    b, ax = rng.standard_normal(fshape), rng.standard_normal(fshape)
    m = (rng.random(fshape) < 0.9).astype(float)
    cm = (rng.random(cshape) < 0.9).astype(float)
    rc = mgops.coarse_residual(*(v.astype(dtype) for v in (b, ax, m, cm)), cshape)
    rf = np.zeros(pshape)
    rf[:, :fshape[1], :fshape[2], :fshape[3]] = (b - ax) * m
    want = restrict(rf) * cm
    tol = 1e-14 if dtype == np.float64 else 1e-6
    assert np.abs(rc - want).max() <= tol * np.abs(want).max()
    xc = rng.standard_normal(cshape)
    x = rng.standard_normal(fshape).astype(dtype)
    x0 = x.astype(np.float64)
    assert mgops.add_correction(xc.astype(dtype), cm.astype(dtype), m.astype(dtype), x)
    want = x0 + prolong(xc * cm)[:, :fshape[1], :fshape[2], :fshape[3]] * m
    assert np.abs(x - want).max() <= tol * np.abs(want).max()
    assert not mgops.add_correction(xc, cm.astype(np.float32), m, x0)     # mixed dtypes: the caller's numpy


def test_dot_cast_and_element_energy(gpu):
    from loadpath_amd import femops, mgops
    from oracle.fem import box_mesh, hex8_ke

    rng = np.random.default_rng(3)
    a, b = rng.standard_normal((3, 7, 6, 5)), rng.standard_normal((3, 7, 6, 5))
    assert mgops.dot(a, b) == pytest.approx(float(np.vdot(a, b)), rel=1e-13)
    a32 = mgops.cast(a, np.float32)
    assert a32.dtype == np.float32 and np.array_equal(a32, a.astype(np.float32))
    assert np.array_equal(mgops.cast(a32, np.float64), a32.astype(np.float64))
    assert mgops.available() and femops.available()
    mesh = box_mesh(5, 4, 3)
    ke = hex8_ke(0.3)
    u, v = rng.standard_normal(mesh.n_dof), rng.standard_normal(mesh.n_dof)
    got = femops.element_energy(u, v, mesh.edof, ke)
    want = np.einsum("ij,jk,ik->i", u[mesh.edof], ke, v[mesh.edof])
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    assert np.allclose(femops.element_energy(u, None, mesh.edof, ke),
                       np.einsum("ij,jk,ik->i", u[mesh.edof], ke, u[mesh.edof]), rtol=1e-13)
