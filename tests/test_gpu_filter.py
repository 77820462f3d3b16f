"""simp.GridFilter(domain, radius) / density_filter(domain, radius, pad)
(simp.py:1143-1222) on the device, for a domain with a pitch and a mesh
that is not a full box, against the oracle's stencil."""
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _domain(dims, pitch, seed=3):
    from loadpath_amd.fem import HexMesh

    rng = np.random.default_rng(seed)
    g = rng.random(dims) < 0.6
    g[0, 0, 0] = g[-1, -1, -1] = True        # the box is the mesh's bounding box
    ijk = np.argwhere(g)
    return SimpleNamespace(mesh=HexMesh(ijk, pitch), dims=dims, pitch=pitch), ijk


@pytest.mark.parametrize("pitch,radius", [(1.0, 1.5), (0.7, 1.6), (0.5, 1.2), (1.3, 2.5)])
@pytest.mark.parametrize("pad", [True, False])
def test_domain_filter_matches_oracle(gpu, pitch, radius, pad):
    from loadpath_amd.simp import density_filter
    from oracle.simp import GridFilter as OFilter

    dims = (9, 7, 6)
    dom, ijk = _domain(dims, pitch)
    h, hs = density_filter(dom, radius, pad=pad)
    o = OFilter(dims, radius, pitch=pitch, pad=pad, ijk=ijk)
    assert h.n == len(ijk) and h.shape == (len(ijk), len(ijk)) and h.T is h
    assert len(h.taps) == len(o.taps)
    v = np.random.default_rng(5).random(len(ijk))
    got, want = h @ v, o @ v
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()
    assert np.abs(hs - o.hs).max() <= 1e-14 * np.abs(o.hs).max()
    # symmetric: v . (H w) == w . (H v)
    w = np.random.default_rng(6).random(len(ijk))
    assert abs(v @ (h @ w) - w @ (h @ v)) <= 1e-12 * abs(v @ (h @ w))


def test_box_filter_keeps_the_dims_form(gpu):
    from loadpath_amd.simp import GridFilter
    from oracle.simp import GridFilter as OFilter

    dims = (8, 5, 4)
    h = GridFilter(dims, 2.0)
    v = np.random.default_rng(7).random(int(np.prod(dims)))
    want = OFilter(dims, 2.0) @ v
    assert np.abs(h @ v - want).max() <= 1e-14 * np.abs(want).max()
