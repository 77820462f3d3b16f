"""26-connected labelling on the device (lp_label26, simp.label26 /
connected) against scipy.ndimage.label with the reference's structure
(simp._CONN26, simp.py:62, used at simp.py:264-282 and 1276-1282)."""
import numpy as np
import pytest
from scipy import ndimage

pytestmark = pytest.mark.gpu

CONN26 = np.ones((3, 3, 3), bool)


def _cases():
    rng = np.random.default_rng(3)
    out = [np.zeros((4, 5, 6), bool), np.ones((7, 3, 5), bool), np.ones((1, 1, 1), bool), np.zeros((1, 9, 1), bool)]
    for dims, p in [((17, 13, 11), 0.1), ((17, 13, 11), 0.3), ((33, 20, 9), 0.6), ((64, 1, 40), 0.45),
                    ((1, 50, 50), 0.5), ((48, 24, 24), 0.25)]:
        out.append(rng.random(dims) < p)
    # corner-touching chain: one component under 26-connectivity, many under 6
    d = np.zeros((9, 9, 9), bool)
    for i in range(9):
        d[i, i, i] = True
    out.append(d)
    # a spiral that winds back: the smallest index is reached late
    s = np.zeros((12, 12, 3), bool)
    s[11, :, 1] = True
    s[:, 11, 1] = True
    s[0, 2:, 1] = True
    s[0:9, 2, 1] = True
    s[8, 2:9, 1] = True
    out.append(s)
    return out


@pytest.mark.parametrize("k", range(12))
def test_label26_matches_scipy(gpu, k):
    from loadpath_amd.simp import label26

    m = _cases()[k]
    want, n = ndimage.label(m, structure=CONN26)
    got, ng = label26(m)
    assert ng == n
    assert np.array_equal(got, want)


def test_connected_matches_the_analysis_rule(gpu):
    """simp._Analysis._connected: the alive cells of components that touch
    an anchor (support) cell."""
    from loadpath_amd.simp import connected

    rng = np.random.default_rng(9)
    alive = rng.random((30, 20, 16)) < 0.3
    anchor = np.zeros_like(alive)
    anchor[0] = True
    lbl, _ = ndimage.label(alive, structure=CONN26)
    keep_l = np.unique(lbl[alive & anchor])
    want = alive & np.isin(lbl, keep_l[keep_l > 0])
    assert np.array_equal(connected(alive, anchor), want)
