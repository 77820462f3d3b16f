"""Host-side parts of the stress path that need no GPU: the C B sample
matrices from liblpmi's host routine against the oracle."""
import numpy as np
import pytest


@pytest.mark.parametrize("points", ["center", "corners"])
@pytest.mark.parametrize("element", ["im", "plain"])
def test_hex8_cb_matches_oracle(points, element):
    from loadpath_amd import _lib
    from oracle.stress import sample_cb

    _lib.load_library()
    from loadpath_amd.stress import hex8_cb

    got = hex8_cb(0.3, points, element, pitch=0.5)
    assert np.abs(got - sample_cb(0.3, points, element, pitch=0.5)).max() < 1e-13
