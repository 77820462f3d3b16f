import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

os.environ.setdefault("OPENBLAS_NUM_THREADS", "4")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built liblpmi.so")
    config.addinivalue_line("markers", "slow: long CPU oracle runs (still part of the default CPU suite)")


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests' entry: the HIP library must load and a GPU must exist;
    no fallback (a missing library or GPU fails the test loudly)."""
    import torch

    from loadpath_amd import _lib

    _lib.load_library()
    assert torch.cuda.is_available(), "GPU test without a visible GPU"
    return _lib


def cantilever_bar(nx, ny, nz, seed=0, contrast=True, pitch=1.0):
    """A bar clamped at x = 0 with a tip shear load, random moduli."""
    from oracle.fem import box_mesh

    mesh = box_mesh(nx, ny, nz, pitch)
    fixed = np.zeros(mesh.n_dof, bool)
    root = np.nonzero(mesh.node_ijk[:, 0] == 0)[0]
    for i in range(3):
        fixed[3 * root + i] = True
    f = np.zeros(mesh.n_dof)
    tip = np.nonzero(mesh.node_ijk[:, 0] == nx)[0]
    f[3 * tip + 2] = -1.0 / len(tip)
    rng = np.random.default_rng(seed)
    if contrast:
        e = 1e-6 + rng.uniform(0.05, 1.0, mesh.n_elem) ** 3
    else:
        e = np.ones(mesh.n_elem)
    return mesh, fixed, f, e
