"""HipBackend in loadpath's backend plugin slot (mg.py:476-484, 587-589,
950-965): a host grid with GridElasticity's attributes builds its own
hierarchy, calls ``_dev.load()`` after each build and ``_dev.solve(b, x0)``
in place of its PCG, and runs the STALL -> strict fallback itself. The
host here is the oracle's restatement of GridElasticity with that slot
added, so the answers are checked against the host's own solve."""
import numpy as np
import pytest

from tests.conftest import cantilever_bar

pytestmark = pytest.mark.gpu


def _host_class():
    from oracle.mg import GridOracle

    class HostGrid(GridOracle):
        """GridOracle with the reference's `_dev` slot: the hierarchy is
        built on the host, then handed to the backend (mg.py:587-589);
        solves go to the backend (mg.py:953-964)."""

        _dev = None

        def set_moduli(self, e):
            super().set_moduli(e)
            if self._dev is not None:
                self._dev.load()

        def _pcg(self, b, x0):
            if self._dev is not None:
                return self._dev.solve(b, x0)
            return super()._pcg(b, x0)

    return HostGrid


def test_backend_solves_like_the_host(gpu):
    from loadpath_amd.gpu import HipBackend, resolve_device

    HostGrid = _host_class()
    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=2)
    host = HostGrid(mesh, 0.3, fixed, rtol=1e-8, coarse_cells=100, maxiter=2000)
    ref = HostGrid(mesh, 0.3, fixed, rtol=1e-8, coarse_cells=100, maxiter=2000)
    assert resolve_device("mi355x") == "mi355x"
    host._dev = HipBackend(host, resolve_device("auto"))
    host.set_moduli(e)
    ref.set_moduli(e)
    u = host.solve_many([f])[0]
    uo = ref.solve_many([f])[0]
    assert np.linalg.norm(u - uo) < 1e-6 * np.linalg.norm(uo)
    assert abs(host.last_iters - ref.last_iters) <= 2, (host.last_iters, ref.last_iters)
    # warm start after a moduli change, as the SIMP loop calls it
    host.set_moduli(e * 1.1)
    ref.set_moduli(e * 1.1)
    w = host.solve_many([f], [u])[0]
    wo = ref.solve_many([f], [uo])[0]
    assert np.linalg.norm(w - wo) < 1e-6 * np.linalg.norm(wo)


def test_backend_follows_rtol_and_maxiter(gpu):
    """The caller's rtol / maxiter reach the device at every solve
    (simp.py:800 sets mg.rtol before each solve)."""
    from loadpath_amd.gpu import HipBackend

    HostGrid = _host_class()
    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=4)
    host = HostGrid(mesh, 0.3, fixed, rtol=1e-3, coarse_cells=100, maxiter=2000)
    host._dev = HipBackend(host)
    host.set_moduli(e)
    b = host.to_grid(f) * host.levels[0].mask
    ku = lambda x: host.levels[0].apply(x)
    x1, it1, ok1 = host._dev.solve(b, None)
    host.rtol = 1e-10
    x2, it2, ok2 = host._dev.solve(b, None)
    assert ok1 and ok2 and it2 > it1
    bn = np.linalg.norm(b)
    assert np.linalg.norm(b - ku(x1)) <= 1e-3 * bn
    assert np.linalg.norm(b - ku(x2)) <= 2e-10 * bn
    host.maxiter = 2
    x3, it3, ok3 = host._dev.solve(b, None)
    assert not ok3 and it3 == 2


def test_backend_strict_after_stall(gpu, monkeypatch):
    """The host's stall switches it to the strict hierarchy; load() hands the
    strict flag to the device, whose rebuild then matches the host's strict
    levels, and the re-solve matches the host's."""
    import oracle.mg as om
    from loadpath_amd.gpu import HipBackend

    monkeypatch.setattr(om, "STALL", 1)
    HostGrid = _host_class()
    mesh, fixed, f, e = cantilever_bar(16, 8, 8, seed=3)
    host = HostGrid(mesh, 0.3, fixed, rtol=1e-9, coarse_cells=100, maxiter=2000)
    ref = HostGrid(mesh, 0.3, fixed, rtol=1e-9, coarse_cells=100, maxiter=2000)
    host._dev = HipBackend(host)
    host.set_moduli(e)
    ref.set_moduli(e)
    u = host.solve_many([f])[0]
    uo = ref.solve_many([f])[0]
    assert host.strict and ref.strict
    assert np.linalg.norm(u - uo) < 1e-7 * np.linalg.norm(uo)
    assert abs(host.last_iters - ref.last_iters) <= 2
