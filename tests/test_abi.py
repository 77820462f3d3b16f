"""The C-ABI library loads and exports every symbol include/loadpath_mi.h
declares (no GPU needed: nothing is computed on a device here)."""
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "loadpath_mi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from loadpath_amd import _lib

    lib = _lib.load_library()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTS), set(names) ^ set(_lib.EXPORTS)
    assert lib.lp_version() >= 1


def test_host_ke_matches_oracle():
    from loadpath_amd.fem import hex8_ke
    from oracle.fem import hex8_ke as ok

    for el in ("im", "plain"):
        for nu in (0.3, 0.35, 0.2):
            a, b = hex8_ke(nu, el), ok(nu, el)
            assert np.allclose(a, b, rtol=0, atol=1e-14 * np.abs(b).max()), (el, nu)


def test_host_mesh_numbering_matches_oracle():
    from loadpath_amd.fem import HexMesh
    from oracle.fem import Mesh

    rng = np.random.default_rng(3)
    for _ in range(3):
        g = rng.random((7, 5, 6)) < 0.55
        ijk = np.argwhere(g)
        a, b = HexMesh(ijk, 0.7, np.zeros(3)), Mesh(ijk, 0.7)
        assert np.array_equal(a.enode, b.enode)
        assert np.array_equal(a.node_ijk, b.node_ijk)
        assert np.array_equal(a.edof, b.edof)
        q = rng.integers(-1, 8, (50, 3))
        assert np.array_equal(a.node_id(q), b.node_id(q))


def test_box_problems_are_consistent():
    from loadpath_amd import problems as P

    p = P.multiload_256x128x128(16, 8, 8)
    loads = p.load_grids()
    assert loads.shape == (8, 3, 17, 9, 9)
    for k in range(8):
        assert np.linalg.norm(loads[k].reshape(3, -1).sum(axis=1)) == np.float64(1.0) or \
            abs(np.linalg.norm(loads[k].reshape(3, -1).sum(axis=1)) - 1.0) < 1e-12
    fx = p.fixed_grid()
    assert fx[:, :, :, 0].all() and not fx[:, :, :, 1:].any()
    c1 = P.cantilever_48x24x24()
    assert c1.load_grids()[0, 2, 48, 0, 12] == -1.0


def test_product_package_never_imports_the_oracle():
    """The oracle is test infrastructure: no module of the package imports
    it (the product path has no CPU fallback through it)."""
    import ast
    import pathlib

    root = pathlib.Path(__file__).resolve().parents[1] / "loadpath_amd"
    for path in root.rglob("*.py"):
        tree = ast.parse(path.read_text(), str(path))
        for node in ast.walk(tree):
            names = []
            if isinstance(node, ast.Import):
                names = [a.name for a in node.names]
            elif isinstance(node, ast.ImportFrom) and node.module and node.level == 0:
                names = [node.module]
            assert not any(n == "oracle" or n.startswith("oracle.") for n in names), f"{path} imports {names}"
