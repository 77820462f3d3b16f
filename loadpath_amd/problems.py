"""Box problems on unit voxel grids: supports, loads and seeded inputs.

Textbook cases follow loadpath/benchmarks/problems.py (cantilever_3d
:107-119, bridge_3d :122-136). The SIMP workloads of this package are named by
what they are:

- ``cantilever_48x24x24``: clamped x = 0 face, unit downward point load at
  node (48, 0, 12); vf 0.3, rmin 1.5.
- ``mbb_192x64x64``: half MBB beam: ux = 0 on x = 0, a roller along the
  bottom edge of x = 192 that moves in x only (uy = uz = 0: with uz alone
  the rigid y-translation is free and K singular; top88's MBB, which this
  extends, is plane), unit line load down along the top edge at x = 0;
  density U(0.05, 1) from default_rng(2).
- ``lbracket_128x64x64``: the box minus x in [64, 128) x z in [32, 64); the
  top face (z = 64, x < 64) clamped; unit downward traction on the x = 128
  face nodes with z in [0, 4]; density U(0.1, 1) from default_rng(3).
- ``multiload_256x128x128``: bottom face z = 0 clamped; 8 load cases, each
  a unit total force with direction standard_normal(3)/|.| spread over an
  8 x 8-node patch of the top face, origin integers(...), default_rng(4);
  vf 0.15, rmin 2.5.
- ``batched_studies``: 32 independent studies from default_rng(5): cells
  nx in [32, 96], ny, nz in [16, 48], multiples of 8; vf U(0.1, 0.4); rmin
  1.5 or 2.0; cantilevers (x = 0 clamped, load on the x = nx face) and half
  MBB beams (ux = uy = 0 on x = 0, uz = 0 on the bottom edge at x = nx, load
  on the top edge at x = 0) alternating; a unit force in a seeded direction
  at one node or spread over a 4 x 4-node patch.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_AX = {"x": 0, "y": 1, "z": 2}


def nodes(i, j, k):
    ii, jj, kk = np.meshgrid(np.atleast_1d(i), np.atleast_1d(j), np.atleast_1d(k), indexing="ij")
    return np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.int64)


@dataclass
class BoxProblem:
    name: str
    dims: tuple                       # cells (nx, ny, nz)
    volfrac: float
    rmin: float
    fixed: list                       # [(nodes (N, 3), "xyz")]
    cases: list                       # [[(nodes, comp, value) or (nodes, vec3)], ...]
    present: np.ndarray | None = None  # bool cells in the mesh (None: the full box)
    density: np.ndarray | None = None  # seeded density field (cells) for solve-only workloads
    extra: dict = field(default_factory=dict)

    @property
    def node_shape(self):
        return tuple(d + 1 for d in self.dims)

    def fixed_grid(self) -> np.ndarray:
        """(3, nx+1, ny+1, nz+1) bool held components."""
        g = np.zeros((3,) + self.node_shape, bool)
        for nd, comps in self.fixed:
            for c in comps:
                g[_AX[c], nd[:, 0], nd[:, 1], nd[:, 2]] = True
        return g

    def mesh(self):
        """HexMesh of the present cells (unit pitch)."""
        from .fem import HexMesh

        if self.present is None:
            ii, jj, kk = np.meshgrid(*[np.arange(d) for d in self.dims], indexing="ij")
            cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)
        else:
            cells = np.argwhere(self.present)
        return HexMesh(cells, 1.0)

    def dof_vectors(self, mesh):
        """(fixed (n_dof,) bool, loads (ncases, n_dof)) on a mesh of this box."""
        at = tuple(mesh.node_ijk.T)
        fg = self.fixed_grid()
        lg = self.load_grids()
        fixed = np.empty(mesh.n_dof, bool)
        loads = np.empty((len(self.cases), mesh.n_dof))
        for c in range(3):
            fixed[c::3] = fg[c][at]
            loads[:, c::3] = lg[:, c][(slice(None),) + at]
        return fixed, loads

    def load_grids(self) -> np.ndarray:
        """(ncases, 3, nx+1, ny+1, nz+1) FP64 nodal forces."""
        out = np.zeros((len(self.cases), 3) + self.node_shape)
        for k, case in enumerate(self.cases):
            for item in case:
                if len(item) == 3:
                    nd, comp, v = item
                    np.add.at(out[k, _AX[comp]], (nd[:, 0], nd[:, 1], nd[:, 2]), v)
                else:
                    nd, vec = item
                    for c in range(3):
                        np.add.at(out[k, c], (nd[:, 0], nd[:, 1], nd[:, 2]), vec[c])
        return out


def cantilever_3d(nx=60, ny=4, nz=20, volfrac=0.3, rmin=1.5):
    return BoxProblem("cantilever-3d", (nx, ny, nz), volfrac, rmin,
                      [(nodes(0, range(ny + 1), range(nz + 1)), "xyz")],
                      [[(nodes(nx, range(ny + 1), 0), "z", -1.0)]])


def bridge_3d(nx=48, ny=12, nz=16, volfrac=0.15, rmin=1.5):
    top = nodes(range(nx + 1), range(ny + 1), nz)
    return BoxProblem("bridge-3d", (nx, ny, nz), volfrac, rmin,
                      [(nodes([0, nx], range(ny + 1), 0), "xyz")],
                      [[(top, "z", -100.0 / len(top))]])


def cantilever_48x24x24():
    nx, ny, nz = 48, 24, 24
    return BoxProblem("cantilever-48x24x24", (nx, ny, nz), 0.3, 1.5,
                      [(nodes(0, range(ny + 1), range(nz + 1)), "xyz")],
                      [[(nodes(48, 0, 12), "z", -1.0)]])


def mbb_192x64x64(seed=2, nx=192, ny=64, nz=64):
    """Scaled copies (other nx, ny, nz) keep the supports, the load and the
    seeded density draw on their own grid."""
    top = nodes(0, range(ny + 1), nz)
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.05, 1.0, (nx, ny, nz))
    return BoxProblem(f"mbb-{nx}x{ny}x{nz}", (nx, ny, nz), 0.5, 2.0,
                      [(nodes(0, range(ny + 1), range(nz + 1)), "x"), (nodes(nx, range(ny + 1), 0), "yz")],
                      [[(top, "z", -1.0 / len(top))]], density=rho)


def lbracket_128x64x64(seed=3, nx=128, ny=64, nz=64):
    """Scaled copies (other nx, ny, nz) keep the shape: the block
    x >= nx/2, z >= nz/2 removed, the top face x <= nx/2 clamped, the load
    on the x = nx face nodes with z <= nz/16."""
    present = np.ones((nx, ny, nz), bool)
    present[nx // 2:, :, nz // 2:] = False
    tip = nodes(nx, range(ny + 1), range(0, max(nz // 16, 1) + 1))
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.1, 1.0, (nx, ny, nz))
    return BoxProblem(f"lbracket-{nx}x{ny}x{nz}", (nx, ny, nz), 0.3, 1.5,
                      [(nodes(range(0, nx // 2 + 1), range(ny + 1), nz), "xyz")],
                      [[(tip, "z", -1.0 / len(tip))]], present=present, density=rho)


def multiload_256x128x128(nx=256, ny=128, nz=128, ncases=8, patch=8, seed=4):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(ncases):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        ox = int(rng.integers(0, nx + 1 - patch + 1))
        oy = int(rng.integers(0, ny + 1 - patch + 1))
        nd = nodes(range(ox, ox + patch), range(oy, oy + patch), nz)
        cases.append([(nd, d / len(nd))])
    return BoxProblem(f"multiload-{nx}x{ny}x{nz}", (nx, ny, nz), 0.15, 2.5,
                      [(nodes(range(nx + 1), range(ny + 1), 0), "xyz")], cases)


def batched_studies(count=32, seed=5):
    """The uneven studies of one batched run (module docstring)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        nx = 8 * int(rng.integers(4, 13))
        ny = 8 * int(rng.integers(2, 7))
        nz = 8 * int(rng.integers(2, 7))
        vf = float(rng.uniform(0.1, 0.4))
        rmin = float(rng.choice([1.5, 2.0]))
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        patch = bool(rng.integers(0, 2))
        cy, cz = ny // 2, nz // 2
        if k % 2 == 0:
            fixed = [(nodes(0, range(ny + 1), range(nz + 1)), "xyz")]
            nd = nodes(nx, range(cy - 2, cy + 2), range(cz - 2, cz + 2)) if patch else nodes(nx, cy, cz)
            name = f"cantilever-{nx}x{ny}x{nz}"
        else:
            fixed = [(nodes(0, range(ny + 1), range(nz + 1)), "xy"), (nodes(nx, range(ny + 1), 0), "z")]
            nd = nodes(0, range(cy - 2, cy + 2), nz) if patch else nodes(0, cy, nz)
            name = f"mbb-{nx}x{ny}x{nz}"
        out.append(BoxProblem(name, (nx, ny, nz), vf, rmin, fixed, [[(nd, d / len(nd))]]))
    return out
