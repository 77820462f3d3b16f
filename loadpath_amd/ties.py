"""Design ties on a box of cells, host side (the constrained loop's design
update; the production loop ties on the device, lp_simp_set_ties).

- ``mirror_tie(v, dims, axes)``: mirror planes through the box centre, one
  axis after another; a cell and its image both take the mean of the two
  (domain.mirror with simp._symmetrize, simp.py:1476-1482, domain.py:
  1981-1988 on a box, where every cell has its image).
- ``column_tie(v, dims, axis, design)``: extrusion columns along ``axis``;
  the design cells of each column take the column's mean over them, summed
  in increasing order along the axis (manufacture.tie_groups(domain,
  extrusion), manufacture.py:413-427, with simp._extrude, simp.py:
  1453-1463).
"""
from __future__ import annotations

import numpy as np

_AXES = {"x": 0, "y": 1, "z": 2}


def axis_index(a) -> int:
    return _AXES[a] if isinstance(a, str) else int(a)


def mirror_tie(v, dims, axes):
    """v (cells, C order over dims) with every mirror plane of ``axes``
    ("xz", [0, 2], ...) applied in order."""
    g = np.asarray(v, float).reshape(dims)
    for a in axes or ():
        g = 0.5 * (g + np.flip(g, axis_index(a)))
    return g.reshape(np.shape(v))


def column_tie(v, dims, axis, design):
    """v with each extrusion column's design cells set to their mean."""
    if axis is None:
        return v
    a = axis_index(axis)
    g = np.asarray(v, float).reshape(dims)
    d = np.asarray(design, bool).reshape(dims)
    run = np.cumsum(np.where(d, g, 0.0), axis=a)            # sequential along the column
    total = np.take(run, -1, axis=a)
    count = d.sum(axis=a)
    mean = np.expand_dims(total / np.maximum(count, 1), a)
    return np.where(d, mean, g).reshape(np.shape(v))


def tie(v, dims, mirror, extrusion, design):
    """Both ties, mirror planes first (simp.py:2223-2226)."""
    return column_tie(mirror_tie(v, dims, mirror), dims, extrusion, design)
