"""Quadtree mesh restatement for the oracle (TEST INFRASTRUCTURE ONLY).

Restates ``Mesh::refineByFunction`` (src/Mesh.hpp:111-180) without p4est:

* ``p4est_new_ext(min_level, fill_uniform)`` -> uniform quadtree at ``min_level``;
* recursive ``p4est_refine``: a quadrant at ``level < max_level`` is split if ANY of
  its ``nx * ny`` cell centres satisfies ``fn`` (Mesh.hpp:131-166), and its children
  are tested again;
* ``p4est_balance(P4EST_CONNECT_CORNER)`` (Mesh.hpp:169): 2:1 balance across faces
  and corners.  The minimal corner-balanced refinement containing a quadtree is
  unique, so the ripple loop below reaches the same forest as p4est;
* leaves are returned in z-order (pre-order with children SW, SE, NW, NE), the
  order of ``Mesh::iteratePatches`` (Mesh.hpp:316-325) and of p4est's local leaves.

``fn`` is vectorised: ``fn(x, y)`` takes numpy arrays of cell centres and returns a
boolean array.
"""
from __future__ import annotations

import numpy as np


def _centres(level, ix, iy, nx, ny, xl, xu, yl, yu):
    # p4est_qcoord_to_vertex on the single square tree (P4est.cpp:7-33): affine map.
    h = 1.0 / (1 << level)
    qx0, qy0 = ix * h, iy * h
    x0 = xl + (xu - xl) * qx0
    x1 = xl + (xu - xl) * (qx0 + h)
    y0 = yl + (yu - yl) * qy0
    y1 = yl + (yu - yl) * (qy0 + h)
    dx = (x1 - x0) / nx
    dy = (y1 - y0) / ny
    xs = x0 + (np.arange(nx) + 0.5) * dx
    ys = y0 + (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    return X.ravel(), Y.ravel()


def refine_by_function(fn, nx, ny, xl, xu, yl, yu, min_level, max_level):
    """Return the balanced leaf list [(level, ix, iy), ...] in z-order."""
    leaves = set()

    def refine(level, ix, iy):
        if level < max_level:
            X, Y = _centres(level, ix, iy, nx, ny, xl, xu, yl, yu)
            if np.any(fn(X, Y)):
                for c in range(4):
                    refine(level + 1, 2 * ix + (c & 1), 2 * iy + (c >> 1))
                return
        leaves.add((level, ix, iy))

    n0 = 1 << min_level
    for ix in range(n0):
        for iy in range(n0):
            refine(min_level, ix, iy)
    leaves = balance_corner(leaves)
    return zorder(leaves)


def covering_leaf(leaves, level, x, y):
    """Leaf of ``leaves`` containing the level-``level`` quadrant (x, y), or None."""
    for l in range(level, -1, -1):
        s = level - l
        key = (l, x >> s, y >> s)
        if key in leaves:
            return key
    return None


def balance_corner(leaves):
    """Minimal 2:1 face+corner balance (p4est_balance with P4EST_CONNECT_CORNER)."""
    leaves = set(leaves)
    changed = True
    while changed:
        changed = False
        split = set()
        for (l, x, y) in leaves:
            if l < 2:
                continue
            n = 1 << l
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    if dx == 0 and dy == 0:
                        continue
                    nxq, nyq = x + dx, y + dy
                    if not (0 <= nxq < n and 0 <= nyq < n):
                        continue
                    cov = covering_leaf(leaves, l, nxq, nyq)
                    if cov is not None and cov[0] < l - 1:
                        split.add(cov)
        if split:
            changed = True
            for (l, x, y) in split:
                leaves.discard((l, x, y))
                for c in range(4):
                    leaves.add((l + 1, 2 * x + (c & 1), 2 * y + (c >> 1)))
    return leaves


def zorder(leaves):
    """Pre-order (z-order) leaf list: children visited SW, SE, NW, NE."""
    out = []
    leaves = set(leaves)

    def visit(level, ix, iy):
        if (level, ix, iy) in leaves:
            out.append((level, ix, iy))
            return
        for c in range(4):
            visit(level + 1, 2 * ix + (c & 1), 2 * iy + (c >> 1))

    visit(0, 0, 0)
    if len(out) != len(leaves):
        raise ValueError("leaf set does not tile the root quadrant")
    return out


def uniform(level):
    return zorder({(level, ix, iy) for ix in range(1 << level) for iy in range(1 << level)})
