"""Scalar eta generators (src/em3d/acoustics/materials.py:11-83 of the reference).

Masks are evaluated on the host from the grid's real-dtype axes, then copied.
"""
from __future__ import annotations

import numpy as np

from ..grid import Grid


def eta_homogeneous(grid: Grid, eta_value):
    be = grid.backend
    return be.array(np.full(grid.N, be.complex_dtype(eta_value), dtype=be.complex_dtype))


def eta_slab(grid: Grid, *, eta_inside, eta_outside=1.0, axis: int = 0, width_fraction: float = 0.5):
    if axis not in (0, 1, 2):
        raise ValueError(f"axis must be 0, 1, or 2, got {axis!r}")
    if not (0.0 < float(width_fraction) <= 1.0):
        raise ValueError(f"width_fraction must be in (0, 1], got {width_fraction!r}")
    be = grid.backend
    coords = grid.host_coords()
    half = 0.5 * float(grid.L[axis]) * float(width_fraction)
    mask = np.abs(coords[axis] - float(grid.center[axis])) <= half
    c = be.complex_dtype
    return be.array(np.where(mask, c(eta_inside), c(eta_outside)).astype(c, copy=False))


def eta_ellipsoid(grid: Grid, *, center, radii, eta_inside, eta_outside=1.0):
    if len(center) != 3:
        raise ValueError(f"center must have length 3, got {center!r}")
    if len(radii) != 3 or any(float(r) <= 0.0 for r in radii):
        raise ValueError(f"radii must contain three positive values, got {radii!r}")
    be = grid.backend
    X, Y, Z = grid.host_coords()
    cx, cy, cz = (float(x) for x in center)
    rx, ry, rz = (float(x) for x in radii)
    metric = ((X - cx) / rx) ** 2 + ((Y - cy) / ry) ** 2 + ((Z - cz) / rz) ** 2
    c = be.complex_dtype
    return be.array(np.where(metric <= 1.0, c(eta_inside), c(eta_outside)).astype(c, copy=False))


def eta_sphere(grid: Grid, *, center, radius: float, eta_inside, eta_outside=1.0):
    return eta_ellipsoid(grid, center=center, radii=(float(radius),) * 3, eta_inside=eta_inside,
                         eta_outside=eta_outside)
