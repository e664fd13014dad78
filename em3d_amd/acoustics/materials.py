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


def eta_inclusions(grid: Grid, inclusions, *, eta_outside=1.0):
    """Device eta of a background holding ellipsoidal inclusions
    ``[(center, radii, eta_inside), ...]``, each tested as ``eta_ellipsoid``
    does (the axes' real type, no tolerance), later inclusions overwriting
    earlier ones; one HIP launch."""
    import torch

    from .. import _lib
    be = grid.backend
    recs = []
    for center, radii, eta in inclusions:
        if len(center) != 3 or len(radii) != 3 or any(float(r) <= 0.0 for r in radii):
            raise ValueError("each inclusion needs a 3-vector centre and three positive radii")
        e = complex(eta)
        recs.append([*map(float, center), *map(float, radii), e.real, e.imag])
    dev = be.torch_device
    inc = torch.as_tensor(np.array(recs, dtype=np.float64).reshape(-1, 8)).to(dev)
    eta = torch.empty(tuple(grid.N), dtype=be.precision.torch_complex_dtype, device=dev)
    o = complex(eta_outside)
    _lib.check(_lib.lib().em3d_eta_inclusions(*grid.N, grid.x.data_ptr(), grid.y.data_ptr(), grid.z.data_ptr(),
                                              be.precision.code, len(recs), inc.data_ptr() if recs else None,
                                              o.real, o.imag, eta.data_ptr(), _lib.stream_handle(dev)),
               "em3d_eta_inclusions")
    return eta


def random_spheres(n: int, seed: int, *, radius=(0.03, 0.08), eta_real=(1.5, 3.0), eta_imag=(0.0, 0.2),
                   box=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0)):
    """``n`` seeded spherical inclusions ``(centre, radii, eta)``: radius and
    eta uniform in the given ranges, centres uniform in the box inset by the
    radius."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        r = rng.uniform(*radius)
        c = [rng.uniform(center[a] - box[a] / 2 + r, center[a] + box[a] / 2 - r) for a in range(3)]
        eta = complex(rng.uniform(*eta_real), rng.uniform(*eta_imag))
        out.append((tuple(c), (r, r, r), eta))
    return out
