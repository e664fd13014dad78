"""Plane wave sampled on a grid (src/em3d/wave.py:9-37 of the reference).

The phase k o.r is a sum of three one-axis terms, so it is formed on the three
cell-centre axes and broadcast, x + y first, then + z: per cell the same
float operations, in the same order and in the precision's real dtype, as the
reference's NumPy backend, hence the same values.  One complex exponential per
cell follows; the field is then copied to the device.
"""
from __future__ import annotations

from typing import Tuple

import numpy as np

from .grid import Grid


def flat_wave_vec(grid: Grid, *, k: float, orient: Tuple[float, float, float],
                  amplitude: Tuple[float, float, float], phi0: float = 0.0, sign: int = 1):
    """``A exp(i sign (k o.r + phi0))``, shape (3, Nx, Ny, Nz); orient must be unit."""
    be = grid.backend
    if sign not in (1, -1):
        raise ValueError(f"sign must be 1 or -1, got {sign!r}")
    norm2 = float(orient[0]) ** 2 + float(orient[1]) ** 2 + float(orient[2]) ** 2
    if abs(norm2 - 1.0) > 1e-9:
        raise ValueError(f"orient must be a unit vector, got norm² = {norm2}")
    # o_a r_a on each axis; the components keep their own type, so NumPy's
    # promotion (and with it the phase's precision) is the reference's
    ox, oy, oz = (orient[a] * grid._host[a] for a in range(3))
    proj = (ox[:, None, None] + oy[None, :, None]) + oz[None, None, :]
    carrier = np.exp(1j * (sign * (k * proj + phi0))).astype(be.complex_dtype, copy=False)
    field = np.zeros((3,) + tuple(grid.N), dtype=be.complex_dtype)
    for c in range(3):
        if amplitude[c] != 0.0:
            field[c] = be.complex_dtype(amplitude[c]) * carrier
    return be.array(field)
