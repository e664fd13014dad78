"""Plane wave sampled on a grid (src/em3d/wave.py:9-37 of the reference).

Evaluated on the host with NumPy in the precision's dtypes, as the reference's
NumPy backend does, then copied to the device.
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
    X, Y, Z = grid.host_coords()
    phase = sign * (k * (orient[0] * X + orient[1] * Y + orient[2] * Z) + phi0)
    carrier = np.exp(1j * phase).astype(be.complex_dtype, copy=False)
    out = np.zeros((3,) + tuple(grid.N), dtype=be.complex_dtype)
    for i, a in enumerate(amplitude):
        if a != 0.0:
            out[i] = be.complex_dtype(a) * carrier
    return be.array(out)
