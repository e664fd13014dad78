"""Scalar Helmholtz kernel (src/em3d/acoustics/kernel.py:7-58 of the reference)."""
from __future__ import annotations

import numpy as np

from ..grid import Grid
from ..operator import build_octant_spectrum, require_hip


def self_cell_coefficient(*, k0: float, dv: float) -> complex:
    """Equivalent-sphere ``k0^2 int_cell G dV``: exp(ika)(1 - ika) - 1."""
    k, v = float(k0), float(dv)
    if k <= 0.0:
        raise ValueError(f"k0 must be positive, got {k0}")
    if v <= 0.0:
        raise ValueError(f"dv must be positive, got {dv}")
    a = (3.0 * v / (4.0 * np.pi)) ** (1.0 / 3.0)
    return complex(np.exp(1j * k * a) * (1.0 - 1j * k * a) - 1.0)


def prep_coeffs_acoustic(grid: Grid, *, k0: float):
    """Octant spectrum of ``k0^2 dv G`` (device, FP64 build, backend precision)."""
    require_hip(grid.backend)
    be = grid.backend
    return build_octant_spectrum(grid.N, grid.L, float(k0), 1, be.precision, be.device_index)
