"""Contrast profiles eta = eps - I (src/em3d/refraction.py:11-175 of the reference).

Masks are evaluated on the host from the grid's real-dtype axes exactly as the
reference's NumPy backend does; the (3, 3, Nx, Ny, Nz) result is copied to the
device.
"""
from __future__ import annotations

from typing import Literal, Tuple

import numpy as np

from .grid import Grid


def _eta_matrix(eps_real, eps_imag) -> np.ndarray:
    rs, is_ = np.ndim(eps_real) == 0, np.ndim(eps_imag) == 0
    if rs != is_:
        raise TypeError("eps_real and eps_imag must both be scalars or both be (3,3) arrays")
    if rs:
        if np.iscomplex(eps_real) or np.iscomplex(eps_imag):
            raise TypeError("eps_real and eps_imag must be real scalars")
        return (float(eps_real) - 1.0 + 1j * float(eps_imag)) * np.eye(3, dtype=np.complex128)
    E = np.asarray(eps_real, dtype=np.float64)
    F = np.asarray(eps_imag, dtype=np.float64)
    if E.shape != (3, 3) or F.shape != (3, 3):
        raise ValueError("eps_real and eps_imag must have shape (3,3)")
    return (E - np.eye(3)) + 1j * F


def _masked(grid: Grid, mask, eta):
    out = np.zeros((3, 3) + tuple(grid.N), dtype=grid.backend.complex_dtype)
    for i in range(3):
        for j in range(3):
            out[i, j] = np.where(mask, eta[i, j], out[i, j])
    return grid.backend.array(out)


def cylinder_refraction(grid: Grid, *, eps_real, eps_imag, radius: float, axis: Literal["x", "y", "z"] = "z"):
    X, Y, Z = grid.host_coords()
    if axis == "z":
        r2 = X * X + Y * Y
    elif axis == "y":
        r2 = X * X + Z * Z
    elif axis == "x":
        r2 = Y * Y + Z * Z
    else:
        raise ValueError(f"axis must be 'x'|'y'|'z', got {axis!r}")
    return _masked(grid, r2 <= radius * radius, _eta_matrix(eps_real, eps_imag))


def step_refraction(grid: Grid, *, eps_real, eps_imag, z_min: float, z_max: float):
    _, _, Z = grid.host_coords()
    return _masked(grid, (Z >= z_min) & (Z <= z_max), _eta_matrix(eps_real, eps_imag))


def ellipsis_refraction(grid: Grid, *, eps_real, eps_imag, center: Tuple[float, float, float],
                        radius: Tuple[float, float, float]):
    X, Y, Z = grid.host_coords()
    cx, cy, cz = center
    rx, ry, rz = radius
    metric = ((X - cx) / rx) ** 2 + ((Y - cy) / ry) ** 2 + ((Z - cz) / rz) ** 2
    return _masked(grid, metric <= 1.0, _eta_matrix(eps_real, eps_imag))


def apply_refraction(grid: Grid, *, scalar_eta=None, tensor_eta=None):
    be = grid.backend
    if (scalar_eta is None) == (tensor_eta is None):
        raise ValueError("apply_refraction requires exactly one of scalar_eta or tensor_eta")
    if scalar_eta is not None:
        if tuple(scalar_eta.shape) != tuple(grid.N):
            raise ValueError(f"scalar_eta.shape {tuple(scalar_eta.shape)} != grid.N {grid.N}")
        out = be.zeros((3, 3) + tuple(grid.N), kind="complex")
        s = be.array(scalar_eta, dtype=be.complex_dtype)
        for i in range(3):
            out[i, i] = s
        return out
    if tuple(tensor_eta.shape) != (3, 3) + tuple(grid.N):
        raise ValueError(f"tensor_eta.shape {tuple(tensor_eta.shape)} != {(3, 3) + tuple(grid.N)}")
    return be.array(tensor_eta, dtype=be.complex_dtype)
