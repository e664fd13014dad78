"""Scalar far field (src/em3d/acoustics/farfield.py:1-96 of the reference):
f(d) = k0^2 dv / (4 pi) sum_q chi_q u_q exp(-i k0 d . r_q), computed by
libem3d_hip's ``em3d_scatter_integral`` with one component."""
from __future__ import annotations

import numpy as np

from ..farfield import _integral, packed_contrast
from ..operator import CHI_ISO, eta_contrast


def _directions_array(directions) -> np.ndarray:
    dirs = np.asarray(directions, dtype=np.float64)
    if dirs.ndim != 2 or dirs.shape[1] != 3:
        raise ValueError(f"directions must have shape (M, 3), got {dirs.shape}")
    norms = np.linalg.norm(dirs, axis=1)
    if np.any(norms == 0.0):
        raise ValueError("directions must be non-zero")
    return dirs / norms[:, None]


def _plane_directions(plane: str, phi: np.ndarray) -> np.ndarray:
    if plane == "xy":
        return np.column_stack([np.cos(phi), np.sin(phi), np.zeros_like(phi)])
    if plane == "xz":
        return np.column_stack([np.cos(phi), np.zeros_like(phi), np.sin(phi)])
    if plane == "yz":
        return np.column_stack([np.zeros_like(phi), np.cos(phi), np.sin(phi)])
    raise ValueError(f"plane must be 'xy', 'xz', or 'yz', got {plane!r}")


def farfield_amplitude(u, problem, directions, *, batch_size: int = 64) -> np.ndarray:
    """Scalar far-field amplitude for observation directions (host complex128)."""
    dirs = _directions_array(directions)
    if int(batch_size) <= 0:
        raise ValueError(f"batch_size must be positive, got {batch_size!r}")
    grid = problem.grid
    shape = getattr(u, "shape", None)
    if shape is not None and tuple(shape) != tuple(grid.N):
        raise ValueError(f"u.shape {tuple(shape)} != expected {tuple(grid.N)}")
    _, chi = packed_contrast(problem.eta, eta_contrast)   # problem.chi, packed once
    F = _integral(u, grid, problem.k0, CHI_ISO, chi, 1, dirs)[0, :, 0]
    coeff = float(problem.k0) ** 2 / (4.0 * np.pi)     # dv is applied by the integral
    return coeff * F.cpu().numpy()


def scattering_pattern(u, problem, directions, *, normalize: str | None = "max") -> np.ndarray:
    """|f(d)|^2 with optional max normalisation."""
    sigma = np.abs(farfield_amplitude(u, problem, directions)) ** 2
    if normalize is None:
        return sigma.astype(np.float64, copy=False)
    if normalize != "max":
        raise ValueError(f"normalize must be None or 'max', got {normalize!r}")
    peak = float(np.max(sigma)) if sigma.size else 0.0
    if peak > 0.0:
        sigma = sigma / peak
    return sigma.astype(np.float64, copy=False)


def pattern_plane(u, problem, *, plane: str = "xy", n_angles: int = 360, normalize: str | None = "max"):
    """Angle grid and scattering pattern in a coordinate plane."""
    if int(n_angles) <= 0:
        raise ValueError(f"n_angles must be positive, got {n_angles!r}")
    phi = np.linspace(0.0, 2.0 * np.pi, int(n_angles), endpoint=False)
    return phi, scattering_pattern(u, problem, _plane_directions(plane, phi), normalize=normalize)
