"""CPU oracle of the far-field scatter integral (TEST INFRASTRUCTURE ONLY).

Restates src/em3d/farfield/_core.py:8-71 (EM) and
src/em3d/acoustics/farfield.py:29-60 (scalar) as a dense NumPy product.
Pinned by tests/test_oracle_farfield.py: the single-cell closed form
F = dv (chi u) exp(-i k0 d . r), zero contrast, linearity in u.
"""
from __future__ import annotations

import numpy as np

from .inputs import coords


def scatter_integral(u, chi, N, L, center, k0: float, directions) -> np.ndarray:
    """F[m, i] = dv sum_q (chi u)_i(r_q) exp(-i k0 d_m . r_q); chi (3, 3, N) or scalar field."""
    X, Y, Z = coords(N, L, center, "double")
    r = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=0)
    dv = float(np.prod(np.asarray(L, dtype=np.float64) / np.asarray(N, dtype=np.float64)))
    d = np.asarray(directions, dtype=np.float64).reshape(-1, 3)
    phase = np.exp(-1j * k0 * (d @ r))                                   # (M, cells)
    u = np.asarray(u, dtype=np.complex128)
    if u.ndim == 3:                                                       # scalar
        return dv * (phase @ (np.asarray(chi) * u).ravel())
    pol = np.einsum("ij...,j...->i...", np.asarray(chi, dtype=np.complex128), u).reshape(3, -1)
    return dv * (pol @ phase.T).T


def rcs_plane(F, dirs, k0: float) -> np.ndarray:
    cross = np.cross(dirs, F)
    return k0 ** 4 / (16.0 * np.pi ** 2) * np.real(np.sum(cross * cross.conj(), axis=-1))
