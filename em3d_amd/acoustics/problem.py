"""Scalar acoustic problem (src/em3d/acoustics/problem.py:11-101 of the reference)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..grid import Grid
from ..problem import _dtype_matches


def _direction(direction, real_dtype):
    vals = tuple(float(x) for x in direction)
    if len(vals) != 3:
        raise ValueError(f"direction must have length 3, got {direction!r}")
    d = np.asarray(vals, dtype=real_dtype)
    n = float(np.linalg.norm(d))
    if n == 0.0:
        raise ValueError("direction must be non-zero")
    return d / n


def plane_wave_scalar(grid: Grid, *, k: float, direction=(0.0, 0.0, 1.0), amplitude: complex = 1.0):
    """``amplitude exp(i k d.x)`` evaluated on the host in the backend dtypes."""
    be = grid.backend
    d = _direction(direction, be.real_dtype)
    X, Y, Z = grid.host_coords()
    phase = d[0] * X + d[1] * Y + d[2] * Z
    ik = be.complex_dtype(1j * float(k))
    wave = (be.complex_dtype(amplitude) * np.exp(ik * phase)).astype(be.complex_dtype, copy=False)
    return be.array(wave)


@dataclass(frozen=True)
class AcousticProblem:
    grid: Grid
    eta: object
    wave: object
    k0: float
    volume: float

    def __post_init__(self) -> None:
        be = self.grid.backend
        expected = tuple(self.grid.N)
        if tuple(getattr(self.eta, "shape", ())) != expected:
            raise ValueError(f"eta.shape {tuple(getattr(self.eta, 'shape', ()))} != expected {expected}")
        if tuple(getattr(self.wave, "shape", ())) != expected:
            raise ValueError(f"wave.shape {tuple(getattr(self.wave, 'shape', ()))} != expected {expected}")
        if not _dtype_matches(self.eta, be):
            raise TypeError(f"eta.dtype {self.eta.dtype} != {be.complex_dtype}")
        if not _dtype_matches(self.wave, be):
            raise TypeError(f"wave.dtype {self.wave.dtype} != {be.complex_dtype}")
        if float(self.k0) <= 0.0:
            raise ValueError(f"k0 must be positive, got {self.k0}")
        if float(self.volume) <= 0.0:
            raise ValueError(f"volume must be positive, got {self.volume}")

    @property
    def backend(self):
        return self.grid.backend

    @property
    def chi(self):
        """Acoustic contrast ``eta - 1``."""
        return self.eta - 1.0


def make_acoustic_problem(grid: Grid, eta, *, k0: float, direction=(0.0, 0.0, 1.0),
                          amplitude: complex = 1.0) -> AcousticProblem:
    be = grid.backend
    eta_arr = be.asarray_of_kind(eta, "complex")
    wave = plane_wave_scalar(grid, k=float(k0), direction=direction, amplitude=amplitude)
    n_cells = int(grid.N[0] * grid.N[1] * grid.N[2])
    return AcousticProblem(grid=grid, eta=eta_arr, wave=wave, k0=float(k0), volume=grid.dv * n_cells)
