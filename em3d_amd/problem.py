"""EM scattering problem container (src/em3d/problem.py:7-46 of the reference)."""
from __future__ import annotations

from dataclasses import dataclass

from .grid import Grid


def _dtype_matches(arr, backend) -> bool:
    import torch
    if isinstance(arr, torch.Tensor):
        return arr.dtype == backend.precision.torch_complex_dtype
    return getattr(arr, "dtype", None) == backend.complex_dtype


@dataclass(frozen=True)
class Problem:
    grid: Grid
    eps_tensor: object
    wave: object
    k0: float
    volume: float

    def __post_init__(self) -> None:
        be = self.grid.backend
        expected_eta = (3, 3) + tuple(self.grid.N)
        expected_wave = (3,) + tuple(self.grid.N)
        if tuple(self.eps_tensor.shape) != expected_eta:
            raise ValueError(f"eps_tensor.shape {tuple(self.eps_tensor.shape)} != expected {expected_eta}")
        if tuple(self.wave.shape) != expected_wave:
            raise ValueError(f"wave.shape {tuple(self.wave.shape)} != expected {expected_wave}")
        if not _dtype_matches(self.eps_tensor, be):
            raise TypeError(f"eps_tensor.dtype {self.eps_tensor.dtype} != {be.complex_dtype}")
        if not _dtype_matches(self.wave, be):
            raise TypeError(f"wave.dtype {self.wave.dtype} != {be.complex_dtype}")

    @property
    def backend(self):
        return self.grid.backend

    @property
    def contrast_tensor(self):
        """Canonical name of the stored tensor ``chi = eps_r - I``."""
        return self.eps_tensor
