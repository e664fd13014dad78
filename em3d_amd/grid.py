"""Structured Cartesian grid (src/em3d/grid.py:10-52 of the reference).

Cell-centre axes are computed on the host in the precision's real dtype with
``numpy.linspace`` exactly as the reference does, then copied to the device,
so device inputs equal the reference's NumPy inputs bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Tuple

import numpy as np

from .backend import Backend


def host_axis(n: int, length: float, centre: float, real_dtype) -> np.ndarray:
    step = length / n
    start = centre - length / 2 + step / 2
    return np.linspace(start, start + step * (n - 1), n, dtype=real_dtype)


@dataclass(frozen=True)
class Grid:
    N: Tuple[int, int, int]
    L: Tuple[float, float, float]
    center: Tuple[float, float, float]
    backend: Backend
    _x: object = field(init=False, repr=False)
    _y: object = field(init=False, repr=False)
    _z: object = field(init=False, repr=False)
    _host: tuple = field(init=False, repr=False)

    def __post_init__(self) -> None:
        if any(n <= 0 for n in self.N):
            raise ValueError(f"Grid.N must be strictly positive, got {self.N}")
        if any(l <= 0 for l in self.L):
            raise ValueError(f"Grid.L must be strictly positive, got {self.L}")
        object.__setattr__(self, "N", tuple(int(n) for n in self.N))
        rd = self.backend.real_dtype
        host = tuple(host_axis(self.N[a], self.L[a], self.center[a], rd) for a in range(3))
        object.__setattr__(self, "_host", host)
        object.__setattr__(self, "_x", self.backend.array(host[0]))
        object.__setattr__(self, "_y", self.backend.array(host[1]))
        object.__setattr__(self, "_z", self.backend.array(host[2]))

    @property
    def dv(self) -> float:
        return (self.L[0] / self.N[0]) * (self.L[1] / self.N[1]) * (self.L[2] / self.N[2])

    @property
    def h(self) -> tuple[float, float, float]:
        return tuple(self.L[a] / self.N[a] for a in range(3))

    @property
    def x(self):
        return self._x

    @property
    def y(self):
        return self._y

    @property
    def z(self):
        return self._z

    def coords(self):
        return self.backend.xp.meshgrid(self._x, self._y, self._z, indexing="ij")

    def host_coords(self):
        """``ij`` meshgrid of the host axes (NumPy, real dtype)."""
        return np.meshgrid(*self._host, indexing="ij")
