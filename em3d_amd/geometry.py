"""Material sampling on the grid (src/em3d/geometry.py:11-252 of the reference).

Each shape answers three questions about a grid given by its cell-centre axes:
which cell centres lie inside it (``cell_inside``), what fraction of each cell
it covers (``cell_fraction``) and its exact volume (``volume``).  Both shapes
are products or sums of one-axis terms, so they are evaluated on the three
axes and combined by broadcasting, never on a meshgrid.  The tests keep the
reference's semantics: centres are compared in float64 with the tie tolerance
64 eps max(1, L) (geometry.py:62-64), boundaries inclusive, and an ellipsoid
sums its three squared normalised offsets in x, y, z order, so masks equal
the reference's bit for bit.  The resulting contrast is built on the host in
complex128 and copied to the device once.
"""
from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from math import pi
from typing import Sequence, Union

import numpy as np

from .grid import Grid


class SamplingMode(str, Enum):
    CELL_CENTER = "cell_center"
    VOLUME_FRACTION = "volume_fraction"


def _outer(parts, op):
    """Combine three per-axis vectors into an (Nx, Ny, Nz) array with ``op``,
    x first: op(op(a, b), c)."""
    a, b, c = parts
    return op(op(a[:, None, None], b[None, :, None]), c[None, None, :])


@dataclass(frozen=True)
class FullDomain:
    """The whole computational box."""

    def cell_inside(self, axes: Sequence[np.ndarray], tol: float) -> np.ndarray:
        return np.ones(tuple(len(a) for a in axes), dtype=bool)

    def cell_fraction(self, axes: Sequence[np.ndarray], steps) -> np.ndarray:
        return np.ones(tuple(len(a) for a in axes), dtype=np.float64)

    def volume(self, L) -> float:
        return float(np.prod(L))


@dataclass(frozen=True)
class AxisAlignedBox:
    center: tuple[float, float, float]
    size: tuple[float, float, float]

    def __post_init__(self) -> None:
        if any(float(v) <= 0.0 for v in self.size):
            raise ValueError(f"box sizes must be positive, got {self.size!r}")

    def cell_inside(self, axes, tol):
        # |x - c| <= s/2 + tol on every axis: a product of three 1-D tests
        half = [float(s) / 2.0 + tol for s in self.size]
        return _outer([np.abs(axes[a] - float(self.center[a])) <= half[a] for a in range(3)], np.logical_and)

    def cell_fraction(self, axes, steps):
        # the overlap of a cell with the box is the product of its three 1-D overlaps
        parts = []
        for a in range(3):
            h, c, s = float(steps[a]), float(self.center[a]), float(self.size[a])
            lo = np.maximum(axes[a] - h / 2.0, c - s / 2.0)
            hi = np.minimum(axes[a] + h / 2.0, c + s / 2.0)
            parts.append(np.clip(np.maximum(0.0, hi - lo) / h, 0.0, 1.0))
        return _outer(parts, np.multiply)

    def volume(self, L) -> float:
        return float(np.prod(self.size))


@dataclass(frozen=True)
class Ellipsoid:
    center: tuple[float, float, float]
    radii: tuple[float, float, float]

    def __post_init__(self) -> None:
        if any(float(v) <= 0.0 for v in self.radii):
            raise ValueError(f"ellipsoid radii must be positive, got {self.radii!r}")

    def cell_inside(self, axes, tol):
        # sum of the squared normalised offsets, x + y first, then + z
        q = [((axes[a] - float(self.center[a])) / float(self.radii[a])) ** 2 for a in range(3)]
        return _outer(q, np.add) <= 1.0 + tol

    def cell_fraction(self, axes, steps):
        raise NotImplementedError("Ellipsoid has no volume-fraction rule; sample it at cell centres "
                                  "(SamplingMode.CELL_CENTER)")

    def volume(self, L) -> float:
        return float(4.0 * pi * np.prod(self.radii) / 3.0)


Geometry = Union[FullDomain, AxisAlignedBox, Ellipsoid]


def _shape(geometry):
    if not isinstance(geometry, (FullDomain, AxisAlignedBox, Ellipsoid)):
        raise TypeError(f"unsupported geometry {type(geometry)!r}")
    return geometry


def _host_axes(grid: Grid):
    """Cell-centre axes in float64 (the grid keeps them in its real dtype)."""
    return [np.asarray(a, dtype=np.float64) for a in grid._host]


def _tolerance(grid: Grid) -> float:
    return 64.0 * np.finfo(np.float64).eps * max(1.0, *(abs(float(v)) for v in grid.L))


def center_mask(geometry: Geometry, grid: Grid) -> np.ndarray:
    """Host mask of the cell centres inside ``geometry`` (geometry.py:67-93)."""
    return _shape(geometry).cell_inside(_host_axes(grid), _tolerance(grid))


def volume_fractions(geometry: Geometry, grid: Grid) -> np.ndarray:
    """Fraction of every cell that ``geometry`` covers (geometry.py:114-148)."""
    return _shape(geometry).cell_fraction(_host_axes(grid), grid.h)


def exact_volume(geometry: Geometry, grid: Grid) -> float:
    return _shape(geometry).volume(grid.L)


def sampling_weights(geometry: Geometry, grid: Grid, mode: SamplingMode = SamplingMode.CELL_CENTER) -> np.ndarray:
    if SamplingMode(mode) is SamplingMode.CELL_CENTER:
        return center_mask(geometry, grid).astype(np.float64)
    return volume_fractions(geometry, grid)


def _material(value, name):
    a = np.asarray(value, dtype=np.complex128)
    if a.ndim == 0:
        return complex(a) * np.eye(3, dtype=np.complex128)
    if a.shape != (3, 3):
        raise ValueError(f"{name} must be a scalar or a (3, 3) tensor, got {a.shape}")
    return a


def sample_relative_permittivity(grid: Grid, *, background_eps_r, feature_eps_r, geometry: Geometry,
                                 mode: SamplingMode = SamplingMode.CELL_CENTER):
    bg = _material(background_eps_r, "background_eps_r")
    ft = _material(feature_eps_r, "feature_eps_r")
    w = sampling_weights(geometry, grid, mode)
    eps = bg[:, :, None, None, None] + (ft - bg)[:, :, None, None, None] * w[None, None, ...]
    return grid.backend.array(eps, dtype=grid.backend.complex_dtype)


def sample_contrast_tensor(grid: Grid, *, background_eps_r, feature_eps_r, geometry: Geometry,
                           mode: SamplingMode = SamplingMode.CELL_CENTER):
    """``chi = eps_r - I`` on the device (geometry.py:233-252)."""
    eps = sample_relative_permittivity(grid, background_eps_r=background_eps_r, feature_eps_r=feature_eps_r,
                                       geometry=geometry, mode=mode)
    ident = grid.backend.array(np.eye(3, dtype=np.complex128)[:, :, None, None, None],
                               dtype=grid.backend.complex_dtype)
    return eps - ident


def _eps_record(value) -> np.ndarray:
    m = _material(value, "eps_r")
    out = np.empty(18, dtype=np.float64)
    out[0::2] = m.real.ravel()
    out[1::2] = m.imag.ravel()
    return out


def ellipsoids_contrast(grid: Grid, inclusions, *, background_eps_r=1.0, kind: str = "full"):
    """Device contrast ``chi = eps_r - I`` of a background holding ellipsoidal
    inclusions ``[(Ellipsoid, eps_r), ...]`` (eps_r scalar or 3x3), tested at
    cell centres exactly as ``center_mask`` (FP64, same tolerance), later
    inclusions overwriting earlier ones; computed by one HIP launch.

    ``kind="full"`` returns the (3, 3, Nx, Ny, Nz) tensor ``Problem`` takes;
    "diag" and "iso" return the packed (3, Nx, Ny, Nz) / (Nx, Ny, Nz) forms the
    kernels read (valid only when every eps_r is diagonal / scalar)."""
    import torch

    from . import _lib
    be = grid.backend
    codes = {"iso": 1, "diag": 2, "full": 3}
    if kind not in codes:
        raise ValueError(f"kind must be one of {sorted(codes)}, got {kind!r}")
    recs = []
    for geom, eps in inclusions:
        if not isinstance(geom, Ellipsoid):
            raise TypeError(f"inclusions must be Ellipsoid geometries, got {type(geom)!r}")
        recs.append(np.concatenate([np.asarray(geom.center, np.float64), np.asarray(geom.radii, np.float64),
                                    _eps_record(eps)]))
    dev = be.torch_device
    ell = torch.as_tensor(np.ascontiguousarray(np.array(recs, dtype=np.float64).reshape(-1, 24))).to(dev)
    bg = torch.as_tensor(_eps_record(background_eps_r)).to(dev)
    shape = {"iso": (), "diag": (3,), "full": (3, 3)}[kind] + tuple(grid.N)
    chi = torch.empty(shape, dtype=be.precision.torch_complex_dtype, device=dev)
    _lib.check(_lib.lib().em3d_ellipsoid_contrast(*grid.N, grid.x.data_ptr(), grid.y.data_ptr(), grid.z.data_ptr(),
                                                  be.precision.code, len(recs), ell.data_ptr() if recs else None,
                                                  bg.data_ptr(), _tolerance(grid), codes[kind], chi.data_ptr(),
                                                  _lib.stream_handle(dev)), "em3d_ellipsoid_contrast")
    return chi
