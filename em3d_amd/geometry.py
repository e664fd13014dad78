"""Material sampling on the grid (src/em3d/geometry.py:11-252 of the reference).

Masks and weights are evaluated on the host in float64 (as the reference does
for every backend) and the resulting contrast is copied to the device once.
"""
from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from math import pi
from typing import Union

import numpy as np

from .grid import Grid


class SamplingMode(str, Enum):
    CELL_CENTER = "cell_center"
    VOLUME_FRACTION = "volume_fraction"


@dataclass(frozen=True)
class FullDomain:
    pass


@dataclass(frozen=True)
class AxisAlignedBox:
    center: tuple[float, float, float]
    size: tuple[float, float, float]

    def __post_init__(self) -> None:
        if any(float(v) <= 0.0 for v in self.size):
            raise ValueError(f"box sizes must be positive, got {self.size!r}")


@dataclass(frozen=True)
class Ellipsoid:
    center: tuple[float, float, float]
    radii: tuple[float, float, float]

    def __post_init__(self) -> None:
        if any(float(v) <= 0.0 for v in self.radii):
            raise ValueError(f"ellipsoid radii must be positive, got {self.radii!r}")


Geometry = Union[FullDomain, AxisAlignedBox, Ellipsoid]


def _host_axes(grid: Grid):
    return [np.asarray(a, dtype=np.float64) for a in grid._host]


def _tolerance(grid: Grid) -> float:
    scale = max(1.0, *(abs(float(v)) for v in grid.L))
    return 64.0 * np.finfo(np.float64).eps * scale


def center_mask(geometry: Geometry, grid: Grid) -> np.ndarray:
    """Host mask of cell centres inside ``geometry`` (geometry.py:67-93)."""
    if isinstance(geometry, FullDomain):
        return np.ones(grid.N, dtype=bool)
    X, Y, Z = np.meshgrid(*_host_axes(grid), indexing="ij")
    tol = _tolerance(grid)
    if isinstance(geometry, AxisAlignedBox):
        out = np.ones(grid.N, dtype=bool)
        for c, c0, s in zip((X, Y, Z), geometry.center, geometry.size):
            out &= np.abs(c - float(c0)) <= float(s) / 2.0 + tol
        return out
    if isinstance(geometry, Ellipsoid):
        v = np.zeros(grid.N, dtype=np.float64)
        for c, c0, r in zip((X, Y, Z), geometry.center, geometry.radii):
            v += ((c - float(c0)) / float(r)) ** 2
        return v <= 1.0 + tol
    raise TypeError(f"unsupported geometry {type(geometry)!r}")


def _overlap(centres, step, fc, fs):
    left, right = centres - step / 2.0, centres + step / 2.0
    ov = np.maximum(0.0, np.minimum(right, fc + fs / 2.0) - np.maximum(left, fc - fs / 2.0))
    return np.clip(ov / step, 0.0, 1.0)


def volume_fractions(geometry: Geometry, grid: Grid) -> np.ndarray:
    """Cell occupation fractions (geometry.py:114-148)."""
    if isinstance(geometry, FullDomain):
        return np.ones(grid.N, dtype=np.float64)
    if isinstance(geometry, AxisAlignedBox):
        axes = _host_axes(grid)
        f = [_overlap(axes[a], float(grid.L[a]) / grid.N[a], float(geometry.center[a]), float(geometry.size[a]))
             for a in range(3)]
        return f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :]
    if isinstance(geometry, Ellipsoid):
        raise NotImplementedError("volume-fraction sampling for Ellipsoid is not implemented; "
                                  "use SamplingMode.CELL_CENTER or add an explicit quadrature rule")
    raise TypeError(f"unsupported geometry {type(geometry)!r}")


def exact_volume(geometry: Geometry, grid: Grid) -> float:
    if isinstance(geometry, FullDomain):
        return float(np.prod(grid.L))
    if isinstance(geometry, AxisAlignedBox):
        return float(np.prod(geometry.size))
    if isinstance(geometry, Ellipsoid):
        return float(4.0 * pi * np.prod(geometry.radii) / 3.0)
    raise TypeError(f"unsupported geometry {type(geometry)!r}")


def sampling_weights(geometry: Geometry, grid: Grid, mode: SamplingMode = SamplingMode.CELL_CENTER) -> np.ndarray:
    mode = SamplingMode(mode)
    if mode is SamplingMode.CELL_CENTER:
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
