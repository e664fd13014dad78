"""Field flattening order (src/em3d/layout.py:14-45 of the reference)."""
from __future__ import annotations

from math import prod
from typing import Any, Sequence


def _normalize_shape(shape: Sequence[int]) -> tuple[int, int, int]:
    normalized = tuple(int(n) for n in shape)
    if len(normalized) != 3 or any(n <= 0 for n in normalized):
        raise ValueError(f"shape must be a positive 3-tuple, got {shape!r}")
    return normalized


def flatten_field(field: Any) -> Any:
    """(3, Nx, Ny, Nz) -> (ix, iy, iz, component) order."""
    if getattr(field, "ndim", None) != 4 or field.shape[0] != 3:
        raise ValueError(f"field must have shape (3, Nx, Ny, Nz), got {getattr(field, 'shape', None)!r}")
    if hasattr(field, "permute"):
        return field.permute(1, 2, 3, 0).reshape(-1)
    return field.transpose(1, 2, 3, 0).reshape(-1)


def unflatten_field(vector: Any, shape: Sequence[int]) -> Any:
    Nx, Ny, Nz = _normalize_shape(shape)
    expected = 3 * prod((Nx, Ny, Nz))
    n = int(vector.numel()) if hasattr(vector, "numel") else int(vector.size)
    if getattr(vector, "ndim", None) != 1 or n != expected:
        raise ValueError(f"vector must be one-dimensional with {expected} entries, got shape {getattr(vector, 'shape', None)!r}")
    v = vector.reshape(Nx, Ny, Nz, 3)
    return v.permute(3, 0, 1, 2) if hasattr(v, "permute") else v.transpose(3, 0, 1, 2)


def field_size(shape: Sequence[int]) -> int:
    return 3 * prod(_normalize_shape(shape))
