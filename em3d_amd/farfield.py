"""Far-field scatter integral and RCS (src/em3d/farfield/__init__.py:1-125,
src/em3d/farfield/_core.py:8-71 of the reference).

``scatter_integral`` keeps the reference's signature and returns a host
(M, 3) complex128 array; the sum runs in libem3d_hip (``em3d_scatter_integral``,
FP64 accumulation) for one field or a batch of fields at once.
"""
from __future__ import annotations

import weakref

import numpy as np

from . import _lib
from .operator import pack_contrast, require_hip

_PACKED: dict = {}
_SELF = object()   # stands for the key tensor inside a cached value


def packed_contrast(t, make):
    """make(t), kept while t is alive and unchanged: the same tensor object
    with the same version counter (torch bumps it on every in-place change).
    Entries are keyed by (t, make), so ``make`` must be a module-level
    function (one cache line per packing rule).  The entry is dropped when t
    is freed.  Repeated far-field calls and operator applications on one
    problem then pack the contrast once.  A value that is t itself, or a
    tuple holding it (a full 3 x 3 contrast packs to itself), is stored
    without that reference, so the cache never keeps t alive."""
    if not hasattr(t, "_version"):
        return make(t)
    key = (id(t), make)
    hit = _PACKED.get(key)
    if hit is not None and hit[0]() is t and hit[1] == t._version:
        v = hit[2]
        if isinstance(v, tuple):
            return tuple(t if x is _SELF else x for x in v)
        return t if v is _SELF else v
    val = make(t)
    if hit is None:
        weakref.finalize(t, _PACKED.pop, key, None)
    kept = tuple(_SELF if x is t else x for x in val) if isinstance(val, tuple) else (_SELF if val is t else val)
    _PACKED[key] = (weakref.ref(t), t._version, kept)
    return val

__all__ = ["scatter_integral", "scatter_integral_device", "rcs", "rcs_plane", "plane_directions"]


def plane_directions(plane: str, n_phi: int):
    """(phi, unit directions) of n_phi equally spaced angles in a coordinate plane."""
    phi = np.linspace(0.0, 2.0 * np.pi, int(n_phi), endpoint=False)
    c, s, z = np.cos(phi), np.sin(phi), np.zeros_like(phi)
    if plane == "xy":
        dirs = np.stack([c, s, z], axis=-1)
    elif plane == "yz":
        dirs = np.stack([z, c, s], axis=-1)
    elif plane == "xz":
        dirs = np.stack([c, z, s], axis=-1)
    else:
        raise ValueError(f"plane must be 'xy'|'yz'|'xz', got {plane!r}")
    return phi, dirs


def shared_axis(dirs) -> int:
    """The axis (2, 0, 1 tried in that order) along which all directions have
    exactly the same component, else -1.  Plane cuts have one: the kernel then
    sums that axis once for all directions."""
    for a in (2, 0, 1):
        if np.all(dirs[:, a] == dirs[0, a]):
            return a
    return -1


def _integral(u, grid, k0: float, chi_kind: int, chi, ncomp: int, directions):
    """Device (batch, M, ncomp) complex128 far-field integrals of fields u."""
    import torch
    be = grid.backend
    require_hip(be)
    dirs = np.asarray(directions, dtype=np.float64)
    if dirs.ndim == 1:
        dirs = dirs[np.newaxis, :]
    if dirs.ndim != 2 or dirs.shape[1] != 3:
        raise ValueError(f"directions must have shape (M, 3), got {dirs.shape}")
    t = u if isinstance(u, torch.Tensor) else be.array(u)
    t = t.to(device=be.torch_device, dtype=be.precision.torch_complex_dtype).contiguous()
    fshape = ((3,) if ncomp == 3 else ()) + tuple(grid.N)
    if tuple(t.shape) == fshape:
        batch = 1
    elif t.dim() == len(fshape) + 1 and tuple(t.shape[1:]) == fshape:
        batch = int(t.shape[0])
    else:
        raise ValueError(f"u must have shape {fshape} (or (batch,) + {fshape}), got {tuple(t.shape)}")
    dev = be.torch_device
    d = torch.as_tensor(np.ascontiguousarray(dirs)).to(dev)
    M = dirs.shape[0]
    out = torch.empty((batch, M, ncomp), dtype=torch.complex128, device=dev)
    lib = _lib.lib()
    work = torch.empty(int(lib.em3d_scatter_work_bytes(*grid.N, batch, M, ncomp)), dtype=torch.uint8, device=dev)
    _lib.check(lib.em3d_scatter_integral(*grid.N, grid.x.data_ptr(), grid.y.data_ptr(), grid.z.data_ptr(),
                                         be.precision.code, batch, ncomp, chi.data_ptr(), chi_kind, t.data_ptr(), M,
                                         d.data_ptr(), shared_axis(dirs), float(k0), float(grid.dv), out.data_ptr(),
                                         work.data_ptr(), _lib.stream_handle(dev)), "em3d_scatter_integral")
    return out


def scatter_integral_device(u, problem, directions):
    """F[b][m] = dv sum_q (chi u_b)_q exp(-i k0 d_m . r_q) as a device tensor
    (batch, M, 3) complex128 (u: (3, N) or (batch, 3, N))."""
    kind, chi = packed_contrast(problem.eps_tensor, pack_contrast)
    return _integral(u, problem.grid, problem.k0, kind, chi, 3, directions)


def scatter_integral(u, problem, directions, *, method: str = "direct", batch_size: int = 64) -> np.ndarray:
    """Far-field scatter integral (M, 3) complex128 on the host (batched u:
    (batch, M, 3)).  ``batch_size`` is accepted for signature compatibility;
    the device kernel tiles directions itself."""
    if method != "direct":
        raise ValueError(f"method must be 'direct' on the HIP backend, got {method!r}")
    if batch_size <= 0:
        raise ValueError(f"batch_size must be positive, got {batch_size}")
    F = scatter_integral_device(u, problem, directions).cpu().numpy()
    shape = getattr(u, "shape", None)
    return F[0] if shape is not None and len(shape) == 4 else F


def rcs(u, problem, direction) -> float:
    """k0^4 / (16 pi^2) |d x F(d)|^2 for one direction (normalised here)."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    F = scatter_integral(u, problem, d[np.newaxis, :])[0]
    cross = np.cross(d, F)
    return float(problem.k0 ** 4 / (16.0 * np.pi ** 2) * np.real(np.dot(cross, cross.conj())))


def rcs_plane(u, problem, n_phi: int = 80, plane: str = "xy", method: str = "direct", batch_size: int = 64):
    """(phi, sigma) over n_phi directions of a coordinate plane; for a batch
    of fields sigma has shape (batch, n_phi)."""
    phi, dirs = plane_directions(plane, n_phi)
    F = scatter_integral(u, problem, dirs, method=method, batch_size=batch_size)
    cross = np.cross(dirs, F)
    sigma = problem.k0 ** 4 / (16.0 * np.pi ** 2) * np.real(np.sum(cross * cross.conj(), axis=-1))
    return phi, sigma
