"""Device-generated right-hand sides: batches of incident plane waves.

``plane_waves`` evaluates many ``flat_wave_vec`` fields (src/em3d/wave.py:9-37
of the reference, one field per call there) in one HIP launch, straight into
a (batch, 3, Nx, Ny, Nz) device tensor, the layout of a batched solve.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .grid import Grid


def _axes(grid: Grid):
    return grid.x, grid.y, grid.z


def plane_waves(grid: Grid, *, k: float, orients, amplitudes, phi0=0.0, sign: int = 1):
    """``A_b exp(i sign (k o_b.r + phi0_b))`` for every row b of ``orients``
    (unit 3-vectors) and ``amplitudes`` (complex 3-vectors); phase in FP64."""
    import torch
    be = grid.backend
    o = np.asarray(orients, dtype=np.float64).reshape(-1, 3)
    a = np.asarray(amplitudes, dtype=np.complex128).reshape(-1, 3)
    if o.shape[0] != a.shape[0]:
        raise ValueError(f"orients ({o.shape[0]}) and amplitudes ({a.shape[0]}) differ in count")
    if sign not in (1, -1):
        raise ValueError(f"sign must be 1 or -1, got {sign!r}")
    n2 = (o ** 2).sum(axis=1)
    if np.any(np.abs(n2 - 1.0) > 1e-9):
        raise ValueError("every orient must be a unit vector")
    B = o.shape[0]
    ph = np.broadcast_to(np.asarray(phi0, dtype=np.float64), (B,))
    prm = np.concatenate([o, a.real, a.imag, ph[:, None]], axis=1)
    prm_d = torch.as_tensor(np.ascontiguousarray(prm)).to(be.torch_device)
    out = torch.empty((B, 3) + tuple(grid.N), dtype=be.precision.torch_complex_dtype, device=be.torch_device)
    x, y, z = _axes(grid)
    _lib.check(_lib.lib().em3d_plane_waves(*grid.N, x.data_ptr(), y.data_ptr(), z.data_ptr(), be.precision.code, B,
                                           prm_d.data_ptr(), float(k), int(sign), out.data_ptr(),
                                           _lib.stream_handle(be.torch_device)), "em3d_plane_waves")
    return out


def incident_set(n_directions: int, seed: int):
    """``n_directions`` seeded incident directions on the unit sphere
    (normalised Gaussian 3-vectors) times two orthonormal transverse
    polarisations each: (orients, amplitudes) of 2 n_directions plane waves,
    direction-major."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n_directions, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ref = np.where(np.abs(d[:, 2:3]) < 0.9, np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]]))
    p1 = np.cross(d, ref)
    p1 /= np.linalg.norm(p1, axis=1, keepdims=True)
    p2 = np.cross(d, p1)
    orients = np.repeat(d, 2, axis=0)
    amps = np.empty((2 * n_directions, 3), dtype=np.complex128)
    amps[0::2] = p1
    amps[1::2] = p2
    return orients, amps
