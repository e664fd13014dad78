"""Oracle restatement of the reference's input builders (grid, waves, materials).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Plain NumPy on the host.
All arrays follow the reference layout: fields ``(3, Nx, Ny, Nz)``, contrast
``(3, 3, Nx, Ny, Nz)``, scalar fields ``(Nx, Ny, Nz)``, C order (z fastest).
"""
from __future__ import annotations

import numpy as np

REAL = {"double": np.float64, "single": np.float32}
CPLX = {"double": np.complex128, "single": np.complex64}


def axis(n: int, length: float, centre: float, precision: str = "double") -> np.ndarray:
    """Cell centres ``c - L/2 + h/2 + i h`` in the real dtype (src/em3d/grid.py:10-13)."""
    step = length / n
    start = centre - length / 2 + step / 2
    return np.linspace(start, start + step * (n - 1), n, dtype=REAL[precision])


def coords(N, L, center, precision: str = "double"):
    """``ij`` meshgrid of cell centres (src/em3d/grid.py:50-52)."""
    xs = [axis(int(N[a]), float(L[a]), float(center[a]), precision) for a in range(3)]
    return np.meshgrid(*xs, indexing="ij")


def cell_volume(N, L) -> float:
    """``dv = prod(L/N)`` in the reference's evaluation order (src/em3d/grid.py:36-38)."""
    return (L[0] / N[0]) * (L[1] / N[1]) * (L[2] / N[2])


def flat_wave_vec(N, L, center, *, k, orient, amplitude, phi0=0.0, sign=1,
                  precision: str = "double") -> np.ndarray:
    """Plane wave ``A exp(i sign (k o.r + phi0))`` (src/em3d/wave.py:9-37)."""
    if sign not in (1, -1):
        raise ValueError(f"sign must be 1 or -1, got {sign!r}")
    norm2 = float(orient[0]) ** 2 + float(orient[1]) ** 2 + float(orient[2]) ** 2
    if abs(norm2 - 1.0) > 1e-9:
        raise ValueError(f"orient must be a unit vector, got norm^2 = {norm2}")
    X, Y, Z = coords(N, L, center, precision)
    phase = sign * (k * (orient[0] * X + orient[1] * Y + orient[2] * Z) + phi0)
    carrier = np.exp(1j * phase).astype(CPLX[precision], copy=False)
    out = np.zeros((3,) + tuple(N), dtype=CPLX[precision])
    for i, a in enumerate(amplitude):
        if a != 0.0:
            out[i] = CPLX[precision](a) * carrier
    return out


def _canonical_tolerance(L) -> float:
    """Cell-centre tie tolerance ``64 eps max(1, L)`` (src/em3d/geometry.py:62-64)."""
    scale = max(1.0, *(abs(float(v)) for v in L))
    return 64.0 * np.finfo(np.float64).eps * scale


def ellipsoid_center_mask(N, L, center, *, ell_center, radii, precision="double") -> np.ndarray:
    """Cell-centre ellipsoid test with tolerance (src/em3d/geometry.py:67-93).

    The reference converts the real-dtype axes to float64 before the test
    (src/em3d/geometry.py:57-59)."""
    xs = [axis(int(N[a]), float(L[a]), float(center[a]), precision).astype(np.float64) for a in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    tol = _canonical_tolerance(L)
    value = np.zeros(tuple(N), dtype=np.float64)
    for c, c0, r in zip((X, Y, Z), ell_center, radii):
        value += ((c - float(c0)) / float(r)) ** 2
    return value <= 1.0 + tol


def box_center_mask(N, L, center, *, box_center, size, precision="double") -> np.ndarray:
    """Cell-centre axis-aligned box test (src/em3d/geometry.py:77-83)."""
    xs = [axis(int(N[a]), float(L[a]), float(center[a]), precision).astype(np.float64) for a in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    tol = _canonical_tolerance(L)
    out = np.ones(tuple(N), dtype=bool)
    for c, c0, s in zip((X, Y, Z), box_center, size):
        out &= np.abs(c - float(c0)) <= float(s) / 2.0 + tol
    return out


def _material(value) -> np.ndarray:
    """Scalar -> scalar * I, else a (3, 3) tensor (src/em3d/geometry.py:203-209)."""
    a = np.asarray(value, dtype=np.complex128)
    if a.ndim == 0:
        return complex(a) * np.eye(3, dtype=np.complex128)
    if a.shape != (3, 3):
        raise ValueError(f"material must be scalar or (3, 3), got {a.shape}")
    return a


def contrast_from_weights(weights, *, background_eps_r, feature_eps_r, precision="double") -> np.ndarray:
    """``chi = eps_bg + (eps_f - eps_bg) w - I`` (src/em3d/geometry.py:212-252)."""
    bg = _material(background_eps_r)
    ft = _material(feature_eps_r)
    eps = bg[:, :, None, None, None] + (ft - bg)[:, :, None, None, None] * np.asarray(weights, np.float64)[None, None]
    eps = eps.astype(CPLX[precision])
    ident = np.eye(3, dtype=CPLX[precision])[:, :, None, None, None]
    return (eps - ident).astype(CPLX[precision], copy=False)


def _eta_matrix(eps_real, eps_imag) -> np.ndarray:
    """(3,3) contrast eta = eps - I from real/imag parts (src/em3d/refraction.py:11-49)."""
    if np.ndim(eps_real) == 0 and np.ndim(eps_imag) == 0:
        return (float(eps_real) - 1.0 + 1j * float(eps_imag)) * np.eye(3, dtype=np.complex128)
    E = np.asarray(eps_real, dtype=np.float64)
    F = np.asarray(eps_imag, dtype=np.float64)
    return (E - np.eye(3)) + 1j * F


def _masked_tensor(mask, eta, precision):
    out = np.zeros((3, 3) + mask.shape, dtype=CPLX[precision])
    for i in range(3):
        for j in range(3):
            out[i, j] = np.where(mask, eta[i, j], out[i, j])
    return out


def ellipsis_refraction(N, L, center, *, eps_real, eps_imag, ell_center, radius, precision="double"):
    """Axis-aligned ellipsoid contrast, no tolerance (src/em3d/refraction.py:112-138)."""
    X, Y, Z = coords(N, L, center, precision)
    cx, cy, cz = ell_center
    rx, ry, rz = radius
    metric = ((X - cx) / rx) ** 2 + ((Y - cy) / ry) ** 2 + ((Z - cz) / rz) ** 2
    return _masked_tensor(metric <= 1.0, _eta_matrix(eps_real, eps_imag), precision)


def cylinder_refraction(N, L, center, *, eps_real, eps_imag, radius, axis_name="z", precision="double"):
    """Infinite cylinder contrast (src/em3d/refraction.py:52-83)."""
    X, Y, Z = coords(N, L, center, precision)
    r2 = {"z": X * X + Y * Y, "y": X * X + Z * Z, "x": Y * Y + Z * Z}[axis_name]
    return _masked_tensor(r2 <= radius * radius, _eta_matrix(eps_real, eps_imag), precision)


def eta_ellipsoid(N, L, center, *, ell_center, radii, eta_inside, eta_outside=1.0, precision="double"):
    """Scalar acoustic eta with an ellipsoidal inclusion (src/em3d/acoustics/materials.py:62-83)."""
    X, Y, Z = coords(N, L, center, precision)
    cx, cy, cz = ell_center
    rx, ry, rz = radii
    metric = ((X - cx) / rx) ** 2 + ((Y - cy) / ry) ** 2 + ((Z - cz) / rz) ** 2
    c = CPLX[precision]
    return np.where(metric <= 1.0, c(eta_inside), c(eta_outside)).astype(c, copy=False)


def plane_wave_scalar(N, L, center, *, k, direction=(0.0, 0.0, 1.0), amplitude=1.0, precision="double"):
    """Scalar plane wave ``A exp(i k d.x)`` (src/em3d/acoustics/problem.py:25-40)."""
    d = np.asarray(direction, dtype=REAL[precision])
    d = d / float(np.linalg.norm(d))
    X, Y, Z = coords(N, L, center, precision)
    phase = d[0] * X + d[1] * Y + d[2] * Z
    c = CPLX[precision]
    return (c(amplitude) * np.exp(c(1j * float(k)) * phase)).astype(c, copy=False)
