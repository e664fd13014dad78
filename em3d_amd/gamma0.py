"""SIM parameters from the spectrum of a coarse operator (src/em3d/gamma0.py:
1-221 and src/em3d/spectral/localization.py, complex_geometry.py of the
reference).

The stationary iteration u <- u - (Au - f)/mu converges when a circle of
centre mu and radius R < |mu| holds the spectrum of A.  As in the reference,
the spectrum is estimated from the operator on a coarse grid (contrast
resampled to the nearest fine cell): here the coarse matrix is assembled on
the GPU as A [e_1 ... e_n] with one batched FFT matvec (the matvec equals the
dense operator to rounding), then its eigenvalues are computed on the host
(LAPACK, numpy.linalg.eigvals) and the circle is chosen on the convex hull:
among the circles through two or three hull vertices that contain the hull
and exclude the origin, the one with the smallest ratio q = R / |mu|.
"""
from __future__ import annotations

from dataclasses import dataclass
from itertools import combinations

import numpy as np

__all__ = ["Gamma0Analysis", "analyze_spectrum", "estimate_from_problem", "find_params", "find_params_from_problem",
           "coarse_operator_matrix", "convex_hull", "circle_two_points", "circle_three_points"]


@dataclass(frozen=True)
class Gamma0Analysis:
    mu: complex
    radius: float
    rho: float
    spectrum: np.ndarray
    hull: np.ndarray
    coarse_N: tuple | None = None
    matrix_shape: tuple | None = None

    @property
    def margin(self) -> float:
        return float(abs(self.mu) - self.radius)

    def as_solver_config_kwargs(self) -> dict:
        return {"mu": self.mu, "radius": self.radius}


# ---- plane geometry on complex numbers -------------------------------------
_ROUND, _CROSS_TOL, _CONTAIN_TOL = 13, 1e-13, 1e-12


def _turn(o, a, b) -> float:
    return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])


def convex_hull(points) -> np.ndarray:
    """Convex hull (counter-clockwise vertices) of complex points by Andrew's
    monotone chain, after rounding to 13 decimals; collinear points are
    dropped (turn <= 1e-13)."""
    z = np.asarray(points, dtype=np.complex128).reshape(-1)
    xy = np.unique(np.round(np.column_stack((z.real, z.imag)), _ROUND), axis=0)
    if len(xy) < 2:
        raise ValueError("at least two distinct points are required")
    pts = [tuple(p) for p in xy[np.lexsort((xy[:, 1], xy[:, 0]))]]
    if len(pts) == 2:
        return np.array([complex(*p) for p in pts])

    def chain(seq):
        out = []
        for p in seq:
            while len(out) >= 2 and _turn(out[-2], out[-1], p) <= _CROSS_TOL:
                out.pop()
            out.append(p)
        return out

    lo, hi = chain(pts), chain(pts[::-1])
    verts = lo[:-1] + hi[:-1]
    if len(verts) < 2:
        verts = [pts[0], pts[-1]]
    return np.array([complex(*p) for p in verts])


def _origin_in_hull(hull: np.ndarray) -> bool:
    if len(hull) == 1:
        return abs(hull[0]) <= _CONTAIN_TOL
    if len(hull) == 2:
        a, b = hull
        d = b - a
        t = min(1.0, max(0.0, float(np.real(np.conj(d) * (-a)) / abs(d) ** 2))) if abs(d) > 0 else 0.0
        return abs(a + t * d) <= _CONTAIN_TOL
    area = 0.5 * float(np.sum(hull.real * np.roll(hull.imag, -1) - hull.imag * np.roll(hull.real, -1)))
    s = 1.0 if area >= 0 else -1.0
    for a, b in zip(hull, np.roll(hull, -1)):
        e, r = b - a, -a
        if s * (e.real * r.imag - e.imag * r.real) < -_CONTAIN_TOL:
            return False
    return True


def circle_two_points(z1: complex, z2: complex) -> tuple[complex, float]:
    """Circle through z1, z2 with the smallest radius / |centre|.

    Centres lie on the bisector mu(t) = (z1 + z2)/2 + i t (z2 - z1); with
    p = z1 conj(z2), setting d/dt (R^2 / |mu|^2) = 0 gives
    t = Im p / (2 (|p| + Re p)) and R^2 = |z2 - z1|^2 |p| / (2 (|p| + Re p))."""
    z1, z2 = complex(z1), complex(z2)
    p = z1 * np.conj(z2)
    den = 2.0 * (abs(p) + p.real)
    if abs(den) < 1e-30:
        raise ValueError("two-point circle undefined (points on opposite rays from the origin)")
    mu = 0.5 * (z1 + z2) + 1j * (p.imag / den) * (z2 - z1)
    r2 = abs(z2 - z1) ** 2 * abs(p) / den
    return complex(mu), float(np.sqrt(max(r2, 0.0)))


def circle_three_points(z1: complex, z2: complex, z3: complex) -> tuple[complex, float]:
    """Circumscribed circle of three points."""
    a, b, c = complex(z1), complex(z2), complex(z3)
    d = 2.0 * (a.real * (b.imag - c.imag) + b.real * (c.imag - a.imag) + c.real * (a.imag - b.imag))
    if abs(d) < 1e-30:
        raise ValueError("three points are collinear")
    na, nb, nc = abs(a) ** 2, abs(b) ** 2, abs(c) ** 2
    ux = (na * (b.imag - c.imag) + nb * (c.imag - a.imag) + nc * (a.imag - b.imag)) / d
    uy = (na * (c.real - b.real) + nb * (a.real - c.real) + nc * (b.real - a.real)) / d
    mu = complex(ux, uy)
    return mu, float(abs(mu - a))


def _best_circle(hull: np.ndarray):
    best = None
    pts = hull.reshape(-1)
    for group, make in ((2, circle_two_points), (3, circle_three_points)):
        for zs in combinations(pts, group):
            try:
                mu, r = make(*zs)
            except ValueError:
                continue
            if np.any(np.abs(pts - mu) > r + _CONTAIN_TOL) or abs(mu) <= r + _CONTAIN_TOL or abs(mu) <= _CONTAIN_TOL:
                continue
            q = r / abs(mu)
            if best is None or q < best[2]:
                best = (mu, r, q)
    return best


def analyze_spectrum(spectrum, *, coarse_N=None, matrix_shape=None) -> Gamma0Analysis:
    z = np.asarray(spectrum, dtype=np.complex128).reshape(-1)
    try:
        hull = convex_hull(z)
    except ValueError:
        raise ValueError("spectrum samples are degenerate") from None
    if _origin_in_hull(hull):
        raise ValueError("could not find a gamma0 circle: the origin belongs to the convex spectral hull")
    best = _best_circle(hull)
    if best is None:
        raise ValueError("could not find a gamma0 circle that excludes the origin")
    mu, r, q = best
    return Gamma0Analysis(mu=complex(mu), radius=float(r), rho=float(q), spectrum=z, hull=hull, coarse_N=coarse_N,
                          matrix_shape=matrix_shape)


def find_params(spectrum) -> dict:
    return analyze_spectrum(spectrum).as_solver_config_kwargs()


# ---- coarse operator ---------------------------------------------------------
def _shape(coarse_N) -> tuple[int, int, int]:
    shape = (coarse_N,) * 3 if isinstance(coarse_N, int) else tuple(int(n) for n in coarse_N)
    if len(shape) != 3 or any(n <= 0 for n in shape):
        raise ValueError(f"coarse_N must be a positive int or 3-tuple, got {coarse_N!r}")
    return shape


def _nearest(source, target) -> np.ndarray:
    """Index of the nearest source point for every target point (ties: lower index)."""
    s, t = np.asarray(source, np.float64), np.asarray(target, np.float64)
    hi = np.clip(np.searchsorted(s, t), 0, len(s) - 1)
    lo = np.clip(hi - 1, 0, len(s) - 1)
    return np.where(np.abs(t - s[lo]) <= np.abs(t - s[hi]), lo, hi)


def coarse_operator_matrix(problem, coarse_N=(4, 4, 4)) -> np.ndarray:
    """Dense A = I - B chi (complex128, 3n x 3n) on a coarse grid of the same box,
    contrast taken from the nearest fine cell."""
    import torch

    from .backend import Backend
    from .dtypes import Precision
    from .grid import Grid
    from .operator import Operator
    from .problem import Problem
    shape = _shape(coarse_N)
    fine = problem.grid
    be = Backend.hip(Precision.DOUBLE, fine.backend.device_index)
    grid = Grid(N=shape, L=fine.L, center=fine.center, backend=be)
    idx = [_nearest(np.asarray(fine._host[a], np.float64), np.asarray(grid._host[a], np.float64)) for a in range(3)]
    eps = problem.eps_tensor.to(torch.complex128)
    dev = eps.device
    eps = eps[:, :, torch.as_tensor(idx[0], device=dev)][:, :, :, torch.as_tensor(idx[1], device=dev)]
    eps = eps[:, :, :, :, torch.as_tensor(idx[2], device=dev)].contiguous()
    n = 3 * int(np.prod(shape))
    op = Operator(Problem(grid=grid, eps_tensor=eps, wave=torch.zeros((3,) + shape, dtype=torch.complex128,
                                                                       device=dev), k0=problem.k0, volume=1.0))
    basis = torch.eye(n, dtype=torch.complex128, device=dev).reshape((n, 3) + shape)
    cols = op.matvec(basis)                     # batch of n fields: column j = A e_j
    return cols.reshape(n, n).T.cpu().numpy()


def estimate_from_problem(problem, coarse_N=(4, 4, 4)) -> Gamma0Analysis:
    shape = _shape(coarse_N)
    A = coarse_operator_matrix(problem, coarse_N=shape)
    return analyze_spectrum(np.linalg.eigvals(A), coarse_N=shape, matrix_shape=A.shape)


def find_params_from_problem(problem, coarse_N=(4, 4, 4)) -> dict:
    return estimate_from_problem(problem, coarse_N=coarse_N).as_solver_config_kwargs()
