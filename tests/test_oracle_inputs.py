"""Pinning of oracle/inputs.py (grid axes, plane waves, masks, contrasts) by
what the mathematics fixes, independently of the restated formulas:

* axes: exact rational cell centres c - L/2 + (i + 1/2) L/N (fractions) to a
  few ulps, the span (n - 1) L/n, antisymmetry of a centred axis;
* plane waves: the value at a cell chosen so that the phase is a known
  multiple of pi/2, unit modulus, the phase step between neighbouring cells,
  and the unit-vector check;
* masks: counts of cells inside a sphere or box obtained by integer
  arithmetic on doubled cell indices (no floating point), including cells
  exactly on the surface, which the centre-test tolerance keeps inside
  (src/em3d/geometry.py:62-93) and the refraction builders, without
  tolerance, decide by rounding;
* contrast: eps_f - I inside and eps_bg - I outside at chosen cells.

A planted mistake in the oracle (a dropped half-cell offset, a sign flip in
the phase, a transposed material) fails one of these checks.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

from oracle import inputs


@pytest.mark.parametrize("n,length,centre", [(4, 1.0, 0.0), (7, 0.9, 0.25), (16, 1.0, 0.0), (5, 2.0, -0.5)])
def test_axis_exact_centres(n, length, centre):
    a = inputs.axis(n, length, centre)
    L, c = Fraction(length), Fraction(centre)
    exact = [float(c - L / 2 + (2 * i + 1) * L / (2 * n)) for i in range(n)]
    # start = c - L/2 + h/2 and the step are rounded once each before linspace
    ulp = np.spacing(max(abs(centre) + length, 1.0))
    assert np.all(np.abs(a - np.array(exact)) <= 4 * ulp)
    assert np.all(np.diff(a) > 0) and abs((a[-1] - a[0]) - (n - 1) * length / n) <= 8 * ulp
    if centre == 0.0:
        assert np.allclose(a, -a[::-1], rtol=0, atol=4 * np.spacing(1.0))
    a32 = inputs.axis(n, length, centre, "single")
    assert a32.dtype == np.float32 and np.allclose(a32, exact, rtol=0, atol=4e-7)


def test_cell_volume_and_coords():
    N, L = (4, 3, 2), (1.0, 0.9, 0.6)
    assert inputs.cell_volume(N, L) == pytest.approx(0.25 * 0.3 * 0.3, rel=1e-15)
    X, Y, Z = inputs.coords(N, L, (0.1, 0.0, -0.2))
    assert X.shape == N and X[2, 0, 0] == inputs.axis(4, 1.0, 0.1)[2] and abs(Y[0, 1, 1]) < 1e-15
    assert abs(X[2, 0, 0] - (0.1 - 0.5 + 2.5 * 0.25)) < 1e-15
    assert Z[3, 2, 1] == pytest.approx(-0.2 + 0.15, abs=1e-15)


def test_plane_wave_known_values():
    # N = 4, L = 1, centred: z centres are -3/8, -1/8, 1/8, 3/8.  k = 4 pi / 3
    # makes k z0 = -pi/2 at the first z cell, so exp(i k z0) = -i.
    N, L = (4, 4, 4), (1.0, 1.0, 1.0)
    k = 4.0 * np.pi / 3.0
    f = inputs.flat_wave_vec(N, L, (0, 0, 0), k=k, orient=(0, 0, 1), amplitude=(2.0, 0, 0.5j))
    assert abs(f[0, 1, 2, 0] - 2.0 * (-1j)) < 1e-15
    assert abs(f[2, 3, 0, 0] - 0.5j * (-1j)) < 1e-15
    assert np.all(f[1] == 0)
    assert np.allclose(np.abs(f[0]), 2.0, rtol=1e-15)
    # phase step along z is k h = pi / 3, and sign = -1 conjugates
    assert np.allclose(f[0, :, :, 1] / f[0, :, :, 0], np.exp(1j * np.pi / 3), atol=1e-14)
    g = inputs.flat_wave_vec(N, L, (0, 0, 0), k=k, orient=(0, 0, 1), amplitude=(2.0, 0, 0.5j), sign=-1)
    assert np.allclose(g[0], np.conj(f[0]), atol=1e-15)
    # phi0 rotates every value by exp(i phi0); an oblique orient adds phases per axis
    h = inputs.flat_wave_vec(N, L, (0, 0, 0), k=k, orient=(0, 0, 1), amplitude=(2.0, 0, 0), phi0=np.pi / 2)
    assert np.allclose(h[0], 1j * f[0], atol=1e-14)
    s = 1.0 / np.sqrt(2.0)
    o = inputs.flat_wave_vec(N, L, (0, 0, 0), k=k, orient=(s, 0, s), amplitude=(1, 0, 0))
    assert abs(o[0, 0, 0, 0] - np.exp(1j * k * s * (-0.375 - 0.375))) < 1e-15
    with pytest.raises(ValueError, match="unit vector"):
        inputs.flat_wave_vec(N, L, (0, 0, 0), k=k, orient=(0, 0, 1.1), amplitude=(1, 0, 0))
    with pytest.raises(ValueError):
        inputs.flat_wave_vec(N, L, (0, 0, 0), k=k, orient=(0, 0, 1), amplitude=(1, 0, 0), sign=2)


def test_scalar_plane_wave_known_value():
    N, L = (4, 4, 4), (1.0, 1.0, 1.0)
    k = 4.0 * np.pi / 3.0
    w = inputs.plane_wave_scalar(N, L, (0, 0, 0), k=k, direction=(0, 0, 3.0), amplitude=1.5)
    assert abs(w[1, 1, 0] - 1.5 * (-1j)) < 1e-15 and abs(w[0, 3, 3] - 1.5 * 1j) < 1e-15


def _odd_offsets(n):
    """Doubled cell-centre offsets 2 i + 1 - n of a centred axis (units L / 2n)."""
    return np.arange(n) * 2 + 1 - n


def _count_inside(n, r2_units):
    o = _odd_offsets(n)
    s = o[:, None, None] ** 2 + o[None, :, None] ** 2 + o[None, None, :] ** 2
    return int(np.count_nonzero(s <= r2_units)), s


@pytest.mark.parametrize("n,radius", [(16, 0.25), (16, 0.3), (12, 0.41), (9, 0.35)])
def test_sphere_mask_counts(n, radius):
    # centred sphere: cell inside iff sum (2 i + 1 - n)^2 <= (2 n r)^2, integers
    N, L = (n, n, n), (1.0, 1.0, 1.0)
    want, s = _count_inside(n, (2 * n * radius) ** 2)
    m = inputs.ellipsoid_center_mask(N, L, (0, 0, 0), ell_center=(0, 0, 0), radii=(radius,) * 3)
    if not np.any(s == (2 * n * radius) ** 2):
        assert int(m.sum()) == want
    assert np.array_equal(m, m[::-1]) and np.array_equal(m, m.transpose(1, 0, 2))
    chi = inputs.ellipsis_refraction(N, L, (0, 0, 0), eps_real=2.0, eps_imag=0.5, ell_center=(0, 0, 0),
                                     radius=(radius,) * 3)
    assert np.count_nonzero(chi[0, 0]) == int(m.sum()) or np.any(s == (2 * n * radius) ** 2)


def test_surface_ties_are_inside_with_tolerance():
    # sphere centred on a cell centre with radius exactly 2 cells: the cells at
    # integer distance 2 lie on the surface; with the tolerance 33 cells are
    # inside (1 + 6 + 12 + 8 + 6 lattice points with |v|^2 <= 4)
    n, h = 16, 1.0 / 16
    N, L = (n, n, n), (1.0, 1.0, 1.0)
    c0 = inputs.axis(n, 1.0, 0.0)[7]
    m = inputs.ellipsoid_center_mask(N, L, (0, 0, 0), ell_center=(c0, c0, c0), radii=(2 * h,) * 3)
    assert int(m.sum()) == 33
    assert m[9, 7, 7] and m[5, 7, 7] and m[7, 7, 9] and not m[9, 8, 7]


def test_box_mask_counts():
    n = 10
    N, L = (n, n, n), (1.0, 1.0, 1.0)
    # half sizes 0.3 / 0.2 / 0.1 hold |2 i + 1 - n| <= 2 n s: 6, 4, 2 cells per axis
    m = inputs.box_center_mask(N, L, (0, 0, 0), box_center=(0, 0, 0), size=(0.6, 0.4, 0.2))
    assert int(m.sum()) == 6 * 4 * 2
    assert m[2, 3, 4] and not m[1, 3, 4] and not m[2, 2, 4] and not m[2, 3, 3]


def test_contrast_values_and_layout():
    N = (4, 4, 4)
    w = np.zeros(N)
    w[1, 2, 3] = 1.0
    eps_f = np.array([[2.0, 0.1j, 0], [0.2, 3.0 + 0.5j, 0], [0, 0, 4.0]])
    chi = inputs.contrast_from_weights(w, background_eps_r=1.5, feature_eps_r=eps_f)
    assert chi.shape == (3, 3) + N
    assert np.allclose(chi[:, :, 1, 2, 3], eps_f - np.eye(3), atol=0)
    assert chi[0, 1, 1, 2, 3] == 0.1j and chi[1, 0, 1, 2, 3] == 0.2       # not transposed
    assert np.allclose(chi[:, :, 0, 0, 0], 0.5 * np.eye(3), atol=0)
    half = inputs.contrast_from_weights(0.5 * np.ones(N), background_eps_r=1.0, feature_eps_r=3.0)
    assert np.allclose(half[0, 0], 1.0) and np.allclose(half[0, 1], 0.0)


def test_refraction_and_eta_builders():
    n = 8
    N, L = (n, n, n), (1.0, 1.0, 1.0)
    chi = inputs.ellipsis_refraction(N, L, (0, 0, 0), eps_real=[[2.0, 0, 0], [0, 3.0, 0], [0, 0, 4.0]],
                                     eps_imag=[[0.1, 0, 0], [0, 0, 0], [0, 0, 0.2]], ell_center=(0, 0, 0),
                                     radius=(0.3, 0.3, 0.3))
    want, _ = _count_inside(n, (2 * n * 0.3) ** 2)
    assert np.count_nonzero(chi[0, 0]) == want
    inside = np.argwhere(chi[0, 0] != 0)[0]
    assert chi[(0, 0) + tuple(inside)] == 1.0 + 0.1j and chi[(1, 1) + tuple(inside)] == 2.0
    assert chi[(2, 2) + tuple(inside)] == 3.0 + 0.2j and chi[(0, 1) + tuple(inside)] == 0
    cyl = inputs.cylinder_refraction(N, L, (0, 0, 0), eps_real=2.0, eps_imag=0.0, radius=0.3)
    o = _odd_offsets(n)
    disc = int(np.count_nonzero(o[:, None] ** 2 + o[None, :] ** 2 <= (2 * n * 0.3) ** 2))
    assert np.count_nonzero(cyl[0, 0]) == disc * n and np.all(cyl[0, 1] == 0)
    eta = inputs.eta_ellipsoid(N, L, (0, 0, 0), ell_center=(0, 0, 0), radii=(0.3,) * 3, eta_inside=2.0 + 0.1j)
    assert np.count_nonzero(eta == 2.0 + 0.1j) == want and np.count_nonzero(eta == 1.0) == n ** 3 - want
