"""Pinning of the far-field oracle and of the host Mie series (CPU)."""
from __future__ import annotations

import json
import os

import numpy as np

from oracle import farfield as ff

HERE = os.path.dirname(os.path.abspath(__file__))


def test_single_cell_closed_form():
    """One cell with chi = I at r0: F(d) = dv u(r0) exp(-i k0 d . r0)."""
    N, L, c, k0 = (3, 4, 5), (0.9, 1.2, 1.0), (0.1, -0.2, 0.3), 4.0
    chi = np.zeros((3, 3) + N, np.complex128)
    idx = (1, 2, 3)
    for a in range(3):
        chi[(a, a) + idx] = 1.0
    rng = np.random.default_rng(0)
    u = rng.standard_normal((3,) + N) + 1j * rng.standard_normal((3,) + N)
    dirs = rng.standard_normal((7, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    h = np.asarray(L) / np.asarray(N)
    r0 = np.asarray(c) - np.asarray(L) / 2 + h / 2 + np.asarray(idx) * h
    dv = float(np.prod(h))
    want = dv * u[(slice(None),) + idx][None, :] * np.exp(-1j * k0 * dirs @ r0)[:, None]
    np.testing.assert_allclose(ff.scatter_integral(u, chi, N, L, c, k0, dirs), want, rtol=1e-12, atol=0)


def test_zero_contrast_and_linearity():
    N, L, c, k0 = (4, 3, 2), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), 3.0
    rng = np.random.default_rng(1)
    chi = rng.standard_normal((3, 3) + N) + 1j * rng.standard_normal((3, 3) + N)
    u1, u2 = (rng.standard_normal((3,) + N) + 1j * rng.standard_normal((3,) + N) for _ in range(2))
    dirs = np.eye(3)
    assert np.all(ff.scatter_integral(u1, np.zeros_like(chi), N, L, c, k0, dirs) == 0)
    lhs = ff.scatter_integral(2 * u1 - 1j * u2, chi, N, L, c, k0, dirs)
    rhs = 2 * ff.scatter_integral(u1, chi, N, L, c, k0, dirs) - 1j * ff.scatter_integral(u2, chi, N, L, c, k0, dirs)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12, atol=1e-15)


def test_mie_matches_reference_output_and_limits():
    """em3d_amd.mie: the reference's own Mie curve (tests/golden), the
    Rayleigh limit of the scattering cross section, no absorption for a
    real eps, nothing scattered for eps = 1."""
    from em3d_amd import mie
    g = json.load(open(os.path.join(HERE, "golden", "rcs_sphere_eps2_k0a1_N8.json")))
    phi, s = mie.mie_rcs_plane(0.3, 2.0, 1 / 0.3, n_phi=90, plane="xy")
    np.testing.assert_allclose(phi, g["phi"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(s, g["sigma_mie"], rtol=1e-12, atol=0)
    a, k0, eps = 0.01, 1.0, 2.5
    cs = mie.mie_cross_sections(a, eps, k0)
    rayleigh = 8 * np.pi / 3 * k0 ** 4 * a ** 6 * abs((eps - 1) / (eps + 2)) ** 2
    assert abs(cs["scat"] / rayleigh - 1) < 1e-3
    assert abs(cs["abs"]) < 1e-12 * cs["ext"] + 1e-30
    c1 = mie.mie_cross_sections(0.3, 1.0, 3.0)
    assert c1["scat"] == 0 and c1["ext"] == 0
    # the xz and yz planes meet the xy plane on its axes: xz at phi = 0 is (1, 0, 0) = xy at phi = 0
    _, sxz = mie.mie_rcs_plane(0.3, 2.0, 3.0, n_phi=4, plane="xz")
    _, sxy = mie.mie_rcs_plane(0.3, 2.0, 3.0, n_phi=4, plane="xy")
    assert abs(sxz[0] - sxy[0]) <= 1e-12 * sxy[0]
