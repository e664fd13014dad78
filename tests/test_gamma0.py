"""SIM parameter pipeline (em3d_amd.gamma0): the circle constructions are
checked against brute force, the analysis against its defining properties,
and (GPU) the coarse operator against the dense oracle."""
from __future__ import annotations

import numpy as np
import pytest

from em3d_amd import gamma0


def test_two_point_circle_is_the_optimum_on_the_bisector():
    rng = np.random.default_rng(0)
    for _ in range(20):
        z1, z2 = (complex(*rng.uniform([0.5, -1], [2.0, 1])) for _ in range(2))
        mu, r = gamma0.circle_two_points(z1, z2)
        assert abs(abs(mu - z1) - r) < 1e-12 and abs(abs(mu - z2) - r) < 1e-12
        t = np.linspace(-50, 50, 400001)
        m = 0.5 * (z1 + z2) + 1j * t * (z2 - z1)
        q = np.abs(m - z1) / np.abs(m)
        assert r / abs(mu) <= q.min() + 1e-9


def test_three_point_circle():
    mu, r = gamma0.circle_three_points(1 + 0j, 1j, -1 + 0j)
    assert abs(mu) < 1e-15 and abs(r - 1) < 1e-15
    with pytest.raises(ValueError):
        gamma0.circle_three_points(0j, 1 + 1j, 2 + 2j)


def test_analysis_contains_spectrum_excludes_origin_and_is_minimal():
    rng = np.random.default_rng(1)
    spec = 1.0 + 0.4 * (rng.standard_normal(60) + 1j * rng.standard_normal(60)) * 0.5
    a = gamma0.analyze_spectrum(spec)
    assert np.all(np.abs(spec - a.mu) <= a.radius + 1e-9) and abs(a.mu) > a.radius
    # no circle centred on a fine grid of candidate centres does better
    cx, cy = np.meshgrid(np.linspace(0.2, 4, 200), np.linspace(-2, 2, 200))
    c = (cx + 1j * cy).ravel()
    R = np.max(np.abs(spec[None, :] - c[:, None]), axis=1)
    ok = np.abs(c) > R
    assert a.rho <= np.min(R[ok] / np.abs(c[ok])) + 1e-6
    assert set(a.as_solver_config_kwargs()) == {"mu", "radius"}


def test_analysis_rejects_origin_in_hull_and_degenerate_spectra():
    with pytest.raises(ValueError, match="origin"):
        gamma0.analyze_spectrum(np.array([1 + 1j, -1 + 1j, -1 - 1j, 1 - 1j]))
    with pytest.raises(ValueError, match="degenerate"):
        gamma0.analyze_spectrum(np.array([2 + 0j, 2 + 0j]))


def test_sim_with_the_circle_converges_on_its_matrix():
    """SIM with (mu, R) from the spectrum converges on a matrix with that spectrum (oracle SIM)."""
    from oracle import solvers
    rng = np.random.default_rng(2)
    lam = 1.0 + 0.3 * rng.standard_normal(30) + 0.3j * rng.standard_normal(30)
    Q, _ = np.linalg.qr(rng.standard_normal((30, 30)) + 1j * rng.standard_normal((30, 30)))
    A = Q @ np.diag(lam) @ Q.conj().T
    p = gamma0.find_params(np.linalg.eigvals(A))
    f = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    res = solvers.sim(solvers.MatrixOperator(A), f, solvers.Config(max_iter=2000, rtol=1e-10, mu=p["mu"],
                                                                   radius=p["radius"]))
    assert res.status == "converged"
