"""The oracle against the reference's own outputs (tests/golden, written by
tests/golden/make_reference_fixtures.py from the reference package): the
README's config-1 solve (16^3 sphere, eps_r 2, k0 4, BiCGStab rtol 1e-8) and
its Mie comparison in the xz plane over 120 angles."""
from __future__ import annotations

import json
import os

import numpy as np

from oracle import farfield, inputs, solvers

HERE = os.path.dirname(os.path.abspath(__file__))


def _config1():
    N, L, c = (16, 16, 16), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    mask = inputs.ellipsoid_center_mask(N, L, c, ell_center=c, radii=(0.25,) * 3)
    chi = inputs.contrast_from_weights(mask.astype(float), background_eps_r=1.0, feature_eps_r=2.0)
    wave = inputs.flat_wave_vec(N, L, c, k=4.0, orient=(0, 0, 1), amplitude=(1, 0, 0))
    return N, L, c, chi, wave


def test_oracle_bicgstab_matches_reference_trajectory():
    g = np.load(os.path.join(HERE, "golden", "bicgstab_sphere16_trajectory.npz"), allow_pickle=False)
    N, L, c, chi, wave = _config1()
    res = solvers.bicgstab(solvers.EMOperator(N, L, 4.0, chi), wave, solvers.Config(max_iter=200, rtol=1e-8))
    assert res.status == str(g["status"]) == "converged"
    assert res.iterations == int(g["iterations"]) and res.matvec_count == int(g["matvec_count"])
    assert res.residual_action_counts == [int(x) for x in g["residual_action_counts"]]
    np.testing.assert_allclose(res.residual_history, g["residual_history"], rtol=1e-11, atol=0)
    assert np.linalg.norm(res.u - g["u"]) / np.linalg.norm(g["u"]) < 1e-12


def test_oracle_rcs_matches_reference_curve():
    g = json.load(open(os.path.join(HERE, "golden", "rcs_xz_sphere16.json")))
    u = np.load(os.path.join(HERE, "golden", "bicgstab_sphere16_trajectory.npz"), allow_pickle=False)["u"]
    N, L, c, chi, _ = _config1()
    phi = np.asarray(g["phi"])
    dirs = np.stack([np.cos(phi), np.zeros_like(phi), np.sin(phi)], axis=-1)
    sigma = farfield.rcs_plane(farfield.scatter_integral(u, chi, N, L, c, 4.0, dirs), dirs, 4.0)
    np.testing.assert_allclose(sigma, g["sigma_num"], rtol=1e-11, atol=1e-14 * max(g["sigma_num"]))
