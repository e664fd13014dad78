"""The MI355X path against the reference's own outputs (tests/golden, written
by tests/golden/make_reference_fixtures.py from the reference package):

* G1 -- the README's config-1 solve (16^3 sphere, eps_r 2, k0 4, plane wave
  along z polarised along x, BiCGStab rtol 1e-8, max_iter 200;
  solvers/bicgstab.py:12-211): same iterations, status, matvec and action
  counts, residual history to 1e-10, solution to 1e-9, on the device solver
  and on the host-driven loop;
* G2 -- compare_rcs_plane of the solution against Mie in the xz plane over
  120 angles (mie.py:179-240): both curves and shape_err / scale_ratio /
  abs_rel_err to 1e-6.
"""
from __future__ import annotations

import json
import os
import sys
from math import prod

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _config1(em):
    from em3d_amd.geometry import Ellipsoid, sample_contrast_tensor
    be = em.Backend.hip(em.Precision.DOUBLE)
    center = (0.0, 0.0, 0.0)
    grid = em.Grid(N=(16, 16, 16), L=(1.0, 1.0, 1.0), center=center, backend=be)
    contrast = sample_contrast_tensor(grid, geometry=Ellipsoid(center=center, radii=(0.25,) * 3),
                                      background_eps_r=1.0, feature_eps_r=2.0)
    wave = em.flat_wave_vec(grid, k=4.0, orient=(0, 0, 1), amplitude=(1, 0, 0))
    return em.Problem(grid=grid, eps_tensor=contrast, wave=wave, k0=4.0, volume=prod(grid.L))


class _Foreign:
    def __init__(self, op):
        self._op, self.backend = op, op.backend

    def matvec(self, x):
        return self._op.matvec(x)


@pytest.mark.parametrize("path", ["device", "host"])
def test_config1_bicgstab_matches_reference(path):
    import em3d_amd as em
    g = np.load(os.path.join(HERE, "golden", "bicgstab_sphere16_trajectory.npz"), allow_pickle=False)
    _say(f"[G1] config-1 BiCGStab, {path} loop")
    prob = _config1(em)
    op = em.Operator(prob)
    res = em.BiCGStab(em.SolverConfig(max_iter=200, rtol=1e-8)).solve(op if path == "device" else _Foreign(op),
                                                                     prob.wave)
    assert res.status == str(g["status"]) == "converged"
    assert res.iterations == int(g["iterations"]) and res.matvec_count == int(g["matvec_count"])
    assert res.residual_action_counts == [int(x) for x in g["residual_action_counts"]]
    np.testing.assert_allclose(res.residual_history, g["residual_history"], rtol=1e-10, atol=0)
    u = res.u.cpu().numpy()
    assert np.linalg.norm(u - g["u"]) / np.linalg.norm(g["u"]) < 1e-9
    true = em.true_relative_residual(op, res.u, prob.wave)
    assert abs(true - float(g["true_final_residual"])) <= 1e-6 * float(g["true_final_residual"])


def test_config1_rcs_matches_reference():
    import em3d_amd as em
    g = json.load(open(os.path.join(HERE, "golden", "rcs_xz_sphere16.json")))
    _say("[G2] config-1 compare_rcs_plane xz 120")
    prob = _config1(em)
    res = em.BiCGStab(em.SolverConfig(max_iter=200, rtol=1e-8)).solve(em.Operator(prob), prob.wave)
    out = em.mie.compare_rcs_plane(res.u, prob, a=0.25, eps_r=2.0, n_phi=120, plane="xz")
    np.testing.assert_allclose(out["phi"], g["phi"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(out["sigma_mie"], g["sigma_mie"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(out["sigma_num"], g["sigma_num"], rtol=1e-8, atol=1e-12 * max(g["sigma_num"]))
    for key in ("shape_err", "scale_ratio", "abs_rel_err"):
        assert abs(out[key] - g[key]) < 1e-6, (key, out[key], g[key])
