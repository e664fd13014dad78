"""Pins the oracle's solver restatement.

* BiCGStab: iterate-by-iterate agreement with scipy.sparse.linalg.bicgstab
  (the same van der Vorst recurrence, r_hat = r0, x0 = 0) and manufactured
  solutions; statuses forced from the host (zero rhs, NaN, A = 0, budgets).
* SIM: closed form on a diagonal matrix, e_n = (1 - lambda/mu)^n e_0.
* TwoStep: the reference's two-update closed form on a 2x2 matrix
  (tests/test_solvers.py:25-45 of the reference).
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import solvers, vie, inputs


def _system(n=40, seed=0):
    rng = np.random.default_rng(seed)
    A = np.eye(n) + 0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return A, b


def test_bicgstab_iterates_match_scipy():
    A, b = _system()
    ours = []
    solvers.bicgstab(solvers.MatrixOperator(A), b, solvers.Config(max_iter=12, rtol=1e-30),
                     callback=lambda u: ours.append(u.copy()))
    theirs = []
    spla.bicgstab(A, b, rtol=1e-30, atol=0.0, maxiter=12, callback=lambda x: theirs.append(x.copy()))
    assert len(ours) == len(theirs) == 12
    for a, c in zip(ours, theirs):
        assert np.linalg.norm(a - c) / np.linalg.norm(c) < 1e-11


def test_bicgstab_manufactured_em():
    N, L, k = (4, 4, 4), (1.0, 1.0, 1.0), 0.5
    chi = inputs.cylinder_refraction(N, L, (0, 0, 0), eps_real=1.1, eps_imag=0.0, radius=0.2)
    op = solvers.EMOperator(N, L, k, chi)
    rng = np.random.default_rng(1)
    ut = 0.1 * (rng.standard_normal((3,) + N) + 1j * rng.standard_normal((3,) + N))
    res = solvers.bicgstab(op, op.matvec(ut), solvers.Config(max_iter=200, rtol=1e-8))
    assert res.converged and res.status == "converged"
    assert np.linalg.norm(res.u - ut) / np.linalg.norm(ut) < 1e-6
    assert res.residual_action_counts[-1] == res.matvec_count
    assert len(res.residual_history) == len(res.residual_action_counts)


def test_bicgstab_forced_statuses():
    A, b = _system(8, 3)
    op = solvers.MatrixOperator(A)
    r = solvers.bicgstab(op, np.zeros(8, complex), solvers.Config())
    assert r.status == "converged" and r.iterations == 0 and r.matvec_count == 0
    bn = b.copy()
    bn[0] = np.nan
    assert solvers.bicgstab(op, bn, solvers.Config()).status == "nonfinite"
    assert solvers.bicgstab(solvers.MatrixOperator(np.zeros((8, 8))), b, solvers.Config()).status == "breakdown_alpha"
    r = solvers.bicgstab(op, b, solvers.Config(max_iter=100, max_operator_actions=2))
    assert r.status == "max_operator_actions" and r.matvec_count <= 2


def test_sim_diagonal_closed_form():
    lam = np.array([0.9, 1.1, 1.05, 0.95], complex)
    A = np.diag(lam)
    b = np.ones(4, complex)
    mu = 1.0
    res = solvers.sim(solvers.MatrixOperator(A), b, solvers.Config(max_iter=5, rtol=1e-30, mu=mu, radius=0.2))
    x = b / lam
    e = (1 - lam / mu) ** 5 * (0 - x)
    np.testing.assert_allclose(res.u - x, e, atol=1e-14)
    assert res.status == "max_iter" and res.matvec_count == 6


def _manual_twostep(A, rhs):
    u0 = np.zeros_like(rhs)
    r0 = A @ u0 - rhs
    g0 = A.conj().T @ r0
    Hg0 = A @ g0
    u1 = u0 - (np.vdot(g0, g0).real / np.vdot(Hg0, Hg0).real) * g0
    r1 = A @ u1 - rhs
    g1 = A.conj().T @ r1
    Hg1 = A @ g1
    d = r1 - r0
    a00, a01, a11 = np.vdot(d, d).real, np.vdot(d, Hg1).real, np.vdot(Hg1, Hg1).real
    b0, b1 = np.vdot(r1, d).real, np.vdot(r1, Hg1).real
    det = a00 * a11 - a01 * a01
    t = (b0 * a11 - b1 * a01) / det
    h = (a00 * b1 - a01 * b0) / det
    return u1 - t * (u1 - u0) - h * g1


def test_twostep_two_update_closed_form():
    A = np.array([[2.0, 0.5], [1.0, 1.5]], complex)
    b = np.array([1.0, -0.5], complex)
    res = solvers.twostep(solvers.MatrixOperator(A), b, solvers.Config(max_iter=2, rtol=1e-15))
    np.testing.assert_allclose(res.u, _manual_twostep(A, b), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", ["sim", "bicgstab", "twostep"])
def test_action_budgets(name):
    N, L, k = (4, 4, 4), (1.0, 1.0, 1.0), 0.5
    chi = inputs.cylinder_refraction(N, L, (0, 0, 0), eps_real=1.1, eps_imag=0.0, radius=0.2)
    op = solvers.EMOperator(N, L, k, chi)
    rhs = inputs.flat_wave_vec(N, L, (0, 0, 0), k=0.5, orient=(0, 0, 1), amplitude=(1, 0, 0))
    cfg = solvers.Config(max_iter=100, max_operator_actions=2, mu=2.0 + 0j, radius=1.0)
    r = getattr(solvers, name)(op, rhs, cfg)
    assert r.matvec_count + r.rmatvec_count <= 2
    assert r.status in {"converged", "max_operator_actions"}


def test_true_relative_residual_closed_form():
    """A = 2 I, u = f / 2 + d: |A u - f| / |f| = 2 |d| / |f|; f = 0 gives |A u|."""
    from oracle import solvers
    rng = np.random.default_rng(1)
    f = rng.standard_normal(10) + 1j * rng.standard_normal(10)
    d = 1e-3 * (rng.standard_normal(10) + 0j)
    op = solvers.MatrixOperator(2.0 * np.eye(10))
    r = solvers.true_relative_residual(op, f / 2 + d, f)
    assert abs(r - 2 * np.linalg.norm(d) / np.linalg.norm(f)) < 1e-15
    assert solvers.true_relative_residual(op, d, np.zeros(10)) == 2 * np.linalg.norm(d)
