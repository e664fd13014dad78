"""The device BiCGStab with its vector updates fused into the matvec passes
(p and s updates in the first z pass, the dot products in the last, one
u/r kernel) against the stand-alone vector kernels and the oracle's
trajectory (src/em3d/solvers/bicgstab.py:12-211)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _problem(em, prec, N, k, kind):
    be = em.Backend.hip(em.Precision.DOUBLE if prec == "double" else em.Precision.SINGLE)
    grid = em.Grid(N=N, L=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), backend=be)
    eps = 2.5 + 0.1j if kind == "iso" else np.diag([2.5 + 0.1j, 3.0 + 0.2j, 2.0])
    chi = em.geometry.sample_contrast_tensor(grid, background_eps_r=1.0, feature_eps_r=eps,
                                             geometry=em.geometry.Ellipsoid(center=(0, 0, 0), radii=(0.3, 0.2, 0.15)))
    wave = em.flat_wave_vec(grid, k=k, orient=(0, 0, 1), amplitude=(1, 0, 0))
    return em.Problem(grid=grid, eps_tensor=chi, wave=wave, k0=k, volume=1.0)


# (N, k, contrast): fast z kernels (16^3, 32^3, 48^3) and generic ones (odd sizes)
CASES = [((16, 16, 16), 8.0, "diag"), ((32, 32, 32), 12.0, "iso"), ((48, 48, 48), 8.0, "diag"),
         ((9, 10, 7), 6.0, "diag")]


@pytest.mark.parametrize("N,k,kind", CASES)
def test_fused_fp64_matches_unfused_and_oracle(N, k, kind):
    import em3d_amd as em
    from oracle import solvers
    _say(f"[fused fp64] N={N} k={k} chi={kind}")
    prob = _problem(em, "double", N, k, kind)
    op = em.Operator(prob)
    assert op.contrast_kind == {"iso": "isotropic", "diag": "diagonal"}[kind]
    fused = em.BiCGStab(em.SolverConfig(max_iter=300, rtol=1e-10)).solve(op, prob.wave)
    plain = em.BiCGStab(em.SolverConfig(max_iter=300, rtol=1e-10, fused=False)).solve(op, prob.wave)
    assert fused.status == plain.status == "converged"
    assert (fused.iterations, fused.matvec_count) == (plain.iterations, plain.matvec_count)
    assert fused.residual_action_counts == plain.residual_action_counts
    # entries are |r| / |f|: rounding of the residual vectors is ~1e-16 absolute
    np.testing.assert_allclose(fused.residual_history, plain.residual_history, rtol=1e-9, atol=1e-14)
    assert _rel(fused.u.cpu().numpy(), plain.u.cpu().numpy()) < 1e-11
    if max(N) <= 16:
        chi = prob.eps_tensor.cpu().numpy()
        ref = solvers.bicgstab(solvers.EMOperator(N, (1, 1, 1), k, chi), prob.wave.cpu().numpy(),
                               solvers.Config(max_iter=300, rtol=1e-10))
        assert ref.iterations == fused.iterations and ref.status == fused.status
        np.testing.assert_allclose(fused.residual_history, ref.residual_history, rtol=1e-8, atol=1e-14)


@pytest.mark.parametrize("N,k,kind", CASES[:3])
def test_fused_fp32_reaches_the_same_tolerance(N, k, kind):
    import em3d_amd as em
    _say(f"[fused fp32] N={N} k={k} chi={kind}")
    prob = _problem(em, "single", N, k, kind)
    op = em.Operator(prob)
    cfg = dict(max_iter=300, rtol=1e-5)
    fused = em.BiCGStab(em.SolverConfig(**cfg)).solve(op, prob.wave)
    plain = em.BiCGStab(em.SolverConfig(**cfg, fused=False)).solve(op, prob.wave)
    assert fused.status == plain.status == "converged"
    assert abs(fused.iterations - plain.iterations) <= 1
    np.testing.assert_allclose(fused.residual_history[:4], plain.residual_history[:4], rtol=1e-4)
    true = em.true_relative_residual(op, fused.u, prob.wave)
    assert true < 2e-5


def test_fused_statuses_and_batch():
    """Zero and NaN right-hand sides, an action budget and max_iter on the
    fused path (diagonal contrast), and a batch whose members equal their
    single fused solves bit for bit."""
    import torch

    import em3d_amd as em
    _say("[fused statuses]")
    N = (16, 16, 16)
    prob = _problem(em, "double", N, 6.0, "diag")
    op = em.Operator(prob)
    be = prob.grid.backend
    r = em.BiCGStab(em.SolverConfig()).solve(op, be.zeros((3,) + N, "complex"))
    assert r.status == "converged" and r.matvec_count == 0
    bad = prob.wave.clone()
    bad[0, 1, 2, 3] = float("nan")
    assert em.BiCGStab(em.SolverConfig()).solve(op, bad).status == "nonfinite"
    r = em.BiCGStab(em.SolverConfig(max_iter=100, max_operator_actions=4)).solve(op, prob.wave)
    assert r.status == "max_operator_actions" and r.operator_action_count <= 4
    r = em.BiCGStab(em.SolverConfig(max_iter=2, rtol=1e-12)).solve(op, prob.wave)
    assert r.status == "max_iter" and r.iterations == 2
    rhs = torch.stack([prob.wave, torch.zeros_like(prob.wave), 0.5j * prob.wave, bad])
    cfg = em.SolverConfig(max_iter=200, rtol=1e-9)
    batch = em.BiCGStab(cfg).solve_batch(op, rhs)
    for b in range(4):
        single = em.BiCGStab(cfg).solve(op, rhs[b].contiguous())
        assert (batch[b].status, batch[b].iterations, batch[b].matvec_count) == (single.status, single.iterations,
                                                                                 single.matvec_count), b
        np.testing.assert_allclose(batch[b].residual_history, single.residual_history, rtol=1e-12, atol=0)
        if single.status != "nonfinite":
            assert torch.equal(batch[b].u, single.u), b
