"""Stationary iteration on the HIP backend (src/em3d/solvers/sim.py:8-91 of the reference)."""
from __future__ import annotations

import math

import torch

from ._vec import lincomb, norm
from .base import SolverConfig, SolverResult


class SIM:
    def __init__(self, config: SolverConfig):
        config.require_gamma()
        self.cfg = config

    def solve(self, operator, rhs) -> SolverResult:
        cfg = self.cfg
        rhs = rhs.contiguous()
        gamma = 1.0 / complex(cfg.mu)
        u = torch.zeros_like(rhs)
        res: list[float] = []
        acts: list[int] = []
        rhs_norm = norm(rhs)
        if rhs_norm == 0.0:
            return SolverResult(u=u, iterations=0, residual_history=[0.0], converged=True, matvec_count=0,
                                residual_action_counts=[0], status="converged", true_final_residual=0.0)
        tol = max(float(cfg.rtol), float(cfg.atol) / rhs_norm)
        updates = mv = 0
        status, conv = "max_iter", False
        while True:
            r = lincomb(1.0, operator.matvec(u), -1.0, rhs)
            mv += 1
            rel = norm(r) / rhs_norm
            res.append(rel)
            acts.append(mv)
            if cfg.log:
                print(f"[SIM] updates={updates}, actions={mv}, rel_res={rel:.3e}")
            if not math.isfinite(rel):
                status = "nonfinite"
                break
            if rel < tol:
                status, conv = "converged", True
                break
            if cfg.divergence_guard is not None and rel > float(cfg.divergence_guard):
                status = "divergence_guard"
                break
            if cfg.max_operator_actions is not None and mv >= int(cfg.max_operator_actions):
                status = "max_operator_actions"
                break
            if updates >= int(cfg.max_iter):
                status = "max_iter"
                break
            u = lincomb(1.0, u, -gamma, r)
            updates += 1
        return SolverResult(u=u, iterations=updates, residual_history=res, converged=conv, matvec_count=mv,
                            rmatvec_count=0, residual_action_counts=acts, status=status,
                            true_final_residual=float(res[-1]))
