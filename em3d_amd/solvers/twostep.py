"""Two-step gradient method on the HIP backend (src/em3d/solvers/twostep.py:12-139 of the reference)."""
from __future__ import annotations

import math

import torch

from ._vec import dots, lincomb, norm
from .base import SolverConfig, SolverResult


class TwoStep:
    def __init__(self, config: SolverConfig):
        self.cfg = config

    def solve(self, operator, rhs) -> SolverResult:
        cfg = self.cfg
        rhs = rhs.contiguous()
        rhs_norm = norm(rhs)
        res: list[float] = []
        acts: list[int] = []
        u = torch.zeros_like(rhs)
        if rhs_norm == 0.0:
            return SolverResult(u=u, iterations=0, residual_history=[0.0], converged=True, matvec_count=0,
                                rmatvec_count=0, residual_action_counts=[0], status="converged",
                                true_final_residual=0.0)
        tol = max(float(cfg.rtol), float(cfg.atol) / rhs_norm)
        pu = pr = None
        updates = mv = rmv = 0
        status = "max_iter"
        while True:
            r = lincomb(1.0, operator.matvec(u), -1.0, rhs)
            mv += 1
            rel = norm(r) / rhs_norm
            res.append(rel)
            acts.append(mv + rmv)
            if cfg.log:
                print(f"[TwoStep] updates={updates}, actions={mv + rmv}, rel_res={rel:.3e}")
            if not math.isfinite(rel):
                status = "nonfinite"
                break
            if rel < tol:
                status = "converged"
                break
            if cfg.divergence_guard is not None and rel > float(cfg.divergence_guard):
                status = "divergence_guard"
                break
            if cfg.max_operator_actions is not None and mv + rmv >= int(cfg.max_operator_actions):
                status = "max_operator_actions"
                break
            if updates >= int(cfg.max_iter):
                status = "max_iter"
                break
            if cfg.max_operator_actions is not None and mv + rmv + 2 > int(cfg.max_operator_actions):
                status = "max_operator_actions"
                break
            g = operator.rmatvec(r)
            rmv += 1
            Hg = operator.matvec(g)
            mv += 1
            if pu is None:
                a11c, ggc = dots([(Hg, Hg), (g, g)])
                a11 = a11c.real
                if not math.isfinite(a11) or a11 <= 0.0:
                    status = "breakdown_gradient"
                    break
                h = ggc.real / a11
                nu = lincomb(1.0, u, -h, g)
            else:
                dr = lincomb(1.0, r, -1.0, pr)
                a11c, a00c, a01c, b0c, b1c = dots([(Hg, Hg), (dr, dr), (dr, Hg), (r, dr), (r, Hg)])
                a11 = a11c.real
                if not math.isfinite(a11) or a11 <= 0.0:
                    status = "breakdown_gradient"
                    break
                a00, a01, b0, b1 = a00c.real, a01c.real, b0c.real, b1c.real
                det = a00 * a11 - a01 * a01
                scale = max(abs(a00 * a11), abs(a01 * a01), 1.0)
                if abs(det) <= 1e-14 * scale:
                    t, h = 0.0, b1 / a11
                else:
                    t = (b0 * a11 - b1 * a01) / det
                    h = (a00 * b1 - a01 * b0) / det
                # u - t (u - pu) - h g = (1 - t) u + t pu - h g
                nu = lincomb(1.0 - t, u, t, pu, -h, g)
            pu, pr, u = u, r, nu
            updates += 1
        return SolverResult(u=u, iterations=updates, residual_history=res, converged=status == "converged",
                            matvec_count=mv, rmatvec_count=rmv, residual_action_counts=acts, status=status,
                            true_final_residual=float(res[-1]))
