"""BiCGStab on the HIP backend (src/em3d/solvers/bicgstab.py:12-211 of the reference).

Same recurrence, status strings, early exit on ||s|| and residual-history
semantics as the reference.  Vector updates and FP64-accumulated reductions
run as libem3d_hip kernels; scalars the control flow needs are read back in
one copy per decision point.
"""
from __future__ import annotations

import math

import torch

from ._vec import dots, lincomb, norm
from .base import SolverConfig, SolverResult


def _finite(x: float) -> bool:
    return math.isfinite(x)


class BiCGStab:
    def __init__(self, config: SolverConfig):
        self.cfg = config

    def solve(self, operator, rhs) -> SolverResult:
        """Solve ``operator u = rhs``.  em3d_amd operators run the device-resident
        solver of libem3d_hip; any other operator object (the reference's duck-typed
        protocol: ``backend`` and ``matvec``) runs the host-driven loop below."""
        fused = getattr(operator, "device_bicgstab", None)
        if fused is not None and getattr(rhs, "dim", lambda: 0)() == len(getattr(operator, "_field_shape")()):
            return fused(rhs, self.cfg)
        return self.solve_host(operator, rhs)

    def solve_batch(self, operator, rhs) -> list[SolverResult]:
        """Solve ``operator u_b = rhs[b]`` for every b of a (batch,) + field-shape
        stack in one device-resident run: each right-hand side keeps its own
        recurrence, status and history (identical to a single solve), and the
        matvecs of all running right-hand sides share one batched launch
        sequence.  No counterpart in the reference, which solves one field at a
        time (solvers/bicgstab.py:12-211)."""
        fused = getattr(operator, "device_bicgstab", None)
        if fused is None:
            return [self.solve_host(operator, rhs[b]) for b in range(rhs.shape[0])]
        return fused(rhs, self.cfg, batched=True)

    def solve_host(self, operator, rhs) -> SolverResult:
        cfg = self.cfg
        rhs = rhs.contiguous()
        rhs_norm = norm(rhs)
        res: list[float] = []
        acts: list[int] = []
        u = torch.zeros_like(rhs)
        if rhs_norm == 0.0:
            return SolverResult(u=u, iterations=0, residual_history=[0.0], converged=True, matvec_count=0,
                                residual_action_counts=[0], status="converged", true_final_residual=0.0)
        tol = max(float(cfg.rtol), float(cfg.atol) / rhs_norm)
        mv = 0
        r = lincomb(1.0, rhs, -1.0, operator.matvec(u))
        mv += 1
        rel0 = norm(r) / rhs_norm
        res.append(rel0)
        acts.append(mv)

        def result(st, it, final=None):
            final = float(res[-1]) if final is None else final
            return SolverResult(u=u, iterations=it, residual_history=res, converged=st == "converged",
                                matvec_count=mv, rmatvec_count=0, residual_action_counts=acts, status=st,
                                true_final_residual=final)

        if not _finite(rel0):
            return result("nonfinite", 0)
        if rel0 < tol:
            return result("converged", 0)
        if cfg.max_operator_actions is not None and mv >= int(cfg.max_operator_actions):
            return result("max_operator_actions", 0)

        r_hat = r.clone()
        rho_prev = 1.0 + 0.0j
        alpha = 1.0 + 0.0j
        omega = 1.0 + 0.0j
        v = torch.zeros_like(rhs)
        p = torch.zeros_like(rhs)
        s = torch.empty_like(rhs)
        status = "max_iter"
        done = 0
        for it in range(int(cfg.max_iter)):
            rho = dots([(r_hat, r)])[0]
            if abs(rho) == 0.0:
                status = "breakdown_rho"
                break
            if it == 0:
                p.copy_(r)
            else:
                if abs(omega) == 0.0:
                    status = "breakdown_omega"
                    break
                beta = (rho / rho_prev) * (alpha / omega)
                lincomb(1.0, r, beta, p, -beta * omega, v, out=p)
            if cfg.max_operator_actions is not None and mv + 1 > int(cfg.max_operator_actions):
                status = "max_operator_actions"
                break
            v = operator.matvec(p)
            mv += 1
            den = dots([(r_hat, v)])[0]
            if abs(den) == 0.0:
                status = "breakdown_alpha"
                break
            alpha = rho / den
            lincomb(1.0, r, -alpha, v, out=s)
            s_rel = norm(s) / rhs_norm
            u_alpha = lincomb(1.0, u, alpha, p)

            stop = None
            if not _finite(s_rel):
                stop = "nonfinite"
            elif cfg.divergence_guard is not None and s_rel > float(cfg.divergence_guard):
                stop = "divergence_guard"
            elif s_rel < tol:
                stop = "converged"
            elif cfg.max_operator_actions is not None and mv + 1 > int(cfg.max_operator_actions):
                stop = "max_operator_actions"
            if stop is not None:
                u, done, status = u_alpha, it + 1, stop
                res.append(s_rel)
                acts.append(mv)
                break
            t = operator.matvec(s)
            mv += 1
            tt, ts = dots([(t, t), (t, s)])
            if abs(tt) < 1e-30 * rhs_norm * rhs_norm:
                u, done, status = u_alpha, it + 1, "breakdown_t"
                res.append(s_rel)
                acts.append(mv)
                break
            omega = ts / tt
            if abs(omega) == 0.0:
                u, done, status = u_alpha, it + 1, "breakdown_omega"
                res.append(s_rel)
                acts.append(mv)
                break
            u = lincomb(1.0, u_alpha, omega, s, out=u_alpha)
            r = lincomb(1.0, s, -omega, t, out=r)
            done = it + 1
            rel = norm(r) / rhs_norm
            res.append(rel)
            acts.append(mv)
            if cfg.log:
                print(f"[BiCGStab] iter={done}, actions={mv}, rel_res={rel:.3e}")
            if not _finite(rel):
                status = "nonfinite"
                break
            if rel < tol:
                status = "converged"
                break
            if cfg.divergence_guard is not None and rel > float(cfg.divergence_guard):
                status = "divergence_guard"
                break
            rho_prev = rho
        else:
            status = "max_iter"
        return result(status, done)
