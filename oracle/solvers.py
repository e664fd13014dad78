"""Oracle restatement of the reference's Krylov / stationary solvers.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Operators are any
object with ``matvec(x)`` (and ``rmatvec(x)`` for TwoStep) on NumPy arrays.
Status strings, residual-history and action-count semantics follow the
reference line by line; ``callback(u)`` (oracle-only) sees every completed
iterate so the recurrence can be pinned against ``scipy.sparse.linalg``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np


@dataclass
class Config:
    """Mirror of ``SolverConfig`` (src/em3d/solvers/base.py:7-25)."""
    max_iter: int = 200
    rtol: float = 1e-6
    atol: float = 0.0
    divergence_guard: float | None = None
    max_operator_actions: int | None = None
    mu: Optional[complex] = None
    radius: Optional[float] = None


@dataclass
class Result:
    """Mirror of ``SolverResult`` (src/em3d/solvers/base.py:28-44)."""
    u: Any
    iterations: int
    residual_history: List[float]
    converged: bool
    matvec_count: int = 0
    rmatvec_count: int = 0
    residual_action_counts: List[int] = field(default_factory=list)
    status: str = "unknown"
    true_final_residual: float | None = None


def _norm(x) -> float:
    return float(np.linalg.norm(x))


def bicgstab(op, rhs, cfg: Config, callback=None) -> Result:
    """Unpreconditioned BiCGStab from zero (src/em3d/solvers/bicgstab.py:12-211)."""
    cdt = rhs.dtype.type
    rhs_norm = _norm(rhs)
    res: list[float] = []
    acts: list[int] = []
    u = np.zeros_like(rhs)
    if rhs_norm == 0.0:
        return Result(u, 0, [0.0], True, 0, 0, [0], "converged", 0.0)
    tol = max(float(cfg.rtol), float(cfg.atol) / rhs_norm)
    mv = 0
    r = rhs - op.matvec(u)
    mv += 1
    rel0 = _norm(r) / rhs_norm
    res.append(rel0)
    acts.append(mv)
    if not np.isfinite(rel0):
        return Result(u, 0, res, False, mv, 0, acts, "nonfinite", rel0)
    if rel0 < tol:
        return Result(u, 0, res, True, mv, 0, acts, "converged", rel0)
    if cfg.max_operator_actions is not None and mv >= int(cfg.max_operator_actions):
        return Result(u, 0, res, False, mv, 0, acts, "max_operator_actions", rel0)
    r_hat = np.asarray(r, dtype=rhs.dtype).copy()
    rho_prev = 1.0 + 0.0j
    alpha = 1.0 + 0.0j
    omega = 1.0 + 0.0j
    v = np.zeros_like(rhs)
    p = np.zeros_like(rhs)
    status = "max_iter"
    done = 0
    for it in range(int(cfg.max_iter)):
        rho = complex(np.vdot(r_hat, r))
        if abs(rho) == 0.0:
            status = "breakdown_rho"
            break
        if it == 0:
            p = r.copy()
        else:
            if abs(omega) == 0.0:
                status = "breakdown_omega"
                break
            beta = (rho / rho_prev) * (alpha / omega)
            p = r + cdt(beta) * (p - cdt(omega) * v)
        if cfg.max_operator_actions is not None and mv + 1 > int(cfg.max_operator_actions):
            status = "max_operator_actions"
            break
        v = op.matvec(p)
        mv += 1
        den = complex(np.vdot(r_hat, v))
        if abs(den) == 0.0:
            status = "breakdown_alpha"
            break
        alpha = rho / den
        s = r - cdt(alpha) * v
        s_rel = _norm(s) / rhs_norm
        u_alpha = u + cdt(alpha) * p

        def _stop(st):
            res.append(s_rel)
            acts.append(mv)
            return st

        if not np.isfinite(s_rel):
            u, done, status = u_alpha, it + 1, _stop("nonfinite")
            break
        if cfg.divergence_guard is not None and s_rel > float(cfg.divergence_guard):
            u, done, status = u_alpha, it + 1, _stop("divergence_guard")
            break
        if s_rel < tol:
            u, done, status = u_alpha, it + 1, _stop("converged")
            if callback is not None:
                callback(u)
            break
        if cfg.max_operator_actions is not None and mv + 1 > int(cfg.max_operator_actions):
            u, done, status = u_alpha, it + 1, _stop("max_operator_actions")
            break
        t = op.matvec(s)
        mv += 1
        tt = complex(np.vdot(t, t))
        if abs(tt) < 1e-30 * rhs_norm * rhs_norm:
            u, done, status = u_alpha, it + 1, _stop("breakdown_t")
            break
        omega = complex(np.vdot(t, s)) / tt
        if abs(omega) == 0.0:
            u, done, status = u_alpha, it + 1, _stop("breakdown_omega")
            break
        u = u_alpha + cdt(omega) * s
        r = s - cdt(omega) * t
        done = it + 1
        rel = _norm(r) / rhs_norm
        res.append(rel)
        acts.append(mv)
        if callback is not None:
            callback(u)
        if not np.isfinite(rel):
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
    final = float(res[-1]) if res else float("inf")
    return Result(u, done, res, status == "converged", mv, 0, acts, status, final)


def sim(op, rhs, cfg: Config) -> Result:
    """Stationary iteration ``u <- u - (1/mu)(Au - f)`` (src/em3d/solvers/sim.py:13-91)."""
    if cfg.mu is None or cfg.radius is None:
        raise ValueError("SolverConfig: mu and radius must be set (call gamma0.find_params)")
    if abs(cfg.mu) == 0.0:
        raise ValueError("SolverConfig: mu must be non-zero")
    cdt = rhs.dtype.type
    gamma = 1.0 / complex(cfg.mu)
    u = np.zeros_like(rhs)
    res: list[float] = []
    acts: list[int] = []
    rhs_norm = _norm(rhs)
    if rhs_norm == 0.0:
        return Result(u, 0, [0.0], True, 0, 0, [0], "converged", 0.0)
    tol = max(float(cfg.rtol), float(cfg.atol) / rhs_norm)
    updates = 0
    mv = 0
    status = "max_iter"
    conv = False
    while True:
        r = op.matvec(u) - rhs
        mv += 1
        rel = _norm(r) / rhs_norm
        res.append(rel)
        acts.append(mv)
        if not np.isfinite(rel):
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
        u = u - cdt(gamma) * r
        updates += 1
    return Result(u, updates, res, conv, mv, 0, acts, status, float(res[-1]))


def _rinner(x, y) -> float:
    return float(np.vdot(x, y).real)


def twostep(op, rhs, cfg: Config) -> Result:
    """Two-step gradient method on A^H A (src/em3d/solvers/twostep.py:16-139)."""
    cdt = rhs.dtype.type
    rhs_norm = _norm(rhs)
    res: list[float] = []
    acts: list[int] = []
    u = np.zeros_like(rhs)
    if rhs_norm == 0.0:
        return Result(u, 0, [0.0], True, 0, 0, [0], "converged", 0.0)
    tol = max(float(cfg.rtol), float(cfg.atol) / rhs_norm)
    pu = pr = None
    updates = mv = rmv = 0
    status = "max_iter"
    while True:
        r = op.matvec(u) - rhs
        mv += 1
        rel = _norm(r) / rhs_norm
        res.append(rel)
        acts.append(mv + rmv)
        if not np.isfinite(rel):
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
        g = op.rmatvec(r)
        rmv += 1
        Hg = op.matvec(g)
        mv += 1
        a11 = _rinner(Hg, Hg)
        if not np.isfinite(a11) or a11 <= 0.0:
            status = "breakdown_gradient"
            break
        if pu is None:
            h = _rinner(g, g) / a11
            nu = u - cdt(h) * g
        else:
            dr = r - pr
            a00 = _rinner(dr, dr)
            a01 = _rinner(dr, Hg)
            b0 = _rinner(r, dr)
            b1 = _rinner(r, Hg)
            det = a00 * a11 - a01 * a01
            scale = max(abs(a00 * a11), abs(a01 * a01), 1.0)
            if abs(det) <= 1e-14 * scale:
                t, h = 0.0, b1 / a11
            else:
                t = (b0 * a11 - b1 * a01) / det
                h = (a00 * b1 - a01 * b0) / det
            nu = u - cdt(t) * (u - pu) - cdt(h) * g
        pu, pr, u = u, r, nu
        updates += 1
    return Result(u, updates, res, status == "converged", mv, rmv, acts, status, float(res[-1]))


def true_relative_residual(op, u, rhs) -> float:
    """``||A u - f|| / ||f||`` with one fresh matvec
    (src/em3d/experiments/stationary_validation.py:172-180)."""
    rn = _norm(rhs)
    rr = _norm(op.matvec(u) - rhs)
    return rr if rn == 0.0 else rr / rn


class MatrixOperator:
    """Dense-matrix operator for solver tests (tests/test_solvers.py:14-23)."""

    def __init__(self, A):
        self.A = np.asarray(A)

    def matvec(self, u):
        return self.A @ np.asarray(u)

    def rmatvec(self, u):
        return self.A.conj().T @ np.asarray(u)


class EMOperator:
    """Oracle EM operator on NumPy arrays: FFT matvec and derived-adjoint rmatvec."""

    def __init__(self, N, L, k0, chi, precision="double"):
        from .vie import kernel_hat
        self.chi = chi
        self.Khat = kernel_hat(N, L, k0, precision)

    def matvec(self, u):
        from .vie import em_matvec
        return em_matvec(self.chi, u, self.Khat)

    def rmatvec(self, v):
        from .vie import em_rmatvec
        return em_rmatvec(self.chi, v, Khat=self.Khat)


class AcousticOperator:
    """Oracle acoustic operator on NumPy arrays."""

    def __init__(self, N, L, k0, chi, precision="double"):
        from .vie import acoustic_kernel_hat
        self.chi = chi
        self.Khat = acoustic_kernel_hat(N, L, k0, precision)
        self.Khat_adj = acoustic_kernel_hat(N, L, k0, precision, adjoint=True)

    def matvec(self, u):
        from .vie import acoustic_matvec
        return acoustic_matvec(self.chi, u, self.Khat)

    def rmatvec(self, v):
        from .vie import acoustic_rmatvec
        return acoustic_rmatvec(self.chi, v, self.Khat_adj)
