"""The solve judge: true relative residual and qualification
(src/em3d/experiments/stationary_validation.py:172-180 and 186-232 of the
reference).

A solver's own stopping flag is not the verdict.  ``true_relative_residual``
recomputes ``|A u - f| / |f|`` with one fresh matvec of the operator it is
given (for the mixed-precision solvers: the FP64 operator), and ``qualify``
applies the reference's rule: converged, finite and at most
``tol (1 + 1e-10)`` with ``tol = max(rtol, atol / |f|)`` unless an explicit
qualification rtol is given.  ``judge_solve`` times ``solve()`` between two
device synchronisations, as the reference's protocol does, then judges it.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

__all__ = ["true_relative_residual", "qualification_tolerance", "qualify", "Judged", "judge_solve"]


def _norm(backend, x) -> float:
    return float(backend.to_host(backend.xp.linalg.norm(x)))


def true_relative_residual(operator, solution, rhs) -> float:
    """``|A u - f| / |f|`` (``|A u - f|`` when ``f = 0``), one fresh matvec."""
    be = operator.backend
    rhs_norm = _norm(be, rhs)
    residual_norm = _norm(be, operator.matvec(solution) - rhs)
    if rhs_norm == 0.0:
        return residual_norm
    return residual_norm / rhs_norm


def qualification_tolerance(config, rhs_norm: float, qualification_rtol: float | None = None) -> float:
    """``qualification_rtol`` if given, else ``max(rtol, atol / max(|f|, 1e-300))``."""
    if qualification_rtol is not None:
        return float(qualification_rtol)
    return max(float(config.rtol), float(config.atol) / max(float(rhs_norm), 1e-300))


def qualify(result, true_residual: float, tolerance: float) -> bool:
    """The reference's rule: converged, finite, ``<= tolerance (1 + 1e-10)``."""
    return bool(result.converged and math.isfinite(true_residual) and true_residual <= tolerance * (1.0 + 1e-10))


@dataclass
class Judged:
    result: object
    elapsed_seconds: float
    true_residual: float
    tolerance: float
    qualified: bool


def judge_solve(solver, operator, rhs, *, config=None, qualification_rtol: float | None = None,
                judge_operator=None) -> Judged:
    """Run ``solver.solve(operator, rhs)`` between device synchronisations,
    recompute the true residual with ``judge_operator`` (default: the solved
    operator) and qualify it.  ``result.true_final_residual`` is set to the
    recomputed value, as the reference's protocol does."""
    be = operator.backend
    be.synchronize()
    t0 = time.perf_counter()
    result = solver.solve(operator, rhs)
    be.synchronize()
    elapsed = time.perf_counter() - t0
    judge = judge_operator if judge_operator is not None else operator
    true = true_relative_residual(judge, result.u, rhs)
    result.true_final_residual = float(true)
    cfg = config if config is not None else solver.cfg
    tol = qualification_tolerance(cfg, _norm(be, rhs), qualification_rtol)
    return Judged(result=result, elapsed_seconds=float(elapsed), true_residual=float(true), tolerance=tol,
                  qualified=qualify(result, true, tol))
