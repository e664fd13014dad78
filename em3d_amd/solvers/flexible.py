"""Right-preconditioned (flexible) BiCGStab with a learned preconditioner.

No counterpart in the reference (its BiCGStab is unpreconditioned,
src/em3d/solvers/bicgstab.py:12-211).  The device-resident recurrence of
``em3d_bicgstab_pc`` applies the preconditioner to p and s and keeps the
preconditioned directions for the update, so a nonlinear network is a valid
preconditioner; the history holds the same recursive residuals of A x = f as
the plain solver.  If the preconditioned run does not converge, plain
BiCGStab solves the system again from zero: it stays the fallback.
"""
from __future__ import annotations

from .base import SolverConfig, SolverResult
from .bicgstab import BiCGStab


def attach(preconditioner, operator):
    """An FNOPreconditioner for `operator` (the network itself, or one already
    attached to that operator; attachments are cached on the network)."""
    from ..precond import FNO, FNOPreconditioner
    if isinstance(preconditioner, FNOPreconditioner):
        if preconditioner.operator is not operator:
            raise ValueError("the preconditioner is attached to another operator")
        return preconditioner
    if not isinstance(preconditioner, FNO):
        raise TypeError(f"unsupported preconditioner {type(preconditioner)!r}")
    cache = preconditioner.__dict__.setdefault("_attached", {})
    pc = cache.get(id(operator))
    if pc is None or pc.operator is not operator:
        pc = FNOPreconditioner(preconditioner, operator)
        cache.clear()
        cache[id(operator)] = pc
    return pc


class FlexibleBiCGStab:
    def __init__(self, config: SolverConfig, preconditioner, *, fallback: bool = True,
                 preconditioned_max_iter: int | None = None):
        self.cfg = config
        self.preconditioner = preconditioner
        self.fallback = fallback
        self.preconditioned_max_iter = preconditioned_max_iter   # budget of the preconditioned attempt

    def solve(self, operator, rhs) -> SolverResult:
        pc = attach(self.preconditioner, operator)
        cfg = self.cfg
        if self.preconditioned_max_iter is not None:
            cfg = SolverConfig(**{**cfg.__dict__, "max_iter": min(cfg.max_iter, self.preconditioned_max_iter)})
        res = operator.device_bicgstab(rhs, cfg, precond=pc)
        res.extras = {"preconditioner": "fno", "fallback": None}
        if res.status == "converged" or not self.fallback:
            return res
        plain = BiCGStab(self.cfg).solve(operator, rhs)
        plain.extras = {"preconditioner": None, "fallback": f"plain BiCGStab after preconditioned {res.status}",
                        "preconditioned_iterations": res.iterations}
        return plain
