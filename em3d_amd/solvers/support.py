"""Solve on the support of the contrast, then extend to the whole grid.

No counterpart in the reference, whose solvers iterate on the whole grid
(src/em3d/solvers/bicgstab.py:12-211 on Operator, operator.py:353-363).
``SupportRestricted(solver)`` runs ``solver`` on ``A_box u_box = f_box``
(``em3d_amd.support``: the same operator on the box of cells that holds
supp(chi)), then forms the whole field with one matvec of the full operator:
u = u_box inside the box, u = f - A u~ outside (u~ = u_box extended by zero).
That matvec transforms forward only the box (``matvec_window``): chi u~ is
zero everywhere else.
The full grid's residual equals the box's inside the box and vanishes
outside, so the inner solve's tolerance is rescaled to the full right-hand
side: ``|r_box| < max(rtol |f|, atol)``.  Histories are reported relative to
|f| as well.  When the box does not pay (the contrast fills more than
``max_fraction`` of the grid, or is zero everywhere), ``solver`` runs on the
full operator unchanged.
"""
from __future__ import annotations

import copy
from dataclasses import replace

from ..support import DEFAULT_MAX_FRACTION, support_operator
from ._vec import dots
from .base import SolverResult
from .refine import MixedBiCGStab


class SupportRestricted:
    def __init__(self, solver, *, max_fraction: float = DEFAULT_MAX_FRACTION):
        if not (0.0 < max_fraction <= 1.0):
            raise ValueError(f"max_fraction must lie in (0, 1], got {max_fraction}")
        self.solver = solver
        self.max_fraction = float(max_fraction)

    @property
    def cfg(self):
        return self.solver.cfg

    def _inner(self, rtol: float):
        s = copy.copy(self.solver)
        s.cfg = replace(self.solver.cfg, rtol=rtol)
        return s

    def solve(self, operator, rhs) -> SolverResult:
        sub = support_operator(operator, self.max_fraction)
        if sub is None:
            return self.solver.solve(operator, rhs)
        f, batch = operator._prepare_input(rhs)
        if batch != 1:
            raise ValueError("SupportRestricted.solve takes one right-hand side; use solve_batch")
        fb = sub.restrict(f)
        nf, nfb = dots([(f, f), (fb, fb)])                    # one read-back for both norms
        fn, fbn = max(nf.real, 0.0) ** 0.5, max(nfb.real, 0.0) ** 0.5
        info = {"box": list(sub.box.n), "offset": list(sub.box.lo), "order": list(sub.box.perm),
                "fraction": round(sub.box.fraction, 5)}
        if fbn == 0.0:
            # u_box = 0 solves the box, and u = f + B(chi 0) = f everywhere
            return SolverResult(u=f.clone(), iterations=0, residual_history=[0.0], converged=True,
                                residual_action_counts=[0], status="converged", true_final_residual=0.0,
                                extras={"support": info})
        scale = fbn / fn
        inner = self._inner(float(self.solver.cfg.rtol) / scale)
        # the mixed solver takes |f_box| as read back above instead of reading it again
        res = inner.solve(sub, fb, rhs_norm=fbn) if isinstance(inner, MixedBiCGStab) else inner.solve(sub, fb)
        # A u~ (u~ = u_box extended by zero), its forward transforms on the box alone
        y = operator.matvec_window(sub.extend(res.u), sub.box.lo, sub.box.n)
        u = sub.extend(res.u, f, y, out=y)                   # u_box inside, f - A u~ outside
        extras = dict(res.extras or {})
        extras["support"] = info
        hist = [float(h) * scale for h in res.residual_history]
        acts = list(res.residual_action_counts)
        return SolverResult(u=u, iterations=res.iterations, residual_history=hist, converged=res.converged,
                            matvec_count=res.matvec_count + 1, rmatvec_count=res.rmatvec_count,
                            residual_action_counts=acts, status=res.status,
                            true_final_residual=None if res.true_final_residual is None
                            else float(res.true_final_residual) * scale,
                            extras=extras)

    def solve_batch(self, operator, rhs) -> list[SolverResult]:
        """Every right-hand side of a (batch,) + field-shape stack on the box in
        one batched device solve (``solver.solve_batch``), extended by one
        batched matvec of the full operator."""
        import torch
        sub = support_operator(operator, self.max_fraction)
        if sub is None:
            return self.solver.solve_batch(operator, rhs)
        f = rhs.contiguous()
        fb = sub.restrict(f)
        fn = torch.linalg.vector_norm(f.reshape(f.shape[0], -1), dim=1).tolist()
        fbn = torch.linalg.vector_norm(fb.reshape(fb.shape[0], -1), dim=1).tolist()
        if any(v == 0.0 for v in fbn) or any(v == 0.0 for v in fn):
            return [self.solve(operator, f[b]) for b in range(f.shape[0])]
        scale = [b / a for a, b in zip(fn, fbn)]
        # one batched solve takes one rtol: the tightest of the rescaled ones
        res = self._inner(min(float(self.solver.cfg.rtol) / s for s in scale)).solve_batch(sub, fb)
        ub = torch.stack([r.u for r in res])
        y = operator.matvec_window(sub.extend(ub), sub.box.lo, sub.box.n)
        u = sub.extend(ub, f, y, out=y)
        info = {"box": list(sub.box.n), "offset": list(sub.box.lo), "order": list(sub.box.perm),
                "fraction": round(sub.box.fraction, 5)}
        out = []
        for b, r in enumerate(res):
            extras = dict(r.extras or {})
            extras["support"] = info
            out.append(SolverResult(u=u[b], iterations=r.iterations,
                                    residual_history=[float(h) * scale[b] for h in r.residual_history],
                                    converged=r.converged, matvec_count=r.matvec_count + 1,
                                    rmatvec_count=r.rmatvec_count,
                                    residual_action_counts=list(r.residual_action_counts), status=r.status,
                                    true_final_residual=None if r.true_final_residual is None
                                    else float(r.true_final_residual) * scale[b], extras=extras))
        return out
