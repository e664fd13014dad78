"""Mixed-precision iterative refinement around the device BiCGStab.

No counterpart in the reference, whose solvers run in the backend's single
precision (src/em3d/solvers/bicgstab.py:12-211).  The bulk of the work -- the
inner Krylov iterations and their FFT matvecs -- runs in complex64 on the
operator's single-precision companion; FP64 is kept for the spectrum build,
the true residual ``r = f - A x`` (one FP64 matvec per outer step, the same
quantity as the reference's true relative residual,
experiments/stationary_validation.py:172-180) and the update ``x += d``:

    x = 0, r = f
    repeat:  solve A32 d = r / |r| to inner_rtol  (complex64 BiCGStab)
             x += |r| d                            (FP64)
             r = f - A64 x                         (FP64 matvec)
    until |r| / |f| < max(rtol, atol / |f|)

With a ``preconditioner`` (a trained FNO) the inner solves are the
right-preconditioned device BiCGStab of ``FlexibleBiCGStab``.  An inner solve
that does not converge within ``preconditioned_max_iter`` iterations ends the
use of the preconditioner: its partial correction is kept when the FP64
residual confirms that it helped (else undone), and the remaining outer steps
run plain inner solves.

Each outer step divides the residual by about ``inner_rtol``, so rtol 1e-8
takes two or three steps.  The default 3e-6 sits near the complex64 limit of
the inner recurrence; over three 128^3 solves (k0 = 4, 8, 12) it took less
time in total than 1e-4, 1e-5 or 1e-6.  When a step fails to reduce the true residual by
``stagnation`` (an inner breakdown, or an operator whose complex64
conditioning caps the attainable accuracy), the remaining correction is
solved by plain FP64 BiCGStab.  That solver is the fallback, and the FP64
true residual decides convergence in every case.
"""
from __future__ import annotations

import math

import torch

from .. import _lib
from ..dtypes import Precision
from ._vec import lincomb, norm
from .base import SolverConfig, SolverResult


def convert_axpy(src, a: complex, dst, accumulate: bool) -> None:
    """``dst = a src (+ dst)`` across complex64/complex128 (libem3d_hip)."""
    a = complex(a)
    code = {torch.complex64: 0, torch.complex128: 1}
    _lib.check(_lib.lib().em3d_convert_axpy(src.numel(), code[src.dtype], src.data_ptr(), a.real, a.imag,
                                            code[dst.dtype], dst.data_ptr(), 1 if accumulate else 0,
                                            _lib.stream_handle(src.device)), "em3d_convert_axpy")


class MixedBiCGStab:
    """complex64 BiCGStab corrections refined against the FP64 operator."""

    def __init__(self, config: SolverConfig, *, inner_rtol: float = 3e-6, max_outer: int = 12,
                 stagnation: float = 0.5, inner_max_iter: int | None = None, preconditioner=None,
                 preconditioned_max_iter: int = 8):
        if not (0.0 < inner_rtol < 1.0):
            raise ValueError(f"inner_rtol must lie in (0, 1), got {inner_rtol}")
        if max_outer < 1:
            raise ValueError(f"max_outer must be >= 1, got {max_outer}")
        self.cfg = config
        self.inner_rtol = float(inner_rtol)
        self.max_outer = int(max_outer)
        self.stagnation = float(stagnation)
        self.inner_max_iter = inner_max_iter
        if preconditioned_max_iter < 1:
            raise ValueError(f"preconditioned_max_iter must be >= 1, got {preconditioned_max_iter}")
        self.preconditioner = preconditioner   # FNO: inner solves become right-preconditioned
        self.preconditioned_max_iter = int(preconditioned_max_iter)   # then plain inner solves take over

    def solve(self, operator, rhs, *, rhs_norm: float | None = None) -> SolverResult:
        """``rhs_norm``: |rhs| when the caller already holds it (one device
        read-back fewer); it must be the FP64 norm of ``rhs``."""
        cfg = self.cfg
        be = operator.backend
        if be.precision is not Precision.DOUBLE:
            raise ValueError("MixedBiCGStab refines in FP64: pass the operator built on Backend.hip(Precision.DOUBLE)")
        f, batch = operator._prepare_input(rhs)
        if batch != 1:
            raise ValueError("MixedBiCGStab solves one right-hand side")
        comp = operator.single_precision_companion()
        pc = None
        if self.preconditioner is not None:
            from .flexible import attach
            pc = attach(self.preconditioner, comp)
        fn = norm(f) if rhs_norm is None else float(rhs_norm)
        x = torch.zeros_like(f)
        if fn == 0.0:
            return SolverResult(u=x, iterations=0, residual_history=[0.0], converged=True,
                                residual_action_counts=[0], status="converged", true_final_residual=0.0,
                                extras={"outer": 0, "fp64_matvecs": 0, "fp32_matvecs": 0})
        tol = max(float(cfg.rtol), float(cfg.atol) / fn)
        r = f.clone()
        r32 = torch.empty(f.shape, dtype=torch.complex64, device=f.device)
        rel = 1.0
        hist, acts = [1.0], [0]
        mv32 = mv64 = inner_it = outer = 0
        inner_status: list[str] = []
        status = "max_iter"
        budget = int(cfg.max_iter if self.inner_max_iter is None else self.inner_max_iter)
        while outer < self.max_outer:
            outer += 1
            rn = rel * fn
            convert_axpy(r, 1.0 / rn, r32, False)
            # solve no further than the remaining reduction asks (with a margin)
            want = max(self.inner_rtol, 0.5 * tol / rel)
            icfg = SolverConfig(max_iter=budget if pc is None else min(budget, self.preconditioned_max_iter),
                                rtol=min(want, 0.5), chunk=cfg.chunk, fused=getattr(cfg, "fused", True))
            saved = (x.clone(), r.clone()) if pc is not None else None   # to undo an unhelpful partial step
            ires = comp.device_bicgstab(r32, icfg, precond=pc, history=False)   # outer history only
            inner_it += ires.iterations
            mv32 += ires.matvec_count
            pc_failed = pc is not None and ires.status != "converged"
            inner_status.append(f"preconditioned {ires.status}" if pc is not None else ires.status)
            if pc_failed:
                pc = None   # the learned preconditioner missed the inner tolerance: plain inner solves from here
            convert_axpy(ires.u, rn, x, True)
            y = operator.matvec(x)
            mv64 += 1
            lincomb(1.0, f, -1.0, y, out=r)
            new = norm(r) / fn
            if pc_failed and not (math.isfinite(new) and new < rel):
                # the partial preconditioned correction did not help: undo it
                x.copy_(saved[0])
                r.copy_(saved[1])
                new = rel
            hist.append(new)
            acts.append(mv32 + mv64)
            if not math.isfinite(new):
                status = "nonfinite"
                break
            if new < tol:
                status = "converged"
                rel = new
                break
            stalled = new > self.stagnation * rel and not pc_failed
            rel = new
            if stalled:
                status = "stagnated"
                break
        fallback = None
        if status in ("stagnated", "max_iter") and math.isfinite(rel):
            # plain FP64 BiCGStab on the remaining correction A d = r
            fcfg = SolverConfig(max_iter=cfg.max_iter, rtol=0.5 * tol / rel, chunk=cfg.chunk,
                                fused=getattr(cfg, "fused", True))
            fres = operator.device_bicgstab(r, fcfg)
            fallback = fres.status
            inner_it += fres.iterations
            mv64 += fres.matvec_count
            lincomb(1.0, x, 1.0, fres.u, out=x)
            y = operator.matvec(x)
            mv64 += 1
            lincomb(1.0, f, -1.0, y, out=r)
            rel = norm(r) / fn
            hist.append(rel)
            acts.append(mv32 + mv64)
            status = "converged" if rel < tol else (fres.status if fres.status != "converged" else "max_iter")
        return SolverResult(u=x, iterations=inner_it, residual_history=hist, converged=status == "converged",
                            matvec_count=mv32 + mv64, rmatvec_count=0, residual_action_counts=acts, status=status,
                            true_final_residual=float(rel),
                            extras={"outer": outer, "fp64_matvecs": mv64, "fp32_matvecs": mv32,
                                    "inner_status": inner_status, "fallback": fallback})
