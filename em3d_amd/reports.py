"""Result records in the reference's table formats, so MI355X runs can be
put beside the reference's published CPU tables.

* ``SolverRun`` / ``save_runs_csv``: one row per solve with the columns of
  experiments/cases.py:52-78 and experiments/scans.py:232-250 (final residual
  = last entry of the residual history, experiments/scans.py:189-203);
* ``matvec_record``: the dictionary of benchmark_matvec
  (experiments/scans.py:275-300), each repetition timed from a synchronised
  device to the end of the matvec (the reference's timer runs on synchronous
  NumPy, so the device is synchronised here to measure the same interval).
"""
from __future__ import annotations

import csv
import json
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np

RUN_COLUMNS = ["case_name", "solver_name", "N", "dof", "converged", "iterations", "final_residual", "elapsed_sec"]


def _n_text(N) -> str:
    return "x".join(str(int(n)) for n in N)


@dataclass(frozen=True)
class SolverRun:
    case_name: str
    solver_name: str
    N: tuple
    dof: int
    converged: bool
    iterations: int
    final_residual: float
    elapsed_sec: float
    residual_history: list = field(default_factory=list)

    def to_row(self) -> dict[str, Any]:
        return {"case_name": self.case_name, "solver_name": self.solver_name, "N": _n_text(self.N),
                "dof": self.dof, "converged": self.converged, "iterations": self.iterations,
                "final_residual": self.final_residual, "elapsed_sec": self.elapsed_sec}


def solver_run(case_name: str, solver_name: str, N, result, elapsed_sec: float, ncomp: int = 3) -> SolverRun:
    """SolverRun of a SolverResult (dof = ncomp * cells)."""
    hist = [float(x) for x in result.residual_history]
    return SolverRun(case_name=case_name, solver_name=solver_name, N=tuple(int(n) for n in N),
                     dof=int(ncomp * np.prod(N)), converged=bool(result.converged), iterations=int(result.iterations),
                     final_residual=float(hist[-1]) if hist else float("inf"), elapsed_sec=float(elapsed_sec),
                     residual_history=hist)


def timed_solve(solver, operator, rhs, *, case_name: str, solver_name: str) -> tuple:
    """(SolverResult, SolverRun) of one solve timed between synchronised device states."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = solver.solve(operator, rhs)
    torch.cuda.synchronize()
    return res, solver_run(case_name, solver_name, operator._N, res, time.perf_counter() - t0, operator._ncomp)


def save_runs_csv(runs, path) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("w", encoding="utf-8", newline="") as f:
        w = csv.DictWriter(f, fieldnames=RUN_COLUMNS)
        w.writeheader()
        for run in runs:
            w.writerow(run.to_row())


def matvec_record(case_name: str, operator, field, *, repeats: int = 5, operator_build_sec: float = 0.0) -> dict:
    """benchmark_matvec-style timing of operator.matvec(field)."""
    import torch
    if repeats <= 0:
        raise ValueError(f"repeats must be positive, got {repeats}")
    d = []
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        operator.matvec(field)
        torch.cuda.synchronize()
        d.append(time.perf_counter() - t0)
    return {"case_name": case_name, "N": _n_text(operator._N), "dof": int(operator._ncomp * np.prod(operator._N)),
            "operator_build_sec": float(operator_build_sec), "matvec_avg_sec": float(np.mean(d)),
            "matvec_min_sec": float(np.min(d)), "matvec_max_sec": float(np.max(d)), "repeats": int(repeats)}


def save_json(data: dict, path) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(data, ensure_ascii=False, indent=2), encoding="utf-8")
