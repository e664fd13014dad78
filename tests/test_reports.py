"""Result records keep the reference's table formats (CPU)."""
from __future__ import annotations

import csv

from em3d_amd import reports
from em3d_amd.solvers.base import SolverResult


def test_solver_run_row_and_csv(tmp_path):
    res = SolverResult(u=None, iterations=7, residual_history=[1.0, 0.1, 3e-9], converged=True, matvec_count=15,
                       status="converged")
    run = reports.solver_run("sphere_eps2_k0a1_N16", "BiCGStab", (16, 16, 16), res, 0.25)
    row = run.to_row()
    assert list(row) == reports.RUN_COLUMNS
    assert row["N"] == "16x16x16" and row["dof"] == 3 * 16 ** 3 and row["final_residual"] == 3e-9
    assert row["iterations"] == 7 and row["converged"] is True and row["elapsed_sec"] == 0.25
    p = tmp_path / "t" / "solver_runs.csv"
    reports.save_runs_csv([run, run], p)
    rows = list(csv.DictReader(p.open()))
    assert len(rows) == 2 and list(rows[0]) == reports.RUN_COLUMNS and rows[0]["N"] == "16x16x16"
    empty = SolverResult(u=None, iterations=0, residual_history=[], converged=False)
    assert reports.solver_run("c", "SIM", (2, 3, 4), empty, 0.0, ncomp=1).to_row()["final_residual"] == float("inf")
