"""Write the config-1 reference fixtures G1 and G2 by running the reference
package itself (qwerty29544/em3d 0.8.0, imported from /root/reference/src
in the build container only; the reference never travels with the repo).

G1 (``bicgstab_sphere16_trajectory.npz``): the README's EM solve -- a
dielectric sphere (radius 0.25, eps_r 2) on a 16^3 grid of the unit box,
plane wave k0 = 4 along z polarised along x, BiCGStab rtol 1e-8, max_iter 200
(README.md:45-81; solvers/bicgstab.py:12-211): iterations, status, matvec
count, residual history, action counts, true final residual and the solution.

G2 (``rcs_xz_sphere16.json``): compare_rcs_plane of that solution against the
Mie series in the xz plane over 120 angles (mie.py:179-240,
farfield/__init__.py:86-125): both curves and shape_err / scale_ratio /
abs_rel_err.

Run from the repository root:  python tests/golden/make_reference_fixtures.py
"""
from __future__ import annotations

import json
import os
import sys
from math import prod

import numpy as np

REF_SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> None:
    sys.dont_write_bytecode = True
    sys.path.insert(0, REF_SRC)
    import em3d
    from em3d.geometry import Ellipsoid, sample_contrast_tensor

    center = (0.0, 0.0, 0.0)
    grid = em3d.Grid(N=(16, 16, 16), L=(1.0, 1.0, 1.0), center=center, backend=em3d.Backend.numpy())
    radius, eps_r, k0 = 0.25, 2.0, 4.0
    contrast = sample_contrast_tensor(grid, geometry=Ellipsoid(center=center, radii=(radius,) * 3),
                                      background_eps_r=1.0, feature_eps_r=eps_r)
    wave = em3d.flat_wave_vec(grid, k=k0, orient=(0, 0, 1), amplitude=(1, 0, 0))
    problem = em3d.Problem(grid=grid, eps_tensor=contrast, wave=wave, k0=k0, volume=prod(grid.L))
    result = em3d.BiCGStab(em3d.SolverConfig(max_iter=200, rtol=1e-8)).solve(em3d.Operator(problem), wave)
    np.savez_compressed(
        os.path.join(HERE, "bicgstab_sphere16_trajectory.npz"),
        u=np.asarray(result.u, np.complex128),
        residual_history=np.asarray(result.residual_history, np.float64),
        residual_action_counts=np.asarray(result.residual_action_counts, np.int64),
        iterations=np.int64(result.iterations), matvec_count=np.int64(result.matvec_count),
        status=np.asarray(result.status), true_final_residual=np.float64(result.true_final_residual))
    cmp = em3d.mie.compare_rcs_plane(result.u, problem, a=radius, eps_r=eps_r, n_phi=120, plane="xz")
    out = {
        "case": "README EM solve: sphere r=0.25 eps_r=2 on 16^3 unit box, k0=4, wave along z, x-polarised; "
                "BiCGStab rtol 1e-8; compare_rcs_plane xz, 120 angles",
        "iterations": int(result.iterations), "status": str(result.status),
        "phi": [float(x) for x in cmp["phi"]],
        "sigma_num": [float(x) for x in cmp["sigma_num"]],
        "sigma_mie": [float(x) for x in cmp["sigma_mie"]],
        "shape_err": float(cmp["shape_err"]), "scale_ratio": float(cmp["scale_ratio"]),
        "abs_rel_err": float(cmp["abs_rel_err"]),
    }
    with open(os.path.join(HERE, "rcs_xz_sphere16.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"G1: {result.status} in {result.iterations} iterations, {result.matvec_count} matvecs; "
          f"G2: shape_err {cmp['shape_err']:.6e}, scale_ratio {cmp['scale_ratio']:.6f}")


if __name__ == "__main__":
    main()
