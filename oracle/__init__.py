"""CPU oracle for the FFT volume-integral operator and Krylov solvers.

TEST INFRASTRUCTURE ONLY.  Nothing in this package is part of the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or, in the bench, as the
timed CPU baseline).  The product path (``em3d_amd``) never imports it and
fails loudly when its HIP library is missing.

The modules restate, in plain NumPy, the algorithm of the reference package
``em3d`` 0.8.0 (qwerty29544/em3d); every function cites the reference file and
lines it follows.  Pinning (how each function is checked, independently of this
restatement) is listed per function in ``PINNING`` below and exercised by
``tests/test_oracle_*.py``:

* kernel closed form    -- finite-difference Hessian of the scalar Green
                           function (dyadic identity), self term -I/3, symmetry
* FFT matvec / rmatvec  -- brute-force dense A = I - B chi assembled cell pair by
                           cell pair from the closed form; adjoint identity
                           <Au, v> = <u, A^H v>; zero contrast gives identity
* acoustic operator     -- brute-force dense H; adjoint identity; eta = 1 gives
                           identity; self-cell low-frequency limit
* BiCGStab              -- iterate-by-iterate agreement with
                           scipy.sparse.linalg.bicgstab (same van der Vorst
                           recurrence); manufactured solutions; forced
                           statuses (zero rhs, NaN rhs, A = 0, action budgets)
* SIM                   -- closed form on a diagonal matrix (geometric error
                           decay (1 - lambda/mu)^n); manufactured solutions
* TwoStep               -- closed-form two-update recurrence on a 2x2 matrix
                           (the reference's own known-answer test); manufactured
                           solutions
* input builders        -- exact rational axes, plane-wave values at cells with
                           known phases, mask counts by integer arithmetic on
                           doubled cell indices (surface ties included)
"""

PINNING = {
    "vie.kernel_block": "finite-difference Hessian of g=exp(ikR)/(4 pi R); self -I/3; symmetry",
    "vie.em_matvec": "dense brute force (vie.dense_operator_matrix) at 1e-12",
    "vie.em_rmatvec": "dense adjoint brute force at 1e-12; <Au,v>=<u,A^H v>",
    "vie.acoustic_matvec": "dense brute force (vie.acoustic_dense_H) at 1e-12",
    "vie.acoustic_rmatvec": "adjoint identity; dense H^H",
    "solvers.bicgstab": "scipy.sparse.linalg.bicgstab iterates; manufactured solution",
    "solvers.sim": "diagonal closed form; manufactured solution",
    "solvers.twostep": "2x2 two-update closed form; manufactured solution",
    "solvers.true_relative_residual": "dense operator residual on small grids (tests/test_oracle_solvers.py)",
    "inputs.axis": "exact rational cell centres (fractions) to 4 ulp; span; antisymmetry",
    "inputs.flat_wave_vec": "values at cells whose phase is a multiple of pi/2; modulus; phase step; sign",
    "inputs.plane_wave_scalar": "values at cells whose phase is a multiple of pi/2",
    "inputs.ellipsoid_center_mask": "integer counts on doubled cell indices; 33 cells for exact surface ties",
    "inputs.box_center_mask": "integer counts per axis",
    "inputs.contrast_from_weights": "eps_f - I / eps_bg - I at chosen cells, untransposed",
    "inputs.ellipsis_refraction": "integer counts; entries of a diagonal lossy tensor",
    "inputs.cylinder_refraction": "integer disc count times the axis length",
    "inputs.eta_ellipsoid": "integer counts inside and outside",
}
