"""em3d_amd -- MI355X-native FFT volume-integral operator and Krylov solvers.

A drop-in for the hot path of em3d (qwerty29544/em3d 0.8.0): the same
``Grid``/``Problem``/``Operator``/``PreparedEMKernel``/``AcousticOperator`` and
``BiCGStab``/``SIM``/``TwoStep`` API on ``Backend.hip()``, computed by the
hand-written gfx950 HIP kernels of ``libem3d_hip.so``.
"""
from __future__ import annotations

from . import acoustics, farfield, gamma0, geometry, judge, mie, precond, support
from .backend import Backend
from .dtypes import Precision
from .grid import Grid
from .judge import judge_solve, qualify, true_relative_residual
from .operator import Operator, PreparedEMKernel
from .problem import Problem
from .refraction import apply_refraction, cylinder_refraction, ellipsis_refraction, step_refraction
from .solvers import (SIM, BaseSolver, BiCGStab, FlexibleBiCGStab, MixedBiCGStab, SolverConfig, SolverResult,
                      SupportRestricted, TwoStep)
from .wave import flat_wave_vec

__version__ = "0.1.0"

__all__ = ["Backend", "Precision", "Grid", "Problem", "Operator", "PreparedEMKernel", "apply_refraction",
           "cylinder_refraction", "ellipsis_refraction", "step_refraction", "flat_wave_vec", "geometry",
           "acoustics", "farfield", "gamma0", "mie", "support", "judge", "judge_solve", "qualify", "true_relative_residual", "BaseSolver", "BiCGStab", "FlexibleBiCGStab", "MixedBiCGStab", "SIM", "SolverConfig", "SolverResult", "SupportRestricted", "TwoStep",
           "__version__"]
