"""Scalar acoustic scattering on the HIP backend."""
from __future__ import annotations

from . import farfield
from .kernel import prep_coeffs_acoustic, self_cell_coefficient
from .materials import eta_ellipsoid, eta_homogeneous, eta_inclusions, eta_slab, eta_sphere, random_spheres
from .operator import AcousticOperator
from .problem import AcousticProblem, make_acoustic_problem, plane_wave_scalar

__all__ = ["AcousticOperator", "AcousticProblem", "eta_ellipsoid", "eta_homogeneous", "eta_inclusions", "eta_slab",
           "farfield", "random_spheres",
           "eta_sphere", "make_acoustic_problem", "plane_wave_scalar", "prep_coeffs_acoustic",
           "self_cell_coefficient"]
