from .base import BaseSolver, SolverConfig, SolverResult
from .bicgstab import BiCGStab
from .flexible import FlexibleBiCGStab
from .refine import MixedBiCGStab
from .sim import SIM
from .support import SupportRestricted
from .twostep import TwoStep

__all__ = ["BaseSolver", "SolverConfig", "SolverResult", "SIM", "BiCGStab", "FlexibleBiCGStab", "MixedBiCGStab", "SupportRestricted", "TwoStep"]
