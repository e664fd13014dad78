"""Acoustic operator ``H u = u - B((eta - 1) u)`` on the HIP backend
(src/em3d/acoustics/operator.py:25-73 of the reference)."""
from __future__ import annotations

from ..operator import CHI_ISO, _FieldApplier, eta_contrast, require_hip
from .kernel import prep_coeffs_acoustic
from .problem import AcousticProblem


class AcousticOperator(_FieldApplier):
    def __init__(self, problem: AcousticProblem):
        self.problem = problem
        be = problem.grid.backend
        require_hip(be)
        self._init_applier(be, problem.grid.N, 1, problem.grid.L, problem.k0)
        self._K_hat = prep_coeffs_acoustic(problem.grid, k0=problem.k0)
        self._spectrum()

    def _spectrum(self):
        """Spectrum and ``chi = eta - 1``, recomputed when ``problem.eta``
        changes in place (the reference forms ``problem.chi`` on every call,
        acoustics/operator.py:47)."""
        _, chi = self._refresh_contrast(self.problem.eta, eta_contrast)
        return self._K_hat, CHI_ISO, chi

    @property
    def backend(self):
        return self._be

    def matvec(self, u):
        """``u - B(chi u)``."""
        return self._apply(u, *self._spectrum(), False)

    def device_bicgstab(self, rhs, cfg, batched: bool = False, precond=None):
        """BiCGStab with all iteration scalars on the device (libem3d_hip)."""
        return self._device_bicgstab(rhs, cfg, *self._spectrum(), batched, precond)

    def rmatvec(self, v):
        """``v - conj(chi) B^H v``."""
        return self._apply(v, *self._spectrum(), True)

    def to_dense(self):
        raise RuntimeError("AcousticOperator.to_dense requires numpy backend")
