"""Precision levels and dtype pairs (src/em3d/dtypes.py:9-19 of the reference)."""
from __future__ import annotations

from enum import Enum

import numpy as np


class Precision(Enum):
    SINGLE = "single"
    DOUBLE = "double"

    @property
    def real_dtype(self) -> type:
        return np.float32 if self is Precision.SINGLE else np.float64

    @property
    def complex_dtype(self) -> type:
        return np.complex64 if self is Precision.SINGLE else np.complex128

    @property
    def torch_real_dtype(self):
        import torch
        return torch.float32 if self is Precision.SINGLE else torch.float64

    @property
    def torch_complex_dtype(self):
        import torch
        return torch.complex64 if self is Precision.SINGLE else torch.complex128

    @property
    def code(self) -> int:
        """Precision code of the C-ABI (0 single, 1 double)."""
        return 0 if self is Precision.SINGLE else 1
