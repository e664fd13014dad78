"""Array backend: ``Backend.hip()`` puts em3d's arrays in MI355X HBM.

Mirrors the reference ``Backend`` (src/em3d/backend.py:10-119): a frozen
record of ``(xp, device, precision)`` with the same helper methods.  For the
HIP backend the arrays are ``torch`` tensors on the ROCm device and ``xp`` is
``em3d_amd.xp_hip``, a namespace whose reductions (``vdot``,
``linalg.norm``) run the library's own HIP kernels.  ``device`` is ``"hip"``
so reference helpers that special-case ``"cuda"`` are not taken.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Literal

import numpy as np

from .dtypes import Precision


@dataclass(frozen=True)
class Backend:
    xp: Any
    device: str
    precision: Precision
    device_index: int = field(default=0)

    @property
    def real_dtype(self) -> type:
        return self.precision.real_dtype

    @property
    def complex_dtype(self) -> type:
        return self.precision.complex_dtype

    @property
    def torch_device(self):
        import torch
        return torch.device("cuda", self.device_index)

    @classmethod
    def hip(cls, precision: Precision = Precision.DOUBLE, device_index: int = 0) -> "Backend":
        """MI355X backend.  Raises if torch sees no ROCm GPU or the HIP library is missing."""
        import torch

        from . import _lib, xp_hip

        if not torch.cuda.is_available():
            raise RuntimeError("Backend.hip(): no ROCm GPU is visible to torch")
        _lib.lib()
        return cls(xp=xp_hip, device="hip", precision=precision, device_index=int(device_index))

    @classmethod
    def auto(cls, precision: Precision = Precision.DOUBLE) -> "Backend":
        return cls.hip(precision)

    # -- array helpers (backend.py:47-81) -----------------------------------
    def array(self, obj, dtype=None):
        return self.xp.asarray(obj, dtype=dtype)

    def _dtype(self, kind):
        if kind == "real":
            return self.real_dtype
        if kind == "complex":
            return self.complex_dtype
        raise ValueError(f"kind must be 'real' or 'complex', got {kind!r}")

    def zeros(self, shape, kind: Literal["real", "complex"]):
        return self.xp.zeros(shape, dtype=self._dtype(kind))

    def empty(self, shape, kind: Literal["real", "complex"]):
        return self.xp.empty(shape, dtype=self._dtype(kind))

    def to_host(self, arr) -> np.ndarray:
        if hasattr(arr, "detach"):
            return arr.detach().cpu().numpy()
        return np.asarray(arr)

    def fftn(self, x, axes=None):
        raise NotImplementedError("Backend.hip has no general fftn: the operators run their own pruned FFT")

    def ifftn(self, x, axes=None):
        raise NotImplementedError("Backend.hip has no general ifftn: the operators run their own pruned FFT")

    def asarray_of_kind(self, obj, kind: Literal["real", "complex"]):
        return self.xp.asarray(obj, dtype=self._dtype(kind))

    # -- runtime helpers (backend.py:83-119) --------------------------------
    def synchronize(self) -> None:
        import torch
        torch.cuda.synchronize(self.torch_device)

    def memory_info(self) -> dict[str, int | str]:
        import torch
        free_b, total_b = torch.cuda.mem_get_info(self.torch_device)
        return {
            "device_id": int(self.device_index),
            "device_name": str(torch.cuda.get_device_name(self.torch_device)),
            "free_bytes": int(free_b),
            "total_bytes": int(total_b),
            "pool_used_bytes": int(torch.cuda.memory_allocated(self.torch_device)),
            "pool_total_bytes": int(torch.cuda.memory_reserved(self.torch_device)),
        }

    def clear_memory_pool(self) -> None:
        import torch
        self.synchronize()
        torch.cuda.empty_cache()
