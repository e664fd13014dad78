"""The ``xp`` namespace of ``Backend.hip()``.

Array creation and elementwise helpers are thin torch calls (device memory
plumbing).  The two reductions the solvers use, ``vdot`` and ``linalg.norm``,
run the library's FP64-accumulating HIP reduction (``em3d_dotc``) and return
0-d device tensors, like CuPy returns 0-d device arrays.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np
import torch

from . import _lib

pi = math.pi

_NP_TO_TORCH = {
    np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64,
    np.dtype(np.complex64): torch.complex64,
    np.dtype(np.complex128): torch.complex128,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.bool_): torch.bool,
}

_device_index = 0


def _dev():
    return torch.device("cuda", _device_index)


def torch_dtype(dtype):
    if dtype is None or isinstance(dtype, torch.dtype):
        return dtype
    return _NP_TO_TORCH[np.dtype(dtype)]


def asarray(obj, dtype=None):
    td = torch_dtype(dtype)
    if isinstance(obj, torch.Tensor):
        out = obj.to(device=_dev())
        return out.to(td) if td is not None else out
    arr = np.asarray(obj)
    if td is None:
        td = _NP_TO_TORCH.get(arr.dtype)
        if td is None:
            raise TypeError(f"unsupported dtype {arr.dtype}")
    return torch.as_tensor(np.ascontiguousarray(arr)).to(device=_dev(), dtype=td)


def zeros(shape, dtype=None):
    return torch.zeros(tuple(shape) if not isinstance(shape, int) else (shape,), dtype=torch_dtype(dtype), device=_dev())


def empty(shape, dtype=None):
    return torch.empty(tuple(shape) if not isinstance(shape, int) else (shape,), dtype=torch_dtype(dtype), device=_dev())


def zeros_like(x):
    return torch.zeros_like(x)


def empty_like(x):
    return torch.empty_like(x)


def conj(x):
    return torch.conj(x).resolve_conj()


def _prec(x) -> int:
    if x.dtype == torch.complex64:
        return 0
    if x.dtype == torch.complex128:
        return 1
    raise TypeError(f"HIP reductions take complex64/complex128 tensors, got {x.dtype}")


_red_work: dict[tuple[int, int], torch.Tensor] = {}


def _work(dev) -> torch.Tensor:
    """Partial-sum scratch of em3d_dotc, one per (device, stream): reductions
    issued on two streams at once never share partials."""
    key = (dev.index if dev.index is not None else 0, _lib.stream_handle(dev))
    w = _red_work.get(key)
    if w is None:
        w = torch.empty(int(_lib.lib().em3d_reduce_work_bytes()), dtype=torch.uint8, device=dev)
        _red_work[key] = w
    return w


def dotc_into(a, b, out) -> None:
    """out (2 float64 on the device) = sum conj(a) b, FP64 accumulation."""
    if a.dtype != b.dtype or a.numel() != b.numel():
        raise ValueError("vdot operands must match in dtype and size")
    a = a.contiguous()
    b = b.contiguous()
    _lib.check(_lib.lib().em3d_dotc(_prec(a), a.numel(), a.data_ptr(), b.data_ptr(), out.data_ptr(),
                                    _work(a.device).data_ptr(), _lib.stream_handle(a.device)), "em3d_dotc")


def vdot(a, b):
    """``sum(conj(a) * b)`` over all elements, as a 0-d complex128 device tensor."""
    out = torch.empty(2, dtype=torch.float64, device=a.device)
    dotc_into(a, b, out)
    return torch.view_as_complex(out)


def _norm(x, *args, **kwargs):
    if args or kwargs:
        return torch.linalg.vector_norm(x, *args, **kwargs)
    out = torch.empty(2, dtype=torch.float64, device=x.device)
    dotc_into(x, x, out)
    return torch.sqrt(out[0].clamp_min(0.0))


linalg = SimpleNamespace(norm=_norm)


def abs(x):  # noqa: A001 - namespace mirrors numpy
    return torch.abs(x)


def isfinite(x):
    return torch.isfinite(x)


def exp(x):
    return torch.exp(x)


def sqrt(x):
    return torch.sqrt(x)


def where(c, a, b):
    return torch.where(c, a, b)


def meshgrid(*xs, indexing="ij"):
    return torch.meshgrid(*xs, indexing=indexing)


def linspace(start, stop, num, dtype=None):
    return torch.linspace(start, stop, num, dtype=torch_dtype(dtype) or torch.float64, device=_dev())
