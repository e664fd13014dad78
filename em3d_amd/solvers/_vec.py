"""Vector primitives of the solvers on HIP tensors (libem3d_hip kernels)."""
from __future__ import annotations

import torch

from .. import _lib
from .. import xp_hip


def _prec(x) -> int:
    return xp_hip._prec(x)


def dots(pairs) -> list[complex]:
    """``[vdot(a, b) for a, b in pairs]`` with one device-to-host copy."""
    dev = pairs[0][0].device
    buf = torch.empty((len(pairs), 2), dtype=torch.float64, device=dev)
    for i, (a, b) in enumerate(pairs):
        xp_hip.dotc_into(a, b, buf[i])
    host = buf.cpu()
    return [complex(float(host[i, 0]), float(host[i, 1])) for i in range(len(pairs))]


def norm(x) -> float:
    d = dots([(x, x)])[0]
    return float(max(d.real, 0.0)) ** 0.5


def lincomb(a, x, b=0.0, y=None, c=0.0, z=None, out=None):
    """``out = a x + b y + c z`` (y, z optional); out may alias an operand."""
    if y is None:
        y, b = x, 0.0
    if out is None:
        out = torch.empty_like(x)
    a, b, c = complex(a), complex(b), complex(c)
    _lib.check(_lib.lib().em3d_axpbypcz(_prec(x), x.numel(), a.real, a.imag, x.data_ptr(), b.real, b.imag,
                                        y.data_ptr(), c.real, c.imag, z.data_ptr() if z is not None else None,
                                        out.data_ptr(), _lib.stream_handle(x.device)), "em3d_axpbypcz")
    return out
