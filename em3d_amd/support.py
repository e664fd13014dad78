"""The operator restricted to the support of its contrast.

``(A u)_i = u_i - sum_j G(x_i - x_j) (chi u)_j`` sums over the cells j where
chi is non-zero only (src/em3d/operator.py:353-363: the contrast multiplies the
field before the padded FFT).  Take a box of cells that holds supp(chi).  The
equations of the box's cells involve u on the box alone, through kernel values
at offsets (i - j) h, and a grid of the box's shape with the same cell size h
has exactly those values.  So

    A_box u_box = f_box

is the operator of a smaller grid (same k0, same h, the contrast cropped), and
outside the box the solution is explicit:

    u = f + B(chi u) = f + u~ - A u~,   u~ = u_box inside the box, 0 outside,

one matvec of the full operator.  The full grid's residual then equals the
box's residual inside the box and vanishes (up to rounding) outside it, so a
box solve to ``|r_box| < rtol |f|`` is a full solve to the same true relative
residual.  The reference has no counterpart: it always iterates on the whole
grid.  Nothing is approximated; the saving is the ratio of grid sizes (about
10x for the 128^3 ellipsoid of radii (0.3, 0.2, 0.15)).

The box's sides are rounded up to lengths whose doubled axis the fast FFT
kernels cover (2^k, 3 2^k, 5 2^k) and kept inside the grid.  The box grid may
take its own axis order (``SupportBox.perm``): the operator is covariant
under axis permutations when the field components and the contrast indices
permute with the axes, and the x pass, the heaviest, is fastest on a
power-of-two doubled axis (an 80 x 64 x 40 box runs its FP64 matvec in 87 us
as 64 x 80 x 40, 115 us as given).
"""
from __future__ import annotations

import ctypes as C
from collections import OrderedDict
from dataclasses import dataclass
from math import prod

from . import _lib
from .dtypes import Precision
from .operator import CHI_FULL, CHI_ISO, HipPlan, _FieldApplier, build_octant_spectrum

__all__ = ["SupportBox", "support_bounds", "plan_box", "SupportOperator", "support_operator"]

# n with 2n in the fast kernels' lengths: 2^k (16..512), 3 2^k (48..384), 5 2^k (80..320)
FAST_SIDES = (8, 16, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256)


def box_side(extent: int, n: int) -> int:
    """Smallest fast side >= extent, at most n (the grid's own side then)."""
    for m in FAST_SIDES:
        if m >= extent:
            return min(m, n)
    return n


@dataclass(frozen=True)
class SupportBox:
    N: tuple            # the full grid
    lo: tuple           # first cell of the box per grid axis
    n: tuple            # box side per grid axis
    bounds: tuple       # ((lo, hi), ...) inclusive bounds of the non-zero cells
    perm: tuple = (0, 1, 2)   # box axis a is grid axis perm[a]

    @property
    def shape(self) -> tuple:
        """Sides of the box grid, in the box's own axis order."""
        return tuple(self.n[p] for p in self.perm)

    @property
    def cells(self) -> int:
        return prod(self.n)

    @property
    def fraction(self) -> float:
        return self.cells / prod(self.N)

    @property
    def is_full(self) -> bool:
        return tuple(self.n) == tuple(self.N)


def _chi_entries(kind: int) -> int:
    return {CHI_ISO: 1, 2: 3, CHI_FULL: 9}[kind]


def support_bounds(N, kind: int, chi, precision: Precision):
    """((x_lo, x_hi), (y_lo, y_hi), (z_lo, z_hi)) of the cells where the packed
    contrast is non-zero (em3d_support_bounds, one 24-byte read-back), or None
    when it is zero everywhere."""
    import torch
    b = torch.empty(6, dtype=torch.int32, device=chi.device)
    _lib.check(_lib.lib().em3d_support_bounds(*[int(v) for v in N], precision.code, chi.data_ptr(), int(kind),
                                               b.data_ptr(), _lib.stream_handle(chi.device)),
               "em3d_support_bounds")
    h = [int(v) for v in b.cpu().tolist()]
    if h[1] < 0:
        return None
    return ((h[0], h[1]), (h[2], h[3]), (h[4], h[5]))


def _pow2(m: int) -> bool:
    return m & (m - 1) == 0


def box_order(n) -> tuple:
    """Axis order of the box grid: a power-of-two side first when the first
    side is not one (it becomes the x pass's axis), else the grid's order."""
    if _pow2(n[0]):
        return (0, 1, 2)
    for a in (1, 2):
        if _pow2(n[a]):
            return (1, 0, 2) if a == 1 else (2, 1, 0)
    return (0, 1, 2)


def plan_box(N, bounds) -> SupportBox:
    """The box of fast sides holding ``bounds``, shifted to lie inside the grid."""
    lo, n = [], []
    for a in range(3):
        b0, b1 = bounds[a]
        m = box_side(b1 - b0 + 1, int(N[a]))
        lo.append(min(b0, int(N[a]) - m))
        n.append(m)
    return SupportBox(N=tuple(int(v) for v in N), lo=tuple(lo), n=tuple(n), bounds=tuple(bounds),
                      perm=box_order(n))


# FP64-built spectra of box grids, keyed by (sides, h, k0, ncomp, precision,
# device); a few boxes are live at a time (one per material in use)
_SPECTRA: OrderedDict = OrderedDict()
_SPECTRA_KEEP = 8


def _box_spectrum(n, h, k0, ncomp, precision, device_index):
    key = (tuple(n), tuple(h), float(k0), int(ncomp), precision, int(device_index))
    hit = _SPECTRA.get(key)
    if hit is not None:
        _SPECTRA.move_to_end(key)
        return hit
    L = tuple(n[a] * h[a] for a in range(3))
    khat = build_octant_spectrum(n, L, k0, ncomp, precision, device_index, h=h)
    _SPECTRA[key] = khat
    while len(_SPECTRA) > _SPECTRA_KEEP:
        _SPECTRA.popitem(last=False)
    return khat


def _ints(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def gather(field, box: SupportBox, planes: int, ncc: int = 1):
    """(planes, box shape) crop of a (planes, grid) field in the box's axis
    order (em3d_support_gather); planes come in groups of ncc = 1, 3 (vector
    components) or 9 (3 x 3 contrast), which follow the axes."""
    import torch
    out = torch.empty((planes,) + box.shape, dtype=field.dtype, device=field.device)
    prec = 1 if field.dtype == torch.complex128 else 0
    _lib.check(_lib.lib().em3d_support_gather(prec, planes, ncc, _ints(box.N), _ints(box.n), _ints(box.lo),
                                               _ints(box.perm), field.data_ptr(), out.data_ptr(),
                                               _lib.stream_handle(field.device)),
               "em3d_support_gather")
    return out


def scatter(box_field, box: SupportBox, planes: int, ncc: int = 1, f=None, y=None, out=None):
    """Full-grid field: ``box_field`` (box axis order) inside the box, ``f - y``
    (or 0) outside (em3d_support_scatter).  ``out`` may be ``f`` or ``y``."""
    import torch
    if out is None:
        out = torch.empty((planes,) + tuple(box.N), dtype=box_field.dtype, device=box_field.device)
    prec = 1 if box_field.dtype == torch.complex128 else 0
    _lib.check(_lib.lib().em3d_support_scatter(prec, planes, ncc, _ints(box.N), _ints(box.n), _ints(box.lo),
                                                _ints(box.perm), box_field.data_ptr(),
                                                None if f is None else f.data_ptr(),
                                                None if y is None else y.data_ptr(), out.data_ptr(),
                                                _lib.stream_handle(box_field.device)),
               "em3d_support_scatter")
    return out


class SupportOperator(_FieldApplier):
    """``A_box``: the parent operator's kernel (same k0 and cell size) on the
    box grid, with the parent's packed contrast cropped to the box.  Applies,
    solves (device BiCGStab, mixed refinement through its complex64
    companion) and maps fields between the grids."""

    def __init__(self, parent, box: SupportBox):
        be = parent.backend
        hg = tuple(parent._L[a] / parent._N[a] for a in range(3))
        h = tuple(hg[p] for p in box.perm)           # cell sizes along the box's axes
        shape = box.shape
        self._init_applier(be, shape, parent._ncomp, tuple(shape[a] * h[a] for a in range(3)), parent._k0)
        self._h = h
        self.box = box
        self.parent = parent
        _, kind, chi = parent._spectrum()
        self._parent_chi = chi
        self._khat = _box_spectrum(shape, h, parent._k0, parent._ncomp, be.precision, be.device_index)
        self._chi_kind = kind
        ne = _chi_entries(kind)
        self._chi = gather(chi.reshape((-1,) + tuple(box.N)), box, ne, ne).reshape(chi.shape[:-3] + shape)

    @property
    def backend(self):
        return self._be

    def _spectrum(self):
        return self._khat, self._chi_kind, self._chi

    def matvec(self, field):
        return self._apply(field, self._khat, self._chi_kind, self._chi, False)

    def rmatvec(self, field):
        return self._apply(field, self._khat, self._chi_kind, self._chi, True)

    def device_bicgstab(self, rhs, cfg, batched: bool = False, precond=None, history: bool = True):
        return self._device_bicgstab(rhs, cfg, self._khat, self._chi_kind, self._chi, batched, precond, history)

    # ---- fields between the grids (leading batch axis optional) ----
    def _planes(self, field, shape_dims):
        lead = field.shape[:field.dim() - shape_dims]
        return lead, int(prod(lead)) if lead else 1

    def restrict(self, field):
        """The box part of a full-grid field ((batch,) + field shape), in the
        box's axis order (components too)."""
        lead, planes = self._planes(field, 3)
        return gather(field.contiguous(), self.box, planes, self._ncomp).reshape(lead + self.box.shape)

    def extend(self, box_field, f=None, y=None, out=None):
        """Full-grid field: the box values inside, ``f - y`` (or zero) outside."""
        lead, planes = self._planes(box_field, 3)
        res = scatter(box_field.contiguous(), self.box, planes, self._ncomp, f, y, out)
        return res.reshape(lead + tuple(self.box.N))


# Largest box fraction restricted by default.  Measured on the 28 held-out
# materials at 128^3 whose box covers 42-75 % of the grid (mixed refinement
# on the box, extension included, against the same solver on the grid): 1.19x
# at 75 %, 1.23-1.50x at 56-62 %, 1.51-1.73x at 42-47 %.  The largest box short
# of the whole grid is 3/4 of it when N is a power of two, 5/6 for N = 96.
DEFAULT_MAX_FRACTION = 0.85


def support_operator(operator, max_fraction: float = DEFAULT_MAX_FRACTION):
    """The cached ``SupportOperator`` of ``operator``'s current contrast, or
    None when restricting does not pay: the contrast is zero everywhere, or
    the box covers more than ``max_fraction`` of the grid.  Re-planned when
    the contrast changes (the operators re-pack it when its version counter
    moves)."""
    _, kind, chi = operator._spectrum()
    cache = getattr(operator, "_support_cache", None)
    if cache is not None and cache[0] is chi and cache[1] == chi._version:
        sub = cache[2]
    else:
        bounds = support_bounds(operator._N, kind, chi, operator.backend.precision)
        sub = None
        if bounds is not None:
            box = plan_box(operator._N, bounds)
            if not box.is_full:
                sub = SupportOperator(operator, box)
        # a full 3 x 3 contrast packs to the problem's own tensor, so an
        # in-place edit keeps the object: its version counter is part of the key
        operator._support_cache = (chi, chi._version, sub)
    if sub is None or sub.box.fraction > max_fraction:
        return None
    return sub
