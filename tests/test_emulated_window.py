"""The matvec whose forward passes read only a box of cells (em3d_matvec_window),
on the host emulator of the HIP passes, against the whole-grid matvec and the
oracle.

With chi zero outside the box, chi u vanishes there whatever u holds, so the
windowed matvec must give the whole-grid result: the forward passes skip
pencils, planes and rows whose input is exactly zero, and the inverse passes
and the epilogue y = u - t cover the whole grid as before.  The cases cover
the register kernels (doubled lengths 16, 32, 256), the generic kernels
(doubled length 24), the small-grid tiles, batches, the scalar operator and
boxes that touch the grid's faces or hold a single cell.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import vie
from tests.emu import emu


@pytest.fixture(scope="module")
def lib():
    return emu.load()


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _boxed(shape, lo, n, seed, lead=()):
    """Random values inside the box [lo, lo + n) of the last three axes, zero outside."""
    x = np.zeros(lead + shape, complex)
    sl = tuple(slice(lo[a], lo[a] + n[a]) for a in range(3))
    x[(Ellipsis,) + sl] = _rand(lead + tuple(n), seed)
    return x


WINDOWS = [((8, 8, 8), (2, 1, 3), (4, 6, 2)), ((8, 8, 8), (0, 0, 0), (8, 8, 8)), ((8, 8, 8), (7, 0, 5), (1, 8, 3)),
           ((8, 8, 8), (3, 3, 3), (1, 1, 1)), ((12, 10, 8), (1, 2, 0), (9, 5, 8)), ((16, 8, 12), (5, 0, 2), (11, 3, 7)),
           ((128, 2, 8), (40, 0, 1), (60, 2, 5))]


@pytest.mark.parametrize("N,lo,n", WINDOWS)
@pytest.mark.parametrize("kind", ["diag", "full"])
def test_window_matvec_equals_whole_grid(lib, N, lo, n, kind):
    L, k = (1.0, 0.8, 1.2), 3.0
    chi = _boxed(N, lo, n, 1, (3, 3)) * 0.3
    if kind == "diag":
        for a in range(3):
            for b in range(3):
                if a != b:
                    chi[a, b] = 0
    u = _rand((3,) + N, 2)          # non-zero outside the box too: chi u is not
    op = emu.EmuOperator(lib, N, L, k, chi, 3)
    y = op.matvec(u)
    yw = op.matvec_window(u[None], lo, n)[0]
    assert np.array_equal(yw, y)
    assert _rel(yw, vie.em_matvec(chi, u, vie.kernel_hat(N, L, k))) < 1e-12


def test_window_batched_single_precision_and_scalar(lib):
    N, lo, n, L, k = (8, 8, 8), (1, 2, 3), (5, 4, 3), (1.0, 1.0, 1.0), 4.0
    chi = _boxed(N, lo, n, 3, (3,))
    full = np.zeros((3, 3) + N, complex)
    for a in range(3):
        full[a, a] = chi[a] * 0.4
    u = _rand((3, 3) + N, 4)
    Kh = vie.kernel_hat(N, L, k)
    op = emu.EmuOperator(lib, N, L, k, full, 3, precision="single", batch=3)
    y, yw = op.matvec(u), op.matvec_window(u, lo, n)
    assert np.array_equal(yw, y)
    for b in range(3):
        assert _rel(yw[b], vie.em_matvec(full, u[b], Kh)) < 1e-6
    eta = 1.0 + _boxed(N, lo, n, 5) * 0.5
    sc = emu.EmuOperator(lib, N, L, k, eta - 1.0, 1, batch=2)
    ua = _rand((2,) + N, 6)
    ya = sc.matvec_window(ua, lo, n)
    assert np.array_equal(ya, sc.matvec(ua))
    assert _rel(ya[1], vie.acoustic_matvec(eta - 1.0, ua[1], vie.acoustic_kernel_hat(N, L, k))) < 1e-12
