"""Host logic of the support restriction: box sides and placement.

A box side must be a length whose doubled axis the fast FFT kernels cover
(2^k, 3 2^k, 5 2^k), at least the support's extent, at most the grid's side,
and the box must hold every non-zero cell while lying inside the grid.
"""
from __future__ import annotations

import itertools

import pytest

from em3d_amd.support import FAST_SIDES, SupportBox, box_side, plan_box


def _fast(L: int) -> bool:
    for odd in (1, 3, 5):
        if L % odd == 0:
            m = L // odd
            if m >= 16 and m & (m - 1) == 0:
                return True
    return False


def test_fast_sides_are_the_fast_lengths():
    assert all(_fast(2 * n) for n in FAST_SIDES)
    assert [n for n in range(1, 257) if _fast(2 * n) and 2 * n <= 512] == list(FAST_SIDES)


@pytest.mark.parametrize("extent,n,want", [(1, 128, 8), (8, 128, 8), (9, 128, 16), (13, 32, 16), (77, 128, 80),
                                           (52, 128, 64), (39, 128, 40), (100, 96, 96), (257, 300, 300)])
def test_box_side(extent, n, want):
    assert box_side(extent, n) == want


def test_plan_box_holds_the_support_inside_the_grid():
    N = (32, 24, 20)
    for lo0, hi0 in itertools.product(range(0, 32, 5), range(0, 32, 7)):
        if hi0 < lo0:
            continue
        bounds = ((lo0, hi0), (3, 10), (18, 19))
        box = plan_box(N, bounds)
        for a in range(3):
            assert 0 <= box.lo[a] and box.lo[a] + box.n[a] <= N[a]
            assert box.lo[a] <= bounds[a][0] and bounds[a][1] < box.lo[a] + box.n[a]
            assert box.n[a] == box_side(bounds[a][1] - bounds[a][0] + 1, N[a])


def test_box_fraction():
    b = SupportBox(N=(128, 128, 128), lo=(24, 32, 44), n=(80, 64, 40), bounds=((25, 101), (38, 89), (45, 83)))
    assert b.cells == 204800 and abs(b.fraction - 204800 / 128 ** 3) < 1e-15 and not b.is_full


@pytest.mark.parametrize("n,perm", [((80, 64, 40), (1, 0, 2)), ((64, 80, 40), (0, 1, 2)), ((80, 40, 32), (2, 1, 0)),
                                    ((24, 48, 96), (0, 1, 2)), ((128, 96, 80), (0, 1, 2))])
def test_box_order_puts_a_power_of_two_side_first(n, perm):
    from em3d_amd.support import box_order
    assert box_order(n) == perm
    assert sorted(perm) == [0, 1, 2]


def test_plan_box_order_and_shape():
    bounds = ((25, 101), (38, 89), (45, 83))        # the 128^3 ellipsoid of radii (0.3, 0.2, 0.15)
    box = plan_box((128, 128, 128), bounds)
    assert box.n == (80, 64, 40) and box.perm == (1, 0, 2) and box.shape == (64, 80, 40)
