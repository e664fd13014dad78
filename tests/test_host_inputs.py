"""Host input builders of the package (plane wave, cell-centre masks, volume
fractions) against the oracle's restatement of the reference
(src/em3d/wave.py:9-37, src/em3d/geometry.py:62-148) and closed forms.

The builders take a Grid; a stand-in with the same host axes and a NumPy
backend lets them run without a GPU.  Masks and waves must equal the
reference's bit for bit (same float operations in the same order); volume
fractions are pinned by exact sums and by cells fully inside or outside.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from em3d_amd import geometry as geo
from em3d_amd.wave import flat_wave_vec
from oracle import inputs

CPLX = {"double": np.complex128, "single": np.complex64}


def _grid(N, L, center, prec):
    axes = tuple(inputs.axis(N[a], L[a], center[a], prec) for a in range(3))
    be = SimpleNamespace(complex_dtype=CPLX[prec], array=lambda x, dtype=None: np.asarray(x, dtype=dtype))
    return SimpleNamespace(N=tuple(N), L=tuple(L), center=tuple(center), backend=be, _host=axes,
                           h=tuple(L[a] / N[a] for a in range(3)))


CASES = [((12, 10, 8), (1.0, 0.8, 1.3), (0.1, -0.2, 0.05)), ((16, 16, 16), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)),
         ((7, 9, 5), (2.0, 1.5, 3.0), (0.5, 0.0, -1.0))]


@pytest.mark.parametrize("prec", ["double", "single"])
@pytest.mark.parametrize("N,L,c", CASES)
def test_plane_wave_bits(prec, N, L, c):
    g = _grid(N, L, c, prec)
    s = np.sqrt(0.5)
    for orient, amp, phi0, sign in (((0, 0, 1), (1, 0, 0), 0.0, 1), ((s, 0.0, s), (2.0, 0, 0.5j), 0.3, -1),
                                    (tuple(np.array([0.6, 0.0, 0.8])), (0, 1.0, 0), 0.0, 1)):
        got = flat_wave_vec(g, k=9.0, orient=orient, amplitude=amp, phi0=phi0, sign=sign)
        want = inputs.flat_wave_vec(N, L, c, k=9.0, orient=orient, amplitude=amp, phi0=phi0, sign=sign,
                                    precision=prec)
        assert got.dtype == want.dtype and np.array_equal(got, want), (orient, prec)
    with pytest.raises(ValueError):
        flat_wave_vec(g, k=1.0, orient=(0, 0, 1.1), amplitude=(1, 0, 0))
    with pytest.raises(ValueError):
        flat_wave_vec(g, k=1.0, orient=(0, 0, 1), amplitude=(1, 0, 0), sign=2)


@pytest.mark.parametrize("prec", ["double", "single"])
@pytest.mark.parametrize("N,L,c", CASES)
def test_center_masks_bits(prec, N, L, c):
    g = _grid(N, L, c, prec)
    # a sphere centred on a cell centre with a radius of two cells puts centres exactly on the surface
    c0 = tuple(float(g._host[a][N[a] // 2]) for a in range(3))
    for ctr, radii in ((c0, tuple(2 * L[a] / N[a] for a in range(3))), ((0.1, 0.0, -0.05), (0.3, 0.2, 0.25))):
        got = geo.center_mask(geo.Ellipsoid(center=ctr, radii=radii), g)
        want = inputs.ellipsoid_center_mask(N, L, c, ell_center=ctr, radii=radii, precision=prec)
        assert np.array_equal(got, want)
    for ctr, size in ((c0, tuple(4 * L[a] / N[a] for a in range(3))), ((0.0, 0.1, 0.0), (0.5, 0.3, 0.9))):
        got = geo.center_mask(geo.AxisAlignedBox(center=ctr, size=size), g)
        want = inputs.box_center_mask(N, L, c, box_center=ctr, size=size, precision=prec)
        assert np.array_equal(got, want)
    assert geo.center_mask(geo.FullDomain(), g).all()


def test_volume_fractions_closed_forms():
    N, L, c = (10, 8, 12), (1.0, 1.0, 1.5), (0.0, 0.0, 0.0)
    g = _grid(N, L, c, "double")
    box = geo.AxisAlignedBox(center=(0.03, -0.11, 0.2), size=(0.37, 0.41, 0.66))
    f = geo.volume_fractions(box, g)
    dv = np.prod(g.h)
    assert abs(f.sum() * dv - geo.exact_volume(box, g)) < 1e-13      # the box lies inside the domain
    assert f.min() >= 0.0 and f.max() <= 1.0
    # cells whose every corner is inside are full, cells entirely outside are empty
    lo = [np.asarray(g._host[a]) - g.h[a] / 2 for a in range(3)]
    hi = [np.asarray(g._host[a]) + g.h[a] / 2 for a in range(3)]
    bl = [box.center[a] - box.size[a] / 2 for a in range(3)]
    bh = [box.center[a] + box.size[a] / 2 for a in range(3)]
    full = np.ones(N, bool)
    empty = np.zeros(N, bool)
    for a in range(3):
        sh = [1, 1, 1]
        sh[a] = N[a]
        full &= ((lo[a] >= bl[a]) & (hi[a] <= bh[a])).reshape(sh)
        empty |= ((hi[a] <= bl[a]) | (lo[a] >= bh[a])).reshape(sh)
    assert full.any() and np.all(np.abs(f[full] - 1.0) < 1e-12) and np.all(f[empty] == 0.0)
    assert np.array_equal(geo.sampling_weights(box, g, "cell_center"), geo.center_mask(box, g).astype(float))
    assert np.all(geo.volume_fractions(geo.FullDomain(), g) == 1.0)
    assert geo.exact_volume(geo.FullDomain(), g) == 1.5
    ell = geo.Ellipsoid(center=(0, 0, 0), radii=(0.1, 0.2, 0.3))
    assert abs(geo.exact_volume(ell, g) - 4.0 * np.pi * 0.006 / 3.0) < 1e-15
    with pytest.raises(NotImplementedError):
        geo.volume_fractions(ell, g)
    with pytest.raises(TypeError):
        geo.center_mask(object(), g)
    with pytest.raises(ValueError):
        geo.AxisAlignedBox(center=(0, 0, 0), size=(0.1, 0.0, 0.1))
