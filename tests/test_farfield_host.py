"""Host logic of the far field (no GPU): the shared-axis detection that picks
the kernel's plane-cut path, and the per-tensor cache of the packed contrast."""
from __future__ import annotations

import gc

import numpy as np
import torch

from em3d_amd import farfield as ff


def test_shared_axis_of_direction_sets():
    for plane, axis in (("xy", 2), ("xz", 1), ("yz", 0)):
        assert ff.shared_axis(ff.plane_directions(plane, 90)[1]) == axis
    rng = np.random.default_rng(0)
    d = rng.standard_normal((50, 3))
    assert ff.shared_axis(d) == -1
    d[:, 1] = 0.25                        # a cone about y: shared, nonzero component
    assert ff.shared_axis(d) == 1
    d[3, 1] = np.nextafter(0.25, 1.0)     # one ulp off: not shared
    assert ff.shared_axis(d) == -1
    assert ff.shared_axis(d[:1]) == 2     # one direction shares every axis


def test_packed_contrast_is_reused_until_the_tensor_changes():
    calls = []

    def make(t):
        calls.append(1)
        return t * 2

    t = torch.arange(6, dtype=torch.float32)
    a = ff.packed_contrast(t, make)
    b = ff.packed_contrast(t, make)
    assert a is b and len(calls) == 1
    t.add_(1.0)                           # in place: version counter moves
    c = ff.packed_contrast(t, make)
    assert len(calls) == 2 and torch.equal(c, (torch.arange(6) + 1.0) * 2)
    u = t.clone()                         # equal values, another tensor: packed on its own
    ff.packed_contrast(u, make)
    assert len(calls) == 3
    key = id(t)
    del t, a, b, c
    gc.collect()
    assert key not in ff._PACKED          # dropped with its tensor
    assert ff.packed_contrast(np.ones(3), lambda x: x + 1).tolist() == [2.0, 2.0, 2.0]


def test_packed_contrast_never_keeps_its_tensor_alive():
    """A full 3 x 3 contrast packs to itself: the cached value must not hold it."""
    from em3d_amd.operator import pack_contrast, CHI_FULL
    t = torch.randn(3, 3, 4, 4, 4, dtype=torch.complex64)
    kind, chi = ff.packed_contrast(t, pack_contrast)
    assert kind == CHI_FULL and chi is t
    kind2, chi2 = ff.packed_contrast(t, pack_contrast)     # from the cache
    assert kind2 == CHI_FULL and chi2 is t
    key = id(t)
    del t, chi, chi2
    gc.collect()
    assert key not in ff._PACKED
