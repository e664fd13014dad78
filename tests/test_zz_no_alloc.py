"""em3d_matvec_noalloc (fixed device scratch, no allocation) against
em3d_matvec with a caller workspace, through the C-ABI: bit-equal outputs on
two calls in a row, the second on a smaller grid, and a clear error for a
plan whose workspace exceeds the fixed scratch."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(em, N, prec, seed):
    import torch
    from em3d_amd import _lib
    from em3d_amd.operator import HipPlan
    be = em.Backend.hip(prec)
    grid = em.Grid(N=N, L=(1.0, 0.9, 0.8), center=(0, 0, 0), backend=be)
    rng = np.random.default_rng(seed)
    chi = (rng.standard_normal((3, 3) + N) + 1j * rng.standard_normal((3, 3) + N)) * 0.3
    u = rng.standard_normal((3,) + N) + 1j * rng.standard_normal((3,) + N)
    prob = em.Problem(grid=grid, eps_tensor=be.array(chi.astype(be.complex_dtype)),
                      wave=be.array(u.astype(be.complex_dtype)), k0=4.0, volume=1.0)
    op = em.Operator(prob)
    khat, kind, chi_d = op._spectrum()
    plan = HipPlan.get(N, 3, 1, be.precision, be.device_index)
    ud = prob.wave.contiguous()
    ref = op.matvec(ud)
    y = torch.full_like(ud, float("nan"))
    lib = _lib.lib()
    _lib.check(lib.em3d_matvec_noalloc(plan.handle, khat.data_ptr(), chi_d.data_ptr(), kind, ud.data_ptr(),
                                       y.data_ptr(), _lib.stream_handle(be.torch_device)), "em3d_matvec_noalloc")
    return y, ref


def test_noalloc_matvec_equals_workspace_matvec():
    import torch

    import em3d_amd as em
    print("[noalloc] 40x36x32 double, then 12x10x8 single, then 64^3 single", flush=True)
    sys.stdout.flush()
    for N, prec, seed in (((40, 36, 32), em.Precision.DOUBLE, 1), ((12, 10, 8), em.Precision.SINGLE, 2),
                          ((64, 64, 64), em.Precision.SINGLE, 3), ((12, 10, 8), em.Precision.DOUBLE, 4)):
        y, ref = _case(em, N, prec, seed)
        torch.cuda.synchronize()
        assert torch.equal(y, ref), N


def test_noalloc_rejects_a_plan_above_the_scratch():
    import torch

    import em3d_amd as em
    from em3d_amd import _lib
    from em3d_amd.operator import HipPlan
    plan = HipPlan.get((64, 64, 64), 3, 1, em.Precision.DOUBLE)
    assert plan.work_bytes > 64 * 1024 * 1024
    u = torch.zeros((3, 64, 64, 64), dtype=torch.complex128, device="cuda")
    y = torch.empty_like(u)
    k = torch.zeros(plan.khat_bytes // 16, dtype=torch.complex128, device="cuda")
    rc = _lib.lib().em3d_matvec_noalloc(plan.handle, k.data_ptr(), None, 0, u.data_ptr(), y.data_ptr(),
                                        _lib.stream_handle(u.device))
    assert rc != 0
    assert b"fixed scratch" in _lib.lib().em3d_last_error()
