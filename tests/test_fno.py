"""GPU tests of the learned preconditioner: the FNO forward and backward
kernels against a plain PyTorch fp32 reference of the same network (with the
bf16 rounding of the MFMA inputs reproduced), the preconditioned solver
against the plain one, and a short training run."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def em():
    import em3d_amd
    return em3d_amd


def _split(fno, p):
    out, at = {}, 0
    for name, size in fno._sizes():
        out[name] = p[at:at + size]
        at += size
    return out


def _bf(t):
    import torch
    return t.to(torch.bfloat16).to(torch.float32)


def _ste(t):
    """bf16 rounding in the forward value, identity in the gradient."""
    return t + (_bf(t) - t).detach()


def _reference(fno, params, x):
    """The network of e3_fno.hip in PyTorch fp32: pointwise products take
    bf16-rounded inputs and weights, the spectral weights are bf16-rounded."""
    import torch
    c = fno.config
    G, H, M = c.grid, c.channels, c.modes
    P = _split(fno, params)
    S = x.shape[0]
    nin, nm = fno.in_channels, fno.material_channels
    xr, xm = x[..., :nin], x[..., nin:nin + nm]
    idx = torch.tensor([m if m < M // 2 else G + m - M for m in range(M)], device=x.device)

    def dense(a, w, k, n):
        return _ste(a) @ _ste(w.reshape(n, k)).T

    v = dense(xr, P["w_in"], nin, H)
    for l in range(c.layers):
        g = torch.sigmoid(dense(xm, P[f"u{l}"], nm, H) + P[f"d{l}"])
        R = _ste(P[f"r{l}"]).reshape(M, M, M, H, H, 2)
        Rc = torch.complex(R[..., 0], R[..., 1])
        vs = v.reshape(S, G, G, G, H)
        A = torch.fft.fftn(vs, dim=(1, 2, 3))
        A = A.index_select(1, idx).index_select(2, idx).index_select(3, idx)
        B = torch.einsum("xyzoi,sxyzi->sxyzo", Rc, A)
        full = torch.zeros((S, G, G, G, H), dtype=B.dtype, device=x.device)
        full[:, idx[:, None, None], idx[None, :, None], idx[None, None, :]] = B
        spec = torch.fft.ifftn(full, dim=(1, 2, 3)).real.reshape(S, G ** 3, H)
        v = g * (dense(v, P[f"w{l}"], H, H) + spec)
    return dense(v, P["w_out"], H, fno.out_channels)


def _small_fno(em, seed=0):
    import torch
    from em3d_amd.precond import FNO, FNOConfig
    fno = FNO(FNOConfig(grid=8, channels=16, modes=4, layers=2), 3, 3, material_channels=7,
              device=torch.device("cuda", 0), seed=seed)
    rng = np.random.default_rng(seed + 1)
    p = fno.params.cpu().numpy()
    P = _split(fno, p)
    P["w_out"][:] = rng.standard_normal(P["w_out"].size) / 4.0     # a non-trivial output layer
    for l in range(fno.config.layers):
        P[f"d{l}"][:] = rng.standard_normal(P[f"d{l}"].size)
        P[f"r{l}"][:] *= 8.0                                          # spectral path of the same scale
    fno.params.copy_(torch.as_tensor(p))
    fno.refresh_bf16()
    return fno


def test_fno_forward_matches_torch_reference(em):
    import torch
    _say("[fno] forward 8^3 x 16 channels, 2 samples")
    fno = _small_fno(em)
    rng = np.random.default_rng(5)
    x = torch.as_tensor(rng.standard_normal((2, 512, 10)).astype(np.float32)).cuda()
    y = fno.forward(x, 2)
    ref = _reference(fno, fno.params, x)
    err = float(torch.linalg.vector_norm(y - ref) / torch.linalg.vector_norm(ref))
    assert err < 2e-3, err


def test_fno_backward_matches_autograd(em):
    import torch
    _say("[fno] backward 8^3 x 16 channels, 3 samples")
    fno = _small_fno(em, seed=3)
    rng = np.random.default_rng(6)
    x = torch.as_tensor(rng.standard_normal((3, 512, 10)).astype(np.float32)).cuda()
    t = torch.as_tensor(rng.standard_normal((3, 512, 3)).astype(np.float32)).cuda()
    fno.grad.zero_()
    y = fno.forward(x, 3)
    loss, dy = fno.loss_and_grad(y, t, 3)
    fno.backward(x, dy, 3)
    p = fno.params.clone().requires_grad_(True)
    yr = _reference(fno, p, x)
    lr = (((yr - t) ** 2).sum(dim=(1, 2)) / (t ** 2).sum(dim=(1, 2))).mean()
    lr.backward()
    assert abs(float(loss) - float(lr.detach())) <= 1e-3 * float(lr.detach())
    G, R = _split(fno, fno.grad), _split(fno, p.grad)
    for name in G:
        g, r = G[name], R[name]
        rel = float(torch.linalg.vector_norm(g - r) / max(float(torch.linalg.vector_norm(r)), 1e-30))
        assert rel < 3e-2, (name, rel)


def test_identity_network_gives_the_plain_trajectory(em):
    """A fresh network has W_out = 0, so M = I exactly: the preconditioned
    device BiCGStab must reproduce the plain solver's trajectory."""
    import torch
    from em3d_amd.precond import FNO, FNOConfig
    _say("[fno] identity preconditioner, 16^3")
    be = em.Backend.hip(em.Precision.SINGLE)
    grid = em.Grid(N=(16, 16, 16), L=(1, 1, 1), center=(0, 0, 0), backend=be)
    chi = em.geometry.ellipsoids_contrast(grid, [(em.geometry.Ellipsoid(center=(0, 0, 0), radii=(0.3, 0.2, 0.25)),
                                                  2.5 + 0.2j)])
    wave = em.flat_wave_vec(grid, k=6.0, orient=(0, 0, 1), amplitude=(1, 0, 0))
    op = em.Operator(em.Problem(grid=grid, eps_tensor=chi, wave=wave, k0=6.0, volume=1.0))
    fno = FNO(FNOConfig(grid=8, channels=8, modes=4, layers=1), 3, 3, material_channels=7,
              device=torch.device("cuda", 0))
    cfg = em.SolverConfig(max_iter=100, rtol=1e-5)
    # the preconditioned solver runs the stand-alone vector kernels: compare with the same kernels
    plain = em.BiCGStab(em.SolverConfig(max_iter=100, rtol=1e-5, fused=False)).solve(op, wave)
    pre = em.FlexibleBiCGStab(cfg, fno).solve(op, wave)
    assert pre.extras["fallback"] is None
    assert (pre.status, pre.iterations, pre.matvec_count) == (plain.status, plain.iterations, plain.matvec_count)
    np.testing.assert_allclose(pre.residual_history, plain.residual_history, rtol=1e-12, atol=0)
    assert torch.equal(pre.u, plain.u)


def test_preconditioner_is_linear_in_the_residual(em):
    """M(a r1 + b r2) = a M r1 + b M r2 for complex a, b and a random network,
    up to the bf16 rounding of the MFMA inputs (8-bit mantissa: the rounding
    of a x and of x differ by up to 2^-9 relative per element)."""
    import torch
    from em3d_amd.precond import FNOPreconditioner
    _say("[fno] linearity of the preconditioner, 16^3")
    fno = _small_fno(em, seed=4)
    be = em.Backend.hip(em.Precision.SINGLE)
    grid = em.Grid(N=(16, 16, 16), L=(1, 1, 1), center=(0, 0, 0), backend=be)
    chi = em.geometry.ellipsoids_contrast(grid, [(em.geometry.Ellipsoid(center=(0, 0, 0), radii=(0.3,) * 3), 3.0)])
    op = em.Operator(em.Problem(grid=grid, eps_tensor=chi, wave=chi[0], k0=5.0, volume=1.0))
    pc = FNOPreconditioner(fno, op)
    rng = np.random.default_rng(9)
    r1, r2 = (be.array((rng.standard_normal((3, 16, 16, 16)) + 1j * rng.standard_normal((3, 16, 16, 16)))
                       .astype(np.complex64)) for _ in range(2))
    a, b = 0.7 - 0.2j, -1.3 + 0.5j
    lhs = pc.apply(a * r1 + b * r2)
    rhs = a * pc.apply(r1) + b * pc.apply(r2)
    assert float(torch.linalg.vector_norm(lhs - rhs) / torch.linalg.vector_norm(rhs)) < 2e-3
    assert float(torch.linalg.vector_norm(pc.apply(r1) - r1) / torch.linalg.vector_norm(r1)) > 1e-3


def test_training_reduces_loss_and_solves_stay_correct(em):
    """Residual-loss training lowers the loss; the learned-preconditioner
    solvers (flexible BiCGStab with its plain fallback, and the mixed
    refinement with preconditioned inner solves) return solutions whose true
    residual meets the tolerance."""
    import torch
    from em3d_amd import learn
    from em3d_amd.precond import FNO, FNOConfig
    _say("[fno] residual-loss training at 32^3 (coarse 16^3), 24 materials")
    be = em.Backend.hip(em.Precision.SINGLE)
    fno = FNO(FNOConfig(grid=16, channels=16, modes=8, layers=2), 3, 3, material_channels=7,
              device=torch.device("cuda", 0), seed=1)
    items = learn.make_residual_set(fno, range(24), 32, be, iterations=(0, 2, 6), noise=True)
    hist = learn.train_residual(fno, items, steps=200, batch=4, lr=2e-3)
    assert hist[-1] < 0.95 * hist[0], hist
    mat = learn.random_material(1000)
    op, f = learn.material_operator(mat, 32, be)
    cfg = em.SolverConfig(max_iter=300, rtol=1e-5)
    res = em.FlexibleBiCGStab(cfg, fno, preconditioned_max_iter=60).solve(op, f)
    assert res.status == "converged", res.status
    true = float(torch.linalg.vector_norm(op.matvec(res.u) - f) / torch.linalg.vector_norm(f))
    assert true < 1e-4
    be64 = em.Backend.hip(em.Precision.DOUBLE)
    op64, f64 = learn.material_operator(mat, 32, be64)
    mixed = em.MixedBiCGStab(em.SolverConfig(max_iter=300, rtol=1e-8), preconditioner=fno).solve(op64, f64)
    assert mixed.status == "converged"
    true64 = float(torch.linalg.vector_norm(op64.matvec(mixed.u) - f64) / torch.linalg.vector_norm(f64))
    assert true64 < 1e-8


def test_mixed_refinement_drops_a_harmful_preconditioner(em):
    """A network with a large random output layer is a harmful M: the
    preconditioned inner solve misses its tolerance within its budget, the
    refinement keeps its partial correction only if the FP64 residual drops
    (else restores x and r exactly), then finishes with plain inner solves
    and meets the FP64 tolerance."""
    import torch
    from em3d_amd import learn
    from em3d_amd.precond import FNO, FNOConfig
    _say("[fno] harmful preconditioner in the mixed refinement, 32^3")
    fno = FNO(FNOConfig(grid=16, channels=16, modes=8, layers=2), 3, 3, material_channels=7,
              device=torch.device("cuda", 0), seed=5)
    p = fno.params.cpu().numpy()
    P = _split(fno, p)
    P["w_out"][:] = np.random.default_rng(9).standard_normal(P["w_out"].size) * 30.0
    fno.params.copy_(torch.as_tensor(p))
    fno.refresh_bf16()
    be64 = em.Backend.hip(em.Precision.DOUBLE)
    op64, f64 = learn.material_operator(learn.random_material(2024), 32, be64)
    cfg = em.SolverConfig(max_iter=300, rtol=1e-8)
    plain = em.MixedBiCGStab(cfg).solve(op64, f64)
    res = em.MixedBiCGStab(cfg, preconditioner=fno, preconditioned_max_iter=4).solve(op64, f64)
    assert res.status == "converged", res.extras
    assert res.extras["inner_status"][0].startswith("preconditioned")
    assert not any(s.startswith("preconditioned") for s in res.extras["inner_status"][1:])
    true = float(torch.linalg.vector_norm(op64.matvec(res.u) - f64) / torch.linalg.vector_norm(f64))
    assert true < 1e-8
    rel = float(torch.linalg.vector_norm(res.u - plain.u) / torch.linalg.vector_norm(plain.u))
    assert rel < 1e-6
    # the refinement's history never rises: an unhelpful partial step is undone
    assert all(b <= a * (1 + 1e-12) for a, b in zip(res.residual_history, res.residual_history[1:]))
    with pytest.raises(ValueError):
        em.MixedBiCGStab(cfg, preconditioner=fno, preconditioned_max_iter=0)
