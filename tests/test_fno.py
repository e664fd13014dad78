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
    import torch.nn.functional as F
    c = fno.config
    G, H, M = c.grid, c.channels, c.modes
    P = _split(fno, params)
    S = x.shape[0]
    idx = torch.tensor([m if m < M // 2 else G + m - M for m in range(M)], device=x.device)

    def dense(a, w, b, k, n):
        return _ste(a) @ _ste(w.reshape(n, k)).T + b

    v = dense(x, P["w_in"], P["b_in"], fno.in_channels, H)
    for l in range(c.layers):
        R = _ste(P[f"r{l}"]).reshape(M, M, M, H, H, 2)
        Rc = torch.complex(R[..., 0], R[..., 1])
        vs = v.reshape(S, G, G, G, H)
        A = torch.fft.fftn(vs, dim=(1, 2, 3))
        A = A.index_select(1, idx).index_select(2, idx).index_select(3, idx)
        B = torch.einsum("xyzoi,sxyzi->sxyzo", Rc, A)
        full = torch.zeros((S, G, G, G, H), dtype=B.dtype, device=x.device)
        full[:, idx[:, None, None], idx[None, :, None], idx[None, None, :]] = B
        spec = torch.fft.ifftn(full, dim=(1, 2, 3)).real.reshape(S, G ** 3, H)
        z = dense(v, P[f"w{l}"], P[f"b{l}"], H, H) + spec
        v = F.gelu(z, approximate="tanh")
    return dense(v, P["w_out"], P["b_out"], H, fno.out_channels)


def _small_fno(em, seed=0):
    import torch
    from em3d_amd.precond import FNO, FNOConfig
    fno = FNO(FNOConfig(grid=8, channels=16, modes=4, layers=2), 12, 6, device=torch.device("cuda", 0), seed=seed)
    rng = np.random.default_rng(seed + 1)
    p = fno.params.cpu().numpy()
    P = _split(fno, p)
    P["w_out"][:] = rng.standard_normal(P["w_out"].size) / 4.0     # a non-trivial output layer
    P["b_out"][:] = rng.standard_normal(P["b_out"].size) * 0.1
    for l in range(fno.config.layers):
        P[f"b{l}"][:] = rng.standard_normal(P[f"b{l}"].size) * 0.1
        P[f"r{l}"][:] *= 8.0                                          # spectral path of the same scale
    fno.params.copy_(torch.as_tensor(p))
    fno.refresh_bf16()
    return fno


def test_fno_forward_matches_torch_reference(em):
    import torch
    _say("[fno] forward 8^3 x 16 channels, 2 samples")
    fno = _small_fno(em)
    rng = np.random.default_rng(5)
    x = torch.as_tensor(rng.standard_normal((2, 512, 12)).astype(np.float32)).cuda()
    y = fno.forward(x, 2)
    ref = _reference(fno, fno.params, x)
    err = float(torch.linalg.vector_norm(y - ref) / torch.linalg.vector_norm(ref))
    assert err < 2e-3, err


def test_fno_backward_matches_autograd(em):
    import torch
    _say("[fno] backward 8^3 x 16 channels, 3 samples")
    fno = _small_fno(em, seed=3)
    rng = np.random.default_rng(6)
    x = torch.as_tensor(rng.standard_normal((3, 512, 12)).astype(np.float32)).cuda()
    t = torch.as_tensor(rng.standard_normal((3, 512, 6)).astype(np.float32)).cuda()
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
    fno = FNO(FNOConfig(grid=8, channels=8, modes=4, layers=1), 12, 6, device=torch.device("cuda", 0))
    cfg = em.SolverConfig(max_iter=100, rtol=1e-5)
    plain = em.BiCGStab(cfg).solve(op, wave)
    pre = em.FlexibleBiCGStab(cfg, fno).solve(op, wave)
    assert pre.extras["fallback"] is None
    assert (pre.status, pre.iterations, pre.matvec_count) == (plain.status, plain.iterations, plain.matvec_count)
    np.testing.assert_allclose(pre.residual_history, plain.residual_history, rtol=1e-12, atol=0)
    assert torch.equal(pre.u, plain.u)


def test_training_reduces_loss_and_the_preconditioned_solve_converges(em):
    import torch
    from em3d_amd import learn
    from em3d_amd.precond import FNO, FNOConfig
    _say("[fno] training at 32^3 (coarse 16^3), 24 materials")
    be = em.Backend.hip(em.Precision.SINGLE)
    fno = FNO(FNOConfig(grid=16, channels=16, modes=8, layers=2), 12, 6, device=torch.device("cuda", 0), seed=1)
    X, T = learn.make_pairs(fno, range(24), 32, be)
    assert X.shape == (48, 16 ** 3, 12) and T.shape == (48, 16 ** 3, 6)
    hist = learn.train(fno, X, T, steps=300, batch=8, lr=2e-3)
    assert hist[-1] < 0.8 * hist[0], hist
    mat = learn.random_material(1000)
    op, f = learn.material_operator(mat, 32, be)
    cfg = em.SolverConfig(max_iter=300, rtol=1e-5)
    res = em.FlexibleBiCGStab(cfg, fno, fallback=False).solve(op, f)
    assert res.status == "converged", res.status
    true = float(torch.linalg.vector_norm(op.matvec(res.u) - f) / torch.linalg.vector_norm(f))
    assert true < 1e-4
