"""Training of the FNO preconditioner on pairs from the library's own solves.

No counterpart in the reference.  Materials follow the learned-preconditioner
workload: a random ellipsoid or a few random spherical inclusions, isotropic
eps_r with real part U[1.5, 4.0] and imaginary part U[0, 0.5], k0 U[4, 12],
one seeded incident plane wave each.  For a material with operator A and
incident field f the training residuals are f and f - A x2 (x2: two
BiCGStab iterations): the residuals the preconditioner meets at the start
and early in a solve.

Two objectives:
* ``make_pairs`` + ``train``: regression of the coarse correction
  R(A^-1 r - r) / s (A^-1 r from a complex64 solve to rtol 1e-4);
* ``make_residual_set`` + ``train_residual``: the preconditioned residual
  |A M r - r|^2 / |r|^2 itself, differentiated through A (one matvec and one
  adjoint matvec per sample and step), so the network is only rewarded for
  corrections that reduce the residual.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Material:
    inclusions: tuple      # ((Ellipsoid, eps_r), ...)
    k0: float
    orient: tuple
    amplitude: tuple


def random_material(seed: int) -> Material:
    from .geometry import Ellipsoid
    rng = np.random.default_rng(seed)
    k0 = float(rng.uniform(4.0, 12.0))
    incs = []
    if rng.uniform() < 0.5:
        radii = tuple(float(v) for v in rng.uniform(0.12, 0.35, size=3))
        centre = tuple(float(rng.uniform(-0.45 + r, 0.45 - r)) for r in radii)
        eps = complex(rng.uniform(1.5, 4.0), rng.uniform(0.0, 0.5))
        incs.append((Ellipsoid(center=centre, radii=radii), eps))
    else:
        for _ in range(int(rng.integers(2, 9))):
            r = float(rng.uniform(0.05, 0.15))
            centre = tuple(float(rng.uniform(-0.45 + r, 0.45 - r)) for _ in range(3))
            eps = complex(rng.uniform(1.5, 4.0), rng.uniform(0.0, 0.5))
            incs.append((Ellipsoid(center=centre, radii=(r, r, r)), eps))
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    ref = np.array([0.0, 0.0, 1.0]) if abs(d[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    p = np.cross(d, ref)
    p /= np.linalg.norm(p)
    return Material(tuple(incs), k0, tuple(float(v) for v in d), tuple(float(v) for v in p))


def material_operator(mat: Material, N: int, backend, prepared=None):
    """(Operator, incident field) of a material on an N^3 grid of the unit box."""
    from .geometry import ellipsoids_contrast
    from .grid import Grid
    from .operator import Operator
    from .problem import Problem
    from .wave import flat_wave_vec
    grid = Grid(N=(N, N, N), L=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), backend=backend)
    chi = ellipsoids_contrast(grid, list(mat.inclusions))
    wave = flat_wave_vec(grid, k=mat.k0, orient=mat.orient, amplitude=mat.amplitude)
    prob = Problem(grid=grid, eps_tensor=chi, wave=wave, k0=mat.k0, volume=1.0)
    return Operator(prob, prepared_kernel=prepared), wave


def make_pairs(fno, seeds, N: int, backend, log=None):
    """Training pairs (features, targets) as fp32 device tensors [P][cells][.]."""
    import torch

    from .precond import FNOPreconditioner
    from .solvers import BiCGStab, SolverConfig
    from .solvers._vec import lincomb
    xs, ts = [], []
    t0 = time.perf_counter()
    for i, seed in enumerate(seeds):
        mat = random_material(seed)
        op, f = material_operator(mat, N, backend)
        pc = FNOPreconditioner(fno, op)
        e = BiCGStab(SolverConfig(max_iter=400, rtol=1e-4)).solve(op, f).u
        x2 = BiCGStab(SolverConfig(max_iter=2, rtol=1e-12)).solve(op, f).u
        r2 = lincomb(1.0, f, -1.0, op.matvec(x2))
        e2 = lincomb(1.0, e, -1.0, x2)
        x, t = pc.restrict_pair(torch.stack([f, r2]), torch.stack([e, e2]))
        xs.append(x)
        ts.append(t)
        if log and (i + 1) % 50 == 0:
            log(f"[fno] pairs from {i + 1} materials, {time.perf_counter() - t0:.1f} s")
        del op, pc
    return torch.cat(xs), torch.cat(ts)


def train(fno, X, T, *, steps: int, batch: int = 8, lr: float = 1e-3, lr_final: float = 1e-4, seed: int = 0,
          log=None, weight_decay: float = 0.0):
    """Adam on random minibatches with a cosine learning-rate decay; returns the
    loss history (host floats, one per 50 steps)."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    P = X.shape[0]
    hist = []
    acc = None
    for step in range(steps):
        idx = torch.randint(0, P, (batch,), generator=g).to(X.device)
        xb = X.index_select(0, idx).contiguous()
        tb = T.index_select(0, idx).contiguous()
        a = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * step / max(1, steps)))
        loss = fno.train_step(xb, tb, lr=a, weight_decay=weight_decay)
        acc = loss if acc is None else acc + loss
        if (step + 1) % 50 == 0:
            hist.append(float(acc.item()) / 50.0)
            acc = None
            if log and (step + 1) % 500 == 0:
                log(f"[fno] step {step + 1}: loss {hist[-1]:.4f}")
    return hist


@dataclass
class ResidualItem:
    op: object          # complex64 operator of the material
    pc: object          # FNOPreconditioner attached to it
    r: object           # residuals (k,) + field shape
    q: object           # A r - r
    x: object           # network input features [2k][cells][in + material] (re, im per residual)
    scale: object       # s per residual (device float64)
    rn2: object         # |r|^2 per residual (device float64)


def make_residual_set(fno, seeds, N: int, backend, log=None, iterations=(0, 2), noise: bool = False):
    """Operators, residuals and features of the residual-loss training set:
    per material the residuals f - A x_k of BiCGStab after k in `iterations`
    steps (k = 0: f), and with `noise` a complex Gaussian field, the kind of
    rough direction the solver feeds the preconditioner late in a solve."""
    import torch

    from .precond import FNOPreconditioner
    from .solvers import BiCGStab, SolverConfig
    from .solvers._vec import lincomb
    items = []
    t0 = time.perf_counter()
    for i, seed in enumerate(seeds):
        mat = random_material(seed)
        op, f = material_operator(mat, N, backend)
        pc = FNOPreconditioner(fno, op)
        rs = []
        for k in iterations:
            if k == 0:
                rs.append(f)
            else:
                xk = BiCGStab(SolverConfig(max_iter=k, rtol=1e-12)).solve(op, f).u
                rs.append(lincomb(1.0, f, -1.0, op.matvec(xk)))
        if noise:
            g = torch.Generator(device=f.device).manual_seed(seed)
            rs.append(torch.randn(f.shape, dtype=f.dtype, device=f.device, generator=g))
        r = torch.stack(rs)
        q = lincomb(1.0, op.matvec(r), -1.0, r)
        x, scale = pc.features(r)
        rn2 = (scale ** 2) * float(r[0].numel())
        items.append(ResidualItem(op, pc, r, q, x, scale, rn2))
        if log and (i + 1) % 100 == 0:
            log(f"[fno] residual set: {i + 1} materials, {time.perf_counter() - t0:.1f} s")
    return items


def residual_loss_step(fno, items, picks):
    """Accumulate the gradient of mean_s |A_s M_s r_s - r_s|^2 / |r_s|^2 over
    the picked (item, residual) pairs; returns the device loss."""
    import torch

    from .solvers._vec import lincomb
    S = len(picks)
    x = torch.cat([items[i].x[2 * j:2 * j + 2] for i, j in picks]).contiguous()   # real, imaginary samples
    y = fno.forward(x, 2 * S)
    dys = []
    loss = torch.zeros((), dtype=torch.float64, device=x.device)
    for k, (i, j) in enumerate(picks):
        it = items[i]
        z = it.pc.prolong(y[2 * k:2 * k + 2].contiguous(), it.scale[j:j + 1])
        rho = lincomb(1.0, it.q[j], 1.0, it.op.matvec(z))
        g = it.op.rmatvec(rho)
        loss = loss + torch.linalg.vector_norm(rho).to(torch.float64) ** 2 / it.rn2[j]
        # dL/dy = 2 s P^T(A^H rho) / (|r|^2 S)
        dys.append(it.pc.prolong_adjoint(g, (it.rn2[j:j + 1] * S / (2.0 * it.scale[j:j + 1])).contiguous()))
    dy = torch.cat(dys).contiguous()
    fno.backward(x, dy, 2 * S)
    return loss / S


def train_residual(fno, items, *, steps: int, batch: int = 8, lr: float = 1e-3, lr_final: float = 1e-5,
                   seed: int = 0, log=None):
    """Adam on the preconditioned-residual loss; returns the loss history
    (mean per 50 steps)."""
    rng = np.random.default_rng(seed)
    hist, acc = [], None
    for step in range(steps):
        picks = [(i, int(rng.integers(items[i].r.shape[0]))) for i in rng.integers(len(items), size=batch)]
        fno.grad.zero_()
        loss = residual_loss_step(fno, items, picks)
        a = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * step / max(1, steps)))
        fno.adam_step(lr=a)
        acc = loss if acc is None else acc + loss
        if (step + 1) % 50 == 0:
            hist.append(float(acc.item()) / 50.0)
            acc = None
            if log and (step + 1) % 250 == 0:
                log(f"[fno] step {step + 1}: preconditioned residual loss {hist[-1]:.4f}")
    return hist
