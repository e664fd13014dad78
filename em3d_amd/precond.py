"""Learned right preconditioner: a Fourier neural operator on a coarse grid.

No counterpart in the reference (its solvers are unpreconditioned,
src/em3d/solvers/bicgstab.py:12-211).  The network and its training run in
libem3d_hip (forward, backward and Adam kernels; bf16 MFMA channel mixing with
fp32 accumulation); this module holds the parameters (device tensors), builds
the (residual, correction) training pairs from the library's own solves and
wraps the trained network as the preconditioner of ``FlexibleBiCGStab``.

Preconditioner:  M r = r + s P(FNO(R r / s, chi)),  s = |r| / sqrt(n), where R
averages factor^3 fine cells into one coarse cell and P interpolates back
trilinearly.  The
network is trained to minimise the preconditioned residual
|A M r - r|^2 / |r|^2 over residuals of the library's own solves
(``em3d_amd.learn``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass(frozen=True)
class FNOConfig:
    """grid: coarse cells per axis; channels: hidden width; modes: Fourier
    modes kept per axis (k in [-modes/2, modes/2)); layers: Fourier layers."""
    grid: int = 32
    channels: int = 32
    modes: int = 16
    layers: int = 4


def _cfg(c: FNOConfig, n_in: int, n_out: int, n_mat: int):
    return _lib.FnoConfig(c.grid, c.channels, c.modes, c.layers, n_in, n_out, n_mat)


class FNO:
    """Parameters (fp32 master + bf16 copy), gradients and Adam state on the device.

    The network is linear in its ``in_channels`` inputs (the residual) and
    conditioned on ``material_channels`` inputs through multiplicative gates."""

    def __init__(self, config: FNOConfig, in_channels: int, out_channels: int, *, material_channels: int = 0,
                 device, seed: int = 0):
        import torch
        self.config = config
        self.in_channels, self.out_channels = int(in_channels), int(out_channels)
        self.material_channels = int(material_channels)
        self.device = device
        self.cstruct = _cfg(config, self.in_channels, self.out_channels, self.material_channels)
        lib = _lib.lib()
        n = int(lib.em3d_fno_param_count(C.byref(self.cstruct)))
        if n <= 0:
            raise ValueError(f"invalid FNO configuration {config} (in {in_channels}, out {out_channels})")
        self.n_params = n
        self.params = torch.as_tensor(self._init(seed)).to(device)
        self.params_bf16 = torch.empty(n, dtype=torch.int16, device=device)
        self.grad = torch.zeros(n, dtype=torch.float32, device=device)
        self.m = torch.zeros_like(self.grad)
        self.v = torch.zeros_like(self.grad)
        self.step = 0
        self._work = {}
        self.refresh_bf16()

    # parameter layout of e3_fno.hip: W_in, per layer (W, U, d, R), W_out
    def _sizes(self):
        c, H = self.config, self.config.channels
        m3 = c.modes ** 3
        out = [("w_in", H * self.in_channels)]
        for l in range(c.layers):
            out += [(f"w{l}", H * H), (f"u{l}", H * self.material_channels), (f"d{l}", H),
                    (f"r{l}", m3 * H * H * 2)]
        out += [("w_out", self.out_channels * H)]
        return out

    def _init(self, seed: int) -> np.ndarray:
        rng = np.random.default_rng(seed)
        H = self.config.channels
        parts = []
        for name, size in self._sizes():
            if name == "w_in":
                parts.append(rng.standard_normal(size) / np.sqrt(self.in_channels))
            elif name.startswith("u"):
                parts.append(rng.standard_normal(size) / np.sqrt(max(1, self.material_channels)))
            elif name.startswith("w") and name != "w_out":
                parts.append(rng.standard_normal(size) / np.sqrt(H))
            elif name.startswith("r"):
                parts.append(rng.standard_normal(size) * (0.5 / H))
            else:   # gate biases, and W_out = 0: the preconditioner starts as the identity
                parts.append(np.zeros(size))
        p = np.concatenate(parts).astype(np.float32)
        assert p.size == self.n_params
        return p

    def refresh_bf16(self):
        _lib.check(_lib.lib().em3d_fno_to_bf16(self.n_params, self.params.data_ptr(), self.params_bf16.data_ptr(),
                                               _lib.stream_handle(self.device)), "em3d_fno_to_bf16")

    def work(self, samples: int):
        import torch
        w = self._work.get(samples)
        if w is None:
            lib = _lib.lib()
            nb = int(lib.em3d_fno_work_bytes(C.byref(self.cstruct), samples))
            w = torch.empty(nb, dtype=torch.uint8, device=self.device)
            _lib.check(lib.em3d_fno_init_work(C.byref(self.cstruct), samples, w.data_ptr(),
                                              _lib.stream_handle(self.device)), "em3d_fno_init_work")
            self._work[samples] = w
        return w

    @property
    def cells(self) -> int:
        return self.config.grid ** 3

    def forward(self, x, samples: int):
        """y[s][cell][out] for x[s][cell][in] (fp32, contiguous)."""
        import torch
        y = torch.empty((samples, self.cells, self.out_channels), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().em3d_fno_forward(C.byref(self.cstruct), self.params.data_ptr(),
                                               self.params_bf16.data_ptr(), x.data_ptr(), y.data_ptr(),
                                               self.work(samples).data_ptr(), samples,
                                               _lib.stream_handle(self.device)), "em3d_fno_forward")
        return y

    def backward(self, x, dy, samples: int):
        """grad += dL/dparams (after forward() of the same x)."""
        _lib.check(_lib.lib().em3d_fno_backward(C.byref(self.cstruct), self.params_bf16.data_ptr(), x.data_ptr(),
                                                dy.data_ptr(), self.grad.data_ptr(), self.work(samples).data_ptr(),
                                                samples, _lib.stream_handle(self.device)), "em3d_fno_backward")

    def loss_and_grad(self, y, target, samples: int):
        """(device loss = mean over samples of |y - t|^2 / |t|^2, dL/dy)."""
        import torch
        lib = _lib.lib()
        dy = torch.empty_like(y)
        loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        lw = torch.empty(int(lib.em3d_fno_loss_work_bytes(samples)), dtype=torch.uint8, device=self.device)
        _lib.check(lib.em3d_fno_loss(samples, y[0].numel(), y.data_ptr(), target.data_ptr(), dy.data_ptr(),
                                     loss.data_ptr(), lw.data_ptr(), _lib.stream_handle(self.device)),
                   "em3d_fno_loss")
        return loss / samples, dy

    def adam_step(self, *, lr: float, weight_decay: float = 0.0):
        """Apply the accumulated gradient (Adam) and refresh the bf16 copy."""
        self.step += 1
        _lib.check(_lib.lib().em3d_fno_adam(self.n_params, self.params.data_ptr(), self.params_bf16.data_ptr(),
                                            self.grad.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), lr, 0.9,
                                            0.999, 1e-8, weight_decay, self.step, _lib.stream_handle(self.device)),
                   "em3d_fno_adam")

    def train_step(self, x, target, *, lr: float = 1e-3, weight_decay: float = 0.0):
        """One Adam step on a minibatch (target loss); returns the device loss before the step."""
        samples = x.shape[0]
        self.grad.zero_()
        y = self.forward(x, samples)
        loss, dy = self.loss_and_grad(y, target, samples)
        self.backward(x, dy, samples)
        self.adam_step(lr=lr, weight_decay=weight_decay)
        return loss


class FNOPreconditioner:
    """The trained network attached to one operator (grid and contrast)."""

    def __init__(self, fno: FNO, operator):
        import torch
        N = operator._N
        c = fno.config
        if len(set(N)) != 1 or N[0] % c.grid:
            raise ValueError(f"the FNO preconditioner needs a cubic grid divisible by {c.grid}, got {N}")
        self.fno = fno
        self.operator = operator
        self.factor = N[0] // c.grid
        self.ncomp = operator._ncomp
        # material channels: the diagonal of chi (re, im), plus k0 / 10 when the network takes it
        self.chi_channels = fno.material_channels
        if (self.chi_channels not in ((6, 7) if self.ncomp == 3 else (2, 3)) or fno.in_channels != self.ncomp
                or fno.out_channels != self.ncomp):
            raise ValueError("FNO channels do not match the operator (in = out = ncomp, material = chi [+ k0])")
        be = operator.backend
        dev = be.torch_device
        _, kind, chi = operator._spectrum()
        self.chi_features = torch.empty((fno.cells, self.chi_channels), dtype=torch.float32, device=dev)
        _lib.check(_lib.lib().em3d_fno_chi_features(*N, self.factor, be.precision.code,
                                                    chi.data_ptr() if chi is not None else None, kind,
                                                    self.chi_channels, float(operator._k0) / 10.0,
                                                    self.chi_features.data_ptr(), _lib.stream_handle(dev)),
                   "em3d_fno_chi_features")
        self._models = {}

    def model(self, batch: int):
        """ctypes em3d_fno_model with work and feature buffers for `batch` RHS
        (two real network samples each: real and imaginary parts)."""
        import torch
        m = self._models.get(batch)
        if m is None:
            f = self.fno
            feats = torch.empty(int(_lib.lib().em3d_fno_features_bytes(C.byref(f.cstruct), 2 * batch)),
                                dtype=torch.uint8, device=f.device)
            w = f.work(2 * batch)
            st = _lib.FnoModel(f.cstruct, self.factor, f.params.data_ptr(), f.params_bf16.data_ptr(), w.data_ptr(),
                               feats.data_ptr(), self.chi_features.data_ptr(), self.chi_channels)
            m = (st, feats)
            self._models[batch] = m
        return m[0]

    def apply(self, r):
        """M r for a field (or a batch of fields) of the operator's shape."""
        import torch
        t, batch = self.operator._prepare_input(r)
        out = torch.empty_like(t)
        N = self.operator._N
        _lib.check(_lib.lib().em3d_fno_precondition(C.byref(self.model(batch)), batch, self.ncomp, *N,
                                                    self.operator.backend.precision.code, t.data_ptr(),
                                                    out.data_ptr(), _lib.stream_handle(t.device)),
                   "em3d_fno_precondition")
        return out

    def features(self, r):
        """(features [2 batch][cells][in + material], scale s [batch]) of
        residuals r: sample 2b holds the real parts of r_b, 2b + 1 the
        imaginary parts."""
        import torch
        lib = _lib.lib()
        op = self.operator
        rb, batch = op._prepare_input(r)
        dev = rb.device
        n = rb.numel() // batch
        scale = torch.empty(batch, dtype=torch.float64, device=dev)
        sw = torch.empty(int(lib.em3d_fno_scale_work_bytes(batch)), dtype=torch.uint8, device=dev)
        st = _lib.stream_handle(dev)
        prec = op.backend.precision.code
        _lib.check(lib.em3d_fno_scale(batch, n, prec, rb.data_ptr(), scale.data_ptr(), sw.data_ptr(), st),
                   "em3d_fno_scale")
        x = torch.empty((2 * batch, self.fno.cells, self.ncomp + self.chi_channels), dtype=torch.float32,
                        device=dev)
        _lib.check(lib.em3d_fno_restrict(batch, self.ncomp, *op._N, self.factor, prec, rb.data_ptr(),
                                         scale.data_ptr(), self.chi_features.data_ptr(), self.chi_channels,
                                         x.data_ptr(), st), "em3d_fno_restrict")
        return x, scale

    def prolong(self, y, scale, base=None):
        """base + scale P(y_re + i y_im) on the operator's grid (y: [2 batch][cells][ncomp])."""
        import torch
        op = self.operator
        batch = y.shape[0] // 2
        shape = ((batch,) if batch > 1 or base is None else ()) + op._field_shape()
        out = torch.empty(shape if base is None else base.shape, dtype=op.backend.precision.torch_complex_dtype,
                          device=y.device)
        _lib.check(_lib.lib().em3d_fno_prolong(batch, self.ncomp, *op._N, self.factor, op.backend.precision.code,
                                               base.data_ptr() if base is not None else None, scale.data_ptr(),
                                               y.data_ptr(), out.data_ptr(), _lib.stream_handle(y.device)),
                   "em3d_fno_prolong")
        return out

    def prolong_adjoint(self, g, scale):
        """[2 batch][cells][ncomp]: P^T of the real / imaginary parts of g, / scale."""
        import torch
        op = self.operator
        gb, batch = op._prepare_input(g)
        out = torch.empty((2 * batch, self.fno.cells, self.ncomp), dtype=torch.float32, device=gb.device)
        _lib.check(_lib.lib().em3d_fno_prolong_adjoint(batch, self.ncomp, *op._N, self.factor,
                                                       op.backend.precision.code, gb.data_ptr(), scale.data_ptr(),
                                                       out.data_ptr(), _lib.stream_handle(gb.device)),
                   "em3d_fno_prolong_adjoint")
        return out

    # ---- training pairs -----------------------------------------------------
    def restrict_pair(self, r, e):
        """(features of r, target R(e - r)/s) for residuals r and corrections e
        (A e = r), both (batch,) + field shape; returns fp32 [2 batch][cells][.]
        (real and imaginary samples)."""
        import torch
        lib = _lib.lib()
        op = self.operator
        rb, batch = op._prepare_input(r)
        eb, _ = op._prepare_input(e)
        if rb.dim() == len(op._field_shape()):
            rb, eb = rb.unsqueeze(0), eb.unsqueeze(0)
        dev = rb.device
        N = op._N
        prec = op.backend.precision.code
        n = rb[0].numel()
        scale = torch.empty(batch, dtype=torch.float64, device=dev)
        sw = torch.empty(int(lib.em3d_fno_scale_work_bytes(batch)), dtype=torch.uint8, device=dev)
        s = _lib.stream_handle(dev)
        _lib.check(lib.em3d_fno_scale(batch, n, prec, rb.data_ptr(), scale.data_ptr(), sw.data_ptr(), s),
                   "em3d_fno_scale")
        cells = self.fno.cells
        x = torch.empty((2 * batch, cells, self.ncomp + self.chi_channels), dtype=torch.float32, device=dev)
        _lib.check(lib.em3d_fno_restrict(batch, self.ncomp, *N, self.factor, prec, rb.data_ptr(), scale.data_ptr(),
                                         self.chi_features.data_ptr(), self.chi_channels, x.data_ptr(), s),
                   "em3d_fno_restrict")
        from .solvers._vec import lincomb
        d = lincomb(1.0, eb, -1.0, rb)
        t = torch.empty((2 * batch, cells, self.ncomp), dtype=torch.float32, device=dev)
        _lib.check(lib.em3d_fno_restrict(batch, self.ncomp, *N, self.factor, prec, d.data_ptr(), scale.data_ptr(),
                                         None, 0, t.data_ptr(), s), "em3d_fno_restrict")
        return x, t
