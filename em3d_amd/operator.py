"""FFT volume-integral operator on MI355X (src/em3d/operator.py of the reference).

``PreparedEMKernel`` and ``Operator`` keep the reference's constructor
signatures, validation rules and error types; the arithmetic runs in
libem3d_hip.so:

* the kernel spectrum is built on the device in FP64 (spatial dyadic blocks,
  3-D FFT) and held as the octant of its 6 unique symmetric components
  (the spectrum of an even/odd block is even/odd per axis once the index-N
  planes of the doubled grid, which the cropped convolution never reads, are
  zeroed);
* ``matvec``/``rmatvec`` run the five pruned FFT passes with the contrast
  product, the 3x3 spectral product and the ``u - .`` epilogue fused in.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from math import prod
from typing import Literal

import numpy as np

from . import _lib
from .dtypes import Precision
from .grid import Grid
from .layout import flatten_field, unflatten_field
from .problem import Problem

AdjointStorage = Literal["explicit", "derived", "none"]
KernelBuildStrategy = Literal["standard", "streamed"]

CHI_NONE, CHI_ISO, CHI_DIAG, CHI_FULL = 0, 1, 2, 3
_OCT_COMPONENTS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def require_hip(backend) -> None:
    if getattr(backend, "device", None) != "hip":
        raise TypeError("em3d_amd computes on Backend.hip() only (got device "
                        f"{getattr(backend, 'device', None)!r}); use the em3d package for NumPy/CuPy")


class HipPlan:
    """Owner of one ``em3d_plan_t`` (radix plans and twiddles of a grid shape)."""

    _cache: dict = {}

    def __init__(self, N, ncomp: int, batch: int, precision: Precision, device_index: int = 0):
        import torch
        self.N = tuple(int(n) for n in N)
        self.ncomp, self.batch, self.precision = int(ncomp), int(batch), precision
        self.device_index = int(device_index)
        lib = _lib.lib()
        h = C.c_void_p()
        with torch.cuda.device(self.device_index):
            _lib.check(lib.em3d_plan_create(*self.N, self.ncomp, self.batch, precision.code, C.byref(h)),
                       "em3d_plan_create")
        self.handle = h
        kb, wb, bb = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _lib.check(lib.em3d_plan_sizes(h, C.byref(kb), C.byref(wb), C.byref(bb)), "em3d_plan_sizes")
        self.khat_bytes, self.work_bytes, self.build_bytes = kb.value, wb.value, bb.value

    def variants(self) -> list[int]:
        v = (C.c_int * 3)()
        _lib.check(_lib.lib().em3d_plan_get_variants(self.handle, v), "em3d_plan_get_variants")
        return [int(x) for x in v]

    def set_variants(self, variants) -> None:
        v = (C.c_int * 3)(*[int(x) for x in variants])
        _lib.check(_lib.lib().em3d_plan_set_variants(self.handle, v), "em3d_plan_set_variants")

    def variant_counts(self) -> list[int]:
        return [int(_lib.lib().em3d_plan_variant_count(self.handle, k)) for k in range(3)]

    @classmethod
    def get(cls, N, ncomp, batch, precision, device_index=0) -> "HipPlan":
        key = (tuple(int(n) for n in N), int(ncomp), int(batch), precision, int(device_index))
        p = cls._cache.get(key)
        if p is None:
            p = cls(N, ncomp, batch, precision, device_index)
            cls._cache[key] = p
        return p

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().em3d_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


# FFT-pass scratch shared by the operators of one plan on one stream
_SHARED_WORK: dict = {}


def khat_storage_index(N, ncomp: int, elem_bytes: int):
    """Storage index of every logical octant entry (comp, ix, iy, iz) as an int64
    array of shape (K, Nx+1, Ny+1, Nz+1).  The library stores the octant as
    [iy][iz // S][ix][comp][iz % S] for 0 <= iz <= Nz (S = 64 bytes of
    entries, G = Nz // S + 1 groups per iy)."""
    K = 6 if ncomp == 3 else 1
    S = 64 // elem_bytes
    nx1, nz = N[0] + 1, N[2]
    G = nz // S + 1
    c = np.arange(K)[:, None, None, None]
    ix = np.arange(nx1)[None, :, None, None]
    iy = np.arange(N[1] + 1)[None, None, :, None]
    iz = np.arange(nz + 1)[None, None, None, :]
    return ((((iy * G + iz // S) * nx1 + ix) * K + c) * S + iz % S).astype(np.int64)


def octant_entries(khat_flat, N, ncomp: int):
    """Logical (K, Nx+1, Ny+1, Nz+1) view (a gathered copy) of a stored octant spectrum."""
    import torch
    idx = khat_storage_index(N, ncomp, khat_flat.element_size())
    return khat_flat[torch.as_tensor(idx, device=khat_flat.device)]


def build_octant_spectrum(N, L, k0: float, ncomp: int, precision: Precision, device_index: int = 0, h=None):
    """Device octant spectrum in the library's storage order (1-D), FP64 build,
    plan precision, pre-scaled by 1/(8 Nx Ny Nz).  ``h``: the cell sizes,
    when given (a sub-grid that must keep its parent's cell size bit for
    bit); else L / N."""
    import torch
    plan = HipPlan.get(N, ncomp, 1, precision, device_index)
    dev = torch.device("cuda", device_index)
    es = 8 if precision is Precision.SINGLE else 16
    khat = torch.zeros(plan.khat_bytes // es, dtype=precision.torch_complex_dtype, device=dev)
    scratch = torch.empty(plan.build_bytes, dtype=torch.uint8, device=dev)
    h = [L[a] / N[a] for a in range(3)] if h is None else [float(v) for v in h]
    _lib.check(_lib.lib().em3d_kernel_build(plan.handle, float(k0), h[0], h[1], h[2], khat.data_ptr(),
                                            scratch.data_ptr(), _lib.stream_handle(dev)), "em3d_kernel_build")
    del scratch
    return khat


def _fold(n: int, device):
    import torch
    Lf = 2 * n
    k = torch.arange(Lf, device=device)
    flipped = k > n
    return torch.where(flipped, Lf - k, k), flipped


def materialize_kernel_hat(khat_flat, N, adjoint: bool = False):
    """Full ``(3, 3, 2Nx, 2Ny, 2Nz)`` spectrum (unscaled, FFT index order) from the octant."""
    import torch
    oct_khat = octant_entries(khat_flat, N, 3)
    dev = oct_khat.device
    (ix, fx), (iy, fy), (iz, fz) = (_fold(N[a], dev) for a in range(3))
    D = 8 * prod(N)
    sx = torch.where(fx, -1.0, 1.0)[:, None, None]
    sy = torch.where(fy, -1.0, 1.0)[None, :, None]
    sz = torch.where(fz, -1.0, 1.0)[None, None, :]
    signs = {(0, 1): sx * sy, (0, 2): sx * sz, (1, 2): sy * sz}
    full = torch.empty((3, 3) + tuple(2 * n for n in N), dtype=oct_khat.dtype, device=dev)
    for c, (a, b) in enumerate(_OCT_COMPONENTS):
        blk = oct_khat[c][ix[:, None, None], iy[None, :, None], iz[None, None, :]] * D
        if a != b:
            blk = blk * signs[(a, b)].to(blk.real.dtype)
        if adjoint:
            blk = torch.conj(blk)
        full[a, b] = blk
        full[b, a] = blk
    return full


@dataclass(frozen=True)
class PreparedEMKernel:
    """Reusable spectral representation of the dyadic convolution kernel.

    Holds the octant spectrum actually read by the kernels.  The adjoint is
    always available from the same storage as ``conj(Khat(k))`` (the derived
    adjoint); ``adjoint_storage`` records what the caller asked for and
    ``include_adjoint=False`` disables ``rmatvec`` as in the reference.
    ``kernel_hat`` materialises the full (3, 3, 2N, 2N, 2N) tensor on request;
    it differs from the reference's spectrum only through the zeroed index-N
    planes, which never reach a cropped output.
    """

    grid_shape: tuple[int, int, int]
    grid_lengths: tuple[float, float, float]
    wave_number: float
    device: str
    precision: str
    khat_octant: object = field(repr=False)
    adjoint_storage: AdjointStorage = "explicit"
    build_strategy: KernelBuildStrategy = "standard"
    device_index: int = 0

    @classmethod
    def build(cls, grid: Grid, *, k: float, include_adjoint: bool = True,
              adjoint_storage: Literal["explicit", "derived"] = "explicit",
              build_strategy: KernelBuildStrategy = "standard") -> "PreparedEMKernel":
        if build_strategy not in {"standard", "streamed"}:
            raise ValueError(f"build_strategy must be 'standard' or 'streamed', got {build_strategy!r}")
        if adjoint_storage not in {"explicit", "derived"}:
            raise ValueError(f"adjoint_storage must be 'explicit' or 'derived', got {adjoint_storage!r}")
        be = grid.backend
        require_hip(be)
        storage: AdjointStorage = adjoint_storage if include_adjoint else "none"
        khat = build_octant_spectrum(grid.N, grid.L, float(k), 3, be.precision, be.device_index)
        return cls(grid_shape=tuple(int(v) for v in grid.N), grid_lengths=tuple(float(v) for v in grid.L),
                   wave_number=float(k), device=str(be.device), precision=str(be.precision.value),
                   khat_octant=khat, adjoint_storage=storage, build_strategy=build_strategy,
                   device_index=be.device_index)

    @property
    def supports_adjoint(self) -> bool:
        return self.adjoint_storage != "none"

    @property
    def persistent_nbytes(self) -> int:
        """Bytes this object keeps in device memory: the octant spectrum.  The
        reference's figure is the dense ``kernel_hat`` (plus its explicit
        adjoint) it stores (operator.py:230-235); that size is
        ``dense_kernel_hat_nbytes`` here, materialised only on request."""
        t = self.khat_octant
        return int(t.numel() * t.element_size())

    @property
    def dense_kernel_hat_nbytes(self) -> int:
        """Size of the (3, 3, 2Nx, 2Ny, 2Nz) tensor ``kernel_hat`` returns."""
        return 9 * 8 * prod(self.grid_shape) * int(self.khat_octant.element_size())

    @property
    def kernel_hat(self):
        return materialize_kernel_hat(self.khat_octant, self.grid_shape)

    @property
    def kernel_hat_adjoint(self):
        if self.adjoint_storage != "explicit":
            return None
        return materialize_kernel_hat(self.khat_octant, self.grid_shape, adjoint=True)

    def validate(self, grid: Grid, *, k: float) -> None:
        errors: list[str] = []
        if tuple(grid.N) != self.grid_shape:
            errors.append(f"grid shape {grid.N!r} != {self.grid_shape!r}")
        if not np.allclose(grid.L, self.grid_lengths, rtol=0.0, atol=0.0):
            errors.append(f"grid lengths {grid.L!r} != {self.grid_lengths!r}")
        if not np.isclose(float(k), self.wave_number, rtol=1e-14, atol=1e-14):
            errors.append(f"wave number {k!r} != {self.wave_number!r}")
        if grid.backend.device != self.device:
            errors.append(f"device {grid.backend.device!r} != {self.device!r}")
        if grid.backend.precision.value != self.precision:
            errors.append(f"precision {grid.backend.precision.value!r} != {self.precision!r}")
        if errors:
            raise ValueError("incompatible PreparedEMKernel: " + "; ".join(errors))


def prep_coeffs_em3d(grid: Grid, *, k: float, volume: float | None = None):
    """Octant spectrum (the storage the HIP kernels read)."""
    return PreparedEMKernel.build(grid, k=k, include_adjoint=False).khat_octant


def pack_contrast(chi):
    """Classify a (3, 3, Nx, Ny, Nz) device contrast as iso / diag / full and
    pack it.  The off-diagonal test reads the six slices in place (no copy of
    the tensor); an isotropic contrast packs to a fresh (Nx, Ny, Nz) tensor,
    so nothing keeps the 3 x 3 input alive."""
    import torch
    for a in range(3):
        for b in range(3):
            if a != b and bool(torch.count_nonzero(chi[a, b]).item()):
                return CHI_FULL, chi.contiguous()
    if bool(torch.equal(chi[0, 0], chi[1, 1])) and bool(torch.equal(chi[0, 0], chi[2, 2])):
        return CHI_ISO, chi[0, 0].clone(memory_format=torch.contiguous_format)
    return CHI_DIAG, torch.stack([chi[a, a] for a in range(3)])


def eta_contrast(eta):
    """Acoustic contrast ``chi = eta - 1`` (acoustics/problem.py:73-79) packed
    as an isotropic contrast."""
    return CHI_ISO, (eta - 1.0).contiguous()


class _FieldApplier:
    """Shared matvec/rmatvec plumbing: shape checks, plans and workspaces."""

    def _init_applier(self, be, N, ncomp, L=None, k0=None):
        self._be = be
        self._N = tuple(int(n) for n in N)
        self._ncomp = ncomp
        self._L = None if L is None else tuple(float(v) for v in L)
        self._k0 = None if k0 is None else float(k0)
        self._work: dict[int, object] = {}
        self._companion = None
        self._chi_src = None
        self._chi_version = None

    def _refresh_contrast(self, source, make):
        """(kind, packed contrast) of ``source``, re-packed when ``source`` is
        another tensor or its version counter moved (torch bumps it on every
        in-place change): the reference reads the problem's contrast on every
        call (operator.py:353-357), so an in-place edit must reach the next
        matvec, adjoint and solve.  The complex64 companion, which holds a
        cast of the old contrast, is dropped then."""
        from .farfield import packed_contrast
        v = getattr(source, "_version", None)
        if source is not self._chi_src or v != self._chi_version:
            self._chi_kind, self._chi = packed_contrast(source, make)
            self._chi_src, self._chi_version = source, v
            self._companion = None
        return self._chi_kind, self._chi

    def _spectrum(self):
        """(octant spectrum, chi kind, packed chi) this operator applies."""
        raise NotImplementedError

    def single_precision_companion(self) -> "_SingleCompanion":
        """The same operator in complex64: spectrum rebuilt in FP64 and cast,
        contrast cast from this operator's.  Built once and cached."""
        import torch
        _, kind, chi = self._spectrum()      # drops a companion built on a contrast edited since
        if self._companion is None:
            if self._be.precision is Precision.SINGLE:
                self._companion = self
            else:
                from .backend import Backend
                be32 = Backend.hip(Precision.SINGLE, self._be.device_index)
                khat = build_octant_spectrum(self._N, self._L, self._k0, self._ncomp, Precision.SINGLE,
                                             self._be.device_index, h=getattr(self, "_h", None))
                chi32 = None if chi is None else chi.to(torch.complex64).contiguous()
                self._companion = _SingleCompanion(be32, self._N, self._ncomp, self._L, self._k0, khat, kind, chi32)
        return self._companion

    def _field_shape(self):
        return ((3,) if self._ncomp == 3 else ()) + self._N

    def _workspace(self, plan, private: bool = False):
        """Scratch of the FFT passes for `plan`.  Operators of one shape share
        one workspace per stream (their matvecs run one after another in
        stream order), so many same-shape materials cost one workspace;
        ``private`` gives this operator its own (packed groups, whose members
        run as concurrent graph branches)."""
        import torch
        if private:
            w = self._work.get(plan.batch)
            if w is None:
                w = torch.empty(plan.work_bytes, dtype=torch.uint8, device=self._be.torch_device)
                self._work[plan.batch] = w
            return w
        key = (id(plan), self._be.device_index, _lib.stream_handle(self._be.torch_device))   # plans are cached for the process
        w = _SHARED_WORK.get(key)
        if w is None:
            w = torch.empty(plan.work_bytes, dtype=torch.uint8, device=self._be.torch_device)
            _SHARED_WORK[key] = w
        return w

    def _prepare_input(self, field):
        import torch
        be = self._be
        t = field if isinstance(field, torch.Tensor) else be.array(field)
        if t.device != be.torch_device:
            t = t.to(be.torch_device)
        if t.dtype != be.precision.torch_complex_dtype:
            t = t.to(be.precision.torch_complex_dtype)
        fs = self._field_shape()
        if tuple(t.shape) == fs:
            batch = 1
        elif t.dim() == len(fs) + 1 and tuple(t.shape[1:]) == fs:
            batch = int(t.shape[0])
        else:
            raise ValueError(f"field shape {tuple(t.shape)} != expected {fs} (or (batch,) + {fs})")
        return t.contiguous(), batch

    def _device_bicgstab(self, rhs, cfg, khat, chi_kind, chi, batched: bool = False, precond=None,
                         history: bool = True):
        """Device-resident BiCGStab of libem3d_hip on this operator.  ``batched``:
        rhs carries a leading batch axis and a list of results (one per RHS) is
        returned; every RHS runs its own recurrence and status.  ``precond``: an
        FNOPreconditioner attached to this operator (right preconditioning).
        ``history=False`` skips reading the residual history back (one device
        synchronisation fewer; the result then holds the last entry only)."""
        import torch

        from .solvers.base import SolverResult
        t, batch = self._prepare_input(rhs)
        if batched != (t.dim() == len(self._field_shape()) + 1):
            raise ValueError("batched solves take (batch,) + field shape; single solves take the field shape")
        be = self._be
        plan = HipPlan.get(self._N, self._ncomp, batch, be.precision, be.device_index)
        lib = _lib.lib()
        if precond is not None and precond.operator is not self:
            raise ValueError("the preconditioner is attached to another operator")
        wb = lib.em3d_bicgstab_pc_work_bytes if precond is not None else lib.em3d_bicgstab_work_bytes
        sw = torch.empty(int(wb(plan.handle)), dtype=torch.uint8, device=be.torch_device)
        cap = int(cfg.max_iter) + 2
        # history and action counts in one buffer (entries past history_len are never read)
        hbuf = torch.empty((2, batch, cap), dtype=torch.float64, device=be.torch_device)
        hist, acts = hbuf[0], hbuf[1].view(torch.int64)
        u = torch.empty_like(t)
        prm = _lib.BiCGStabParams(int(cfg.max_iter), float(cfg.rtol), float(cfg.atol),
                                  -1.0 if cfg.divergence_guard is None else float(cfg.divergence_guard),
                                  -1 if cfg.max_operator_actions is None else int(cfg.max_operator_actions),
                                  int(getattr(cfg, "chunk", 4) or 4), 1 if getattr(cfg, "fused", True) else 0)
        res = (_lib.SolverResultC * batch)()
        args = (plan.handle, khat.data_ptr(), chi.data_ptr() if chi is not None else None, chi_kind, t.data_ptr(),
                u.data_ptr(), self._workspace(plan).data_ptr(), sw.data_ptr(), C.byref(prm), hist.data_ptr(),
                acts.data_ptr(), cap, res)
        if precond is None:
            _lib.check(lib.em3d_bicgstab(*args, _lib.stream_handle(be.torch_device)), "em3d_bicgstab")
        else:
            _lib.check(lib.em3d_bicgstab_pc(*args, C.byref(precond.model(batch)),
                                            _lib.stream_handle(be.torch_device)), "em3d_bicgstab_pc")
        n_max = max(int(res[b].history_len) for b in range(batch))
        if history:
            # one plain copy of a small buffer; a slice of the last axis would add a gather kernel
            hb = hbuf.cpu() if hbuf.numel() <= (1 << 19) else hbuf[:, :, :max(n_max, 1)].cpu()
            hist_h, acts_h = hb[0], hb[1].contiguous().view(torch.int64)
        out = []
        for b in range(batch):
            n = int(res[b].history_len)
            if history:
                hist = [float(x) for x in hist_h[b, :n].tolist()]
                counts = [int(x) for x in acts_h[b, :n].tolist()]
            else:
                hist, counts = [float(res[b].true_final_residual)], [int(res[b].matvec_count)]
            status = _lib.STATUS[int(res[b].status)]
            out.append(SolverResult(u=u[b] if batched else u, iterations=int(res[b].iterations),
                                    residual_history=hist, converged=status == "converged",
                                    matvec_count=int(res[b].matvec_count), rmatvec_count=0,
                                    residual_action_counts=counts, status=status,
                                    true_final_residual=float(hist[-1]) if hist else float("inf")))
        return out if batched else out[0]

    def autotune(self, reps: int = 5, batch: int = 1):
        """Pick the fastest FFT kernel variant of each pass kind for this grid
        shape, precision and batch by timing whole matvecs (libem3d_hip's
        em3d_plan_autotune); the choice is kept by the shared plan.  Returns
        {"variants": [strided, x, z], "ms": time per matvec}."""
        import torch
        be = self._be
        plan = HipPlan.get(self._N, self._ncomp, batch, be.precision, be.device_index)
        khat, kind, chi = self._spectrum()
        shape = ((batch,) if batch > 1 else ()) + self._field_shape()
        g = torch.Generator(device=be.torch_device).manual_seed(0)
        u = torch.randn(shape, dtype=be.precision.torch_complex_dtype, device=be.torch_device, generator=g)
        y = torch.empty_like(u)
        best = C.c_double()
        lib = _lib.lib()
        _lib.check(lib.em3d_plan_autotune(plan.handle, khat.data_ptr(), chi.data_ptr() if chi is not None else None,
                                          kind, u.data_ptr(), y.data_ptr(), self._workspace(plan).data_ptr(),
                                          int(reps), C.byref(best), _lib.stream_handle(be.torch_device)),
                   "em3d_plan_autotune")
        return {"variants": plan.variants(), "ms": best.value}

    def _apply(self, field, khat, chi_kind, chi, adjoint: bool, window=None):
        import torch
        t, batch = self._prepare_input(field)
        plan = HipPlan.get(self._N, self._ncomp, batch, self._be.precision, self._be.device_index)
        y = torch.empty_like(t)
        args = (plan.handle, khat.data_ptr(), chi.data_ptr() if chi is not None else None, chi_kind,
                t.data_ptr(), y.data_ptr(), self._workspace(plan).data_ptr())
        stream = _lib.stream_handle(self._be.torch_device)
        if window is not None:
            lo, n = (C.c_int * 3)(*[int(v) for v in window[0]]), (C.c_int * 3)(*[int(v) for v in window[1]])
            _lib.check(_lib.lib().em3d_matvec_window(*args, lo, n, stream), "em3d_matvec_window")
            return y
        fn = _lib.lib().em3d_rmatvec if adjoint else _lib.lib().em3d_matvec
        _lib.check(fn(*args, stream), "em3d_rmatvec" if adjoint else "em3d_matvec")
        return y

    def matvec_window(self, field, lo, n):
        """``matvec`` with the forward transforms restricted to the box of cells
        [lo, lo + n) (em3d_matvec_window; no counterpart in the reference): chi u
        is taken as zero outside the box, so the result equals ``matvec``
        whenever the box holds the contrast's support (or u's)."""
        return self._apply(field, *self._spectrum(), False, window=(lo, n))


class _SingleCompanion(_FieldApplier):
    """complex64 twin of an FP64 operator (inner solves of the mixed-precision refinement)."""

    def __init__(self, be, N, ncomp, L, k0, khat, chi_kind, chi):
        self._init_applier(be, N, ncomp, L, k0)
        self._khat, self._chi_kind, self._chi = khat, chi_kind, chi

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


class Operator(_FieldApplier):
    """``A u = u - B(chi u)`` on the HIP backend (operator.py:320-428)."""

    def __init__(self, problem: Problem, *, prepared_kernel: PreparedEMKernel | None = None):
        self.problem = problem
        grid = problem.grid
        require_hip(grid.backend)
        if prepared_kernel is None:
            prepared_kernel = PreparedEMKernel.build(grid, k=problem.k0)
        else:
            prepared_kernel.validate(grid, k=problem.k0)
        self._prepared_kernel = prepared_kernel
        self._init_applier(grid.backend, grid.N, 3, grid.L, problem.k0)
        self._refresh_contrast(problem.eps_tensor, pack_contrast)

    def _spectrum(self):
        kind, chi = self._refresh_contrast(self.problem.eps_tensor, pack_contrast)
        return self._prepared_kernel.khat_octant, kind, chi

    @property
    def backend(self):
        return self._be

    @property
    def shape(self) -> tuple[int, int]:
        d = 3 * prod(self._N)
        return d, d

    @property
    def prepared_kernel(self) -> PreparedEMKernel:
        return self._prepared_kernel

    @property
    def contrast_kind(self) -> str:
        return {CHI_ISO: "isotropic", CHI_DIAG: "diagonal", CHI_FULL: "full"}[self._spectrum()[1]]

    def matvec(self, field):
        return self._apply(field, *self._spectrum(), False)

    def device_bicgstab(self, rhs, cfg, batched: bool = False, precond=None):
        """BiCGStab with all iteration scalars on the device (libem3d_hip)."""
        return self._device_bicgstab(rhs, cfg, *self._spectrum(), batched, precond)

    def rmatvec(self, field):
        if not self._prepared_kernel.supports_adjoint:
            raise RuntimeError("adjoint action is unavailable: build PreparedEMKernel with include_adjoint=True")
        return self._apply(field, *self._spectrum(), True)

    def matvec_flat(self, vector):
        return flatten_field(self.matvec(unflatten_field(vector, self._N)))

    def rmatvec_flat(self, vector):
        return flatten_field(self.rmatvec(unflatten_field(vector, self._N)))

    def iteration_matvec_flat(self, vector, *, mu: complex):
        if abs(mu) == 0.0:
            raise ValueError("mu must be non-zero")
        return vector - self.matvec_flat(vector) / complex(mu)

    def to_dense_kernel(self):
        raise RuntimeError("dense matrix construction requires a NumPy backend")

    def to_dense_operator(self):
        raise RuntimeError("dense matrix construction requires a NumPy backend")

    def to_dense(self):
        return self.to_dense_kernel()
