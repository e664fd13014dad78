"""Oracle restatement of the reference's volume-integral operators.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Plain NumPy, host only.
``precision="double"`` is complex128 throughout.  ``precision="single"``
follows the reference's SINGLE semantics step by step: scalar factors rounded
to complex64 where the reference casts, NumPy promotion to complex128 where the
reference mixes a complex64 scalar with float64 arrays (NEP 50), complex64 FFTs.
"""
from __future__ import annotations

import numpy as np

from .inputs import CPLX, axis, cell_volume

_ROW_SHAPES = ((-1, 1, 1), (1, -1, 1), (1, 1, -1))


# ---------------------------------------------------------------------------
# EM dyadic kernel
# ---------------------------------------------------------------------------

def b_coeff(x, y, *, k: float, dv: float) -> np.ndarray:
    """3x3 dyadic block ``dv G(R) [C1 a a^T + C2 I]``, self block -I/3
    (src/em3d/kernel.py:22-45, green function src/em3d/kernel.py:7-19)."""
    d = np.asarray(x, np.float64) - np.asarray(y, np.float64)
    r = float(np.linalg.norm(d))
    ident = np.eye(3, dtype=np.complex128)
    if r < 1e-15:
        return (-1.0 / 3.0) * ident
    a = d / r
    c1 = 3.0 / (r * r) - 3.0j * k / r - k * k
    c2 = k * k + 1.0j * k / r - 1.0 / (r * r)
    g = np.exp(1j * k * r) / (4.0 * np.pi * r)
    return g * dv * (c1 * np.outer(a, a) + c2 * ident)


def doubled_offsets(N, L):
    """Offsets ``[0..N-1, -N..-1] * h`` per axis (src/em3d/operator.py:16-25)."""
    out = []
    for n, l in zip(N, L):
        h = l / n
        out.append(np.concatenate([np.arange(n) * h, -(np.arange(n, 0, -1)) * h]))
    return out


def kernel_tensor(N, L, k: float, precision: str = "double") -> np.ndarray:
    """Block-circulant spatial kernel (3, 3, 2Nx, 2Ny, 2Nz)
    (src/em3d/operator.py:28-107)."""
    c = CPLX[precision]
    sx, sy, sz = doubled_offsets(N, L)
    r2 = sx[:, None, None] ** 2 + sy[None, :, None] ** 2 + sz[None, None, :] ** 2
    radius = np.sqrt(r2)
    is_self = radius < 1e-15
    radius = np.where(is_self, np.ones_like(radius), radius)
    inv_r = 1.0 / radius
    ik = c(1j * k)
    dv = cell_volume(N, L)
    wg = (np.exp(ik * radius) / (4.0 * np.pi * radius) * dv).astype(c, copy=False)
    c1 = (3.0 * inv_r * inv_r - 3.0 * ik * inv_r - k * k).astype(c, copy=False)
    c2 = (k * k + ik * inv_r - inv_r * inv_r).astype(c, copy=False)
    offsets = (sx, sy, sz)
    out = np.empty((3, 3) + radius.shape, dtype=c)
    for a in range(3):
        for b in range(3):
            ar = offsets[a].reshape(_ROW_SHAPES[a]) * inv_r
            ac = offsets[b].reshape(_ROW_SHAPES[b]) * inv_r
            blk = wg * c1 * ar * ac
            if a == b:
                blk = blk + wg * c2
                blk = np.where(is_self, -1.0 / 3.0, blk)
            else:
                blk = np.where(is_self, 0.0, blk)
            out[a, b] = blk.astype(c, copy=False)
    return out


def kernel_hat(N, L, k: float, precision: str = "double", adjoint: bool = False) -> np.ndarray:
    """``FFT3(K)`` (or ``FFT3(conj K)``) per block (src/em3d/operator.py:196-211)."""
    K = kernel_tensor(N, L, k, precision)
    if adjoint:
        K = np.conj(K)
    return np.fft.fftn(K, axes=(-3, -2, -1)).astype(CPLX[precision], copy=False)


def reverse_frequency_axes(block: np.ndarray) -> np.ndarray:
    """``F(k) -> F(-k)`` in FFT index order (src/em3d/operator.py:141-148)."""
    return np.roll(np.flip(block, axis=(-3, -2, -1)), shift=(1, 1, 1), axis=(-3, -2, -1))


def _pad(field, N):
    out = np.zeros(field.shape[:-3] + tuple(2 * n for n in N), dtype=field.dtype)
    out[..., : N[0], : N[1], : N[2]] = field
    return out


def _apply_blocks(Khat, fhat):
    """``y_r = sum_c K_rc f_c`` (src/em3d/operator.py:294-302)."""
    out = np.zeros_like(fhat)
    for r in range(3):
        acc = None
        for c in range(3):
            t = Khat[r, c] * fhat[c]
            acc = t if acc is None else acc + t
        out[r] = acc
    return out


def em_matvec(chi, u, Khat) -> np.ndarray:
    """``A u = u - crop(IFFT(Khat . FFT(pad(chi u))))`` (src/em3d/operator.py:353-363)."""
    N = u.shape[-3:]
    cu = np.einsum("ab...,b...->a...", chi, u)
    fh = np.fft.fftn(_pad(cu, N), axes=(-3, -2, -1))
    big = np.fft.ifftn(_apply_blocks(Khat, fh), axes=(-3, -2, -1))
    return (u - big[:, : N[0], : N[1], : N[2]]).astype(u.dtype, copy=False)


def em_rmatvec(chi, v, Khat=None, Khat_adj=None) -> np.ndarray:
    """``A^H v = v - chi^H crop(IFFT(conjK . FFT(pad v)))`` with the explicit
    adjoint spectrum or, if only ``Khat`` is given, the derived one
    ``conj(Khat[-k])`` (src/em3d/operator.py:365-398, 305-317)."""
    N = v.shape[-3:]
    fh = np.fft.fftn(_pad(v, N), axes=(-3, -2, -1))
    if Khat_adj is None:
        Khat_adj = np.empty_like(Khat)
        for r in range(3):
            for c in range(3):
                Khat_adj[r, c] = np.conj(reverse_frequency_axes(Khat[r, c]))
    big = np.fft.ifftn(_apply_blocks(Khat_adj, fh), axes=(-3, -2, -1))
    kv = big[:, : N[0], : N[1], : N[2]]
    chiH = np.conj(chi).swapaxes(0, 1)
    return (v - np.einsum("ab...,b...->a...", chiH, kv)).astype(v.dtype, copy=False)


# ---------------------------------------------------------------------------
# Dense brute force (independent of the circulant embedding)
# ---------------------------------------------------------------------------

def cell_centres(N, L, center) -> np.ndarray:
    xs = [axis(int(N[a]), float(L[a]), float(center[a])) for a in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=-1)


def dense_kernel_matrix(N, L, center, k: float) -> np.ndarray:
    """Cell-by-cell ``b_coeff`` assembly, cell-major/component-minor
    (src/em3d/dense.py:100-118)."""
    cen = cell_centres(N, L, center)
    M = cen.shape[0]
    dv = cell_volume(N, L)
    out = np.zeros((3 * M, 3 * M), dtype=np.complex128)
    for i in range(M):
        for j in range(M):
            out[3 * i: 3 * i + 3, 3 * j: 3 * j + 3] = b_coeff(cen[i], cen[j], k=k, dv=dv)
    return out


def flatten_field(f) -> np.ndarray:
    """(3, Nx, Ny, Nz) -> (ix, iy, iz, comp) order (src/em3d/layout.py:14-26)."""
    return np.asarray(f).transpose(1, 2, 3, 0).reshape(-1)


def unflatten_field(v, N) -> np.ndarray:
    """Inverse of ``flatten_field`` (src/em3d/layout.py:29-39)."""
    return np.asarray(v).reshape(tuple(N) + (3,)).transpose(3, 0, 1, 2)


def dense_contrast_matrix(chi) -> np.ndarray:
    """Block-diagonal chi (src/em3d/dense.py:156-171)."""
    chi = np.asarray(chi, np.complex128)
    M = int(np.prod(chi.shape[2:]))
    blocks = chi.transpose(2, 3, 4, 0, 1).reshape(M, 3, 3)
    out = np.zeros((3 * M, 3 * M), dtype=np.complex128)
    idx = np.arange(M)
    for r in range(3):
        for c in range(3):
            out[3 * idx + r, 3 * idx + c] = blocks[:, r, c]
    return out


def dense_operator_matrix(N, L, center, k, chi) -> np.ndarray:
    """``A = I - B chi`` (src/em3d/dense.py:180-209)."""
    B = dense_kernel_matrix(N, L, center, k)
    return np.eye(B.shape[0], dtype=np.complex128) - B @ dense_contrast_matrix(chi)


# ---------------------------------------------------------------------------
# Acoustic scalar operator
# ---------------------------------------------------------------------------

def self_cell_coefficient(*, k0: float, dv: float) -> complex:
    """Equivalent-sphere ``exp(ika)(1 - ika) - 1``, ``a = (3 dv / 4 pi)^(1/3)``
    (src/em3d/acoustics/kernel.py:7-18)."""
    if k0 <= 0.0:
        raise ValueError(f"k0 must be positive, got {k0}")
    if dv <= 0.0:
        raise ValueError(f"dv must be positive, got {dv}")
    a = (3.0 * dv / (4.0 * np.pi)) ** (1.0 / 3.0)
    return complex(np.exp(1j * k0 * a) * (1.0 - 1j * k0 * a) - 1.0)


def acoustic_kernel(N, L, k0: float, precision: str = "double") -> np.ndarray:
    """Scalar ``k0^2 dv G(R)`` on the doubled grid with the self-cell value
    (src/em3d/acoustics/kernel.py:21-40)."""
    c = CPLX[precision]
    sx, sy, sz = doubled_offsets(N, L)
    SX, SY, SZ = np.meshgrid(sx, sy, sz, indexing="ij")
    R = np.sqrt(SX * SX + SY * SY + SZ * SZ)
    is_self = R < 1e-15
    Rs = np.where(is_self, np.ones_like(R), R)
    ik = c(1j * float(k0))
    dv = cell_volume(N, L)
    g = (float(k0) ** 2) * dv * np.exp(ik * Rs) / (4.0 * np.pi * Rs)
    sv = c(self_cell_coefficient(k0=float(k0), dv=dv))
    return np.where(is_self, sv, g).astype(c, copy=False)


def acoustic_kernel_hat(N, L, k0, precision="double", adjoint=False):
    """FFT of the scalar kernel or its conjugate (src/em3d/acoustics/kernel.py:43-58)."""
    K = acoustic_kernel(N, L, k0, precision)
    if adjoint:
        K = np.conj(K)
    return np.fft.fftn(K, axes=(-3, -2, -1)).astype(CPLX[precision], copy=False)


def acoustic_matvec(chi, u, Khat) -> np.ndarray:
    """``u - crop(IFFT(Khat FFT(pad(chi u))))`` (src/em3d/acoustics/operator.py:40-54)."""
    N = u.shape
    big = np.fft.ifftn(Khat * np.fft.fftn(_pad(chi * u, N), axes=(-3, -2, -1)), axes=(-3, -2, -1))
    return (u - big[: N[0], : N[1], : N[2]]).astype(u.dtype, copy=False)


def acoustic_rmatvec(chi, v, Khat_adj) -> np.ndarray:
    """``v - conj(chi) crop(IFFT(conjKhat FFT(pad v)))`` (src/em3d/acoustics/operator.py:56-67)."""
    N = v.shape
    big = np.fft.ifftn(Khat_adj * np.fft.fftn(_pad(v, N), axes=(-3, -2, -1)), axes=(-3, -2, -1))
    return (v - np.conj(chi) * big[: N[0], : N[1], : N[2]]).astype(v.dtype, copy=False)


def acoustic_dense_B(N, L, center, k0) -> np.ndarray:
    """Cell-pair scalar B (src/em3d/acoustics/dense.py:31-44)."""
    cen = cell_centres(N, L, center)
    M = cen.shape[0]
    dv = cell_volume(N, L)
    out = np.zeros((M, M), dtype=np.complex128)
    for i in range(M):
        for j in range(M):
            R = float(np.linalg.norm(cen[i] - cen[j]))
            if R < 1e-15:
                out[i, j] = self_cell_coefficient(k0=k0, dv=dv)
            else:
                out[i, j] = (k0 * k0) * dv * np.exp(1j * k0 * R) / (4.0 * np.pi * R)
    return out


def acoustic_dense_H(N, L, center, k0, chi) -> np.ndarray:
    """``H = I - B diag(chi)`` (src/em3d/acoustics/dense.py:47-51)."""
    B = acoustic_dense_B(N, L, center, k0)
    return np.eye(B.shape[0], dtype=np.complex128) - B @ np.diag(np.asarray(chi, np.complex128).reshape(-1))
