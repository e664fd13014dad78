"""Pins the oracle's operator restatement independently of its own formulas.

* dyadic kernel vs a finite-difference Hessian of g = exp(ikR)/(4 pi R)
  (K_ab = dv (k^2 delta_ab g + d_a d_b g)), self term -I/3, symmetry;
* FFT matvec / rmatvec vs brute-force dense A = I - B chi (cell pairs);
* adjoint identity, zero-contrast identity, derived vs explicit adjoint;
* acoustic FFT operator vs dense H, adjoint identity, self-cell limit.

A planted mistake (a flipped sign in C1/C2, a transposed chi, a wrong offset
order) fails at least one of these; see test_planted_mistakes_are_caught.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import inputs, vie


def _green(p, k):
    r = np.linalg.norm(p)
    return np.exp(1j * k * r) / (4 * np.pi * r)


def _fd_hessian(p, k, h=1e-4):
    H = np.zeros((3, 3), dtype=np.complex128)
    E = np.eye(3) * h
    for a in range(3):
        for b in range(3):
            H[a, b] = (_green(p + E[a] + E[b], k) - _green(p + E[a] - E[b], k)
                       - _green(p - E[a] + E[b], k) + _green(p - E[a] - E[b], k)) / (4 * h * h)
    return H


@pytest.mark.parametrize("p,k", [((0.2, 0.3, 0.0), 1.0), ((0.31, -0.17, 0.22), 4.0), ((0.05, 0.4, -0.3), 12.0)])
def test_dyadic_block_matches_finite_difference_hessian(p, k):
    p = np.asarray(p)
    dv = 0.01
    expected = dv * (k * k * np.eye(3) * _green(p, k) + _fd_hessian(p, k))
    got = vie.b_coeff(p, np.zeros(3), k=k, dv=dv)
    np.testing.assert_allclose(got, expected, rtol=0, atol=1e-6 * np.abs(expected).max())


def test_self_term_and_symmetry():
    x = np.array([0.1, 0.2, 0.3])
    np.testing.assert_allclose(vie.b_coeff(x, x, k=1.0, dv=0.01), -np.eye(3) / 3, atol=1e-15)
    y = np.array([0.4, 0.6, 1.0])
    np.testing.assert_allclose(vie.b_coeff(x, y, k=1.0, dv=0.01), vie.b_coeff(y, x, k=1.0, dv=0.01), atol=1e-15)
    B = vie.b_coeff(x, y, k=2.0, dv=0.01)
    np.testing.assert_allclose(B, B.T, atol=1e-15)


def _rand_field(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _lossy_aniso_chi(N, seed=3):
    rng = np.random.default_rng(seed)
    chi = (rng.standard_normal((3, 3) + tuple(N)) + 1j * rng.standard_normal((3, 3) + tuple(N))) * 0.3
    return chi


CASES = [((4, 4, 4), (1.0, 1.0, 1.0), 1.0), ((4, 3, 2), (1.0, 0.8, 0.6), 2.1), ((3, 3, 3), (1, 1, 1), 1.7)]


@pytest.mark.parametrize("N,L,k", CASES)
def test_fft_matvec_matches_dense(N, L, k):
    center = (0.0, 0.0, 0.0)
    chi = _lossy_aniso_chi(N)
    u = _rand_field((3,) + N, 42)
    A = vie.dense_operator_matrix(N, L, center, k, chi)
    expected = vie.unflatten_field(A @ vie.flatten_field(u), N)
    got = vie.em_matvec(chi, u, vie.kernel_hat(N, L, k))
    assert np.linalg.norm(got - expected) / np.linalg.norm(expected) < 1e-12


@pytest.mark.parametrize("N,L,k", CASES)
def test_fft_rmatvec_matches_dense_adjoint(N, L, k):
    center = (0.0, 0.0, 0.0)
    chi = _lossy_aniso_chi(N, 5)
    v = _rand_field((3,) + N, 7)
    A = vie.dense_operator_matrix(N, L, center, k, chi)
    expected = vie.unflatten_field(A.conj().T @ vie.flatten_field(v), N)
    Kh = vie.kernel_hat(N, L, k)
    derived = vie.em_rmatvec(chi, v, Khat=Kh)
    explicit = vie.em_rmatvec(chi, v, Khat_adj=vie.kernel_hat(N, L, k, adjoint=True))
    assert np.linalg.norm(derived - expected) / np.linalg.norm(expected) < 1e-12
    assert np.linalg.norm(explicit - expected) / np.linalg.norm(expected) < 1e-12


def test_zero_contrast_is_identity():
    N = (4, 3, 2)
    u = _rand_field((3,) + N, 1)
    chi = np.zeros((3, 3) + N, complex)
    np.testing.assert_allclose(vie.em_matvec(chi, u, vie.kernel_hat(N, (1, 1, 1), 1.0)), u, atol=1e-15)


def test_acoustic_matches_dense_and_adjoint():
    N, L, center, k0 = (3, 3, 2), (1.0, 0.8, 0.6), (0, 0, 0), 1.2
    eta = inputs.eta_ellipsoid(N, L, center, ell_center=(0, 0, 0), radii=(0.45,) * 3, eta_inside=2.0 + 0.2j)
    chi = eta - 1.0
    u = _rand_field(N, 123)
    H = vie.acoustic_dense_H(N, L, center, k0, chi)
    got = vie.acoustic_matvec(chi, u, vie.acoustic_kernel_hat(N, L, k0))
    np.testing.assert_allclose(got.reshape(-1), H @ u.reshape(-1), rtol=1e-12, atol=1e-12)
    v = _rand_field(N, 456)
    adj = vie.acoustic_rmatvec(chi, v, vie.acoustic_kernel_hat(N, L, k0, adjoint=True))
    np.testing.assert_allclose(adj.reshape(-1), H.conj().T @ v.reshape(-1), rtol=1e-12, atol=1e-12)


def test_acoustic_self_cell_low_frequency_limit():
    # exp(ika)(1-ika)-1 = (ka)^2/2 + O((ka)^3): small and finite as k -> 0
    v = vie.self_cell_coefficient(k0=1e-3, dv=1e-6)
    a = (3e-6 / (4 * np.pi)) ** (1 / 3)
    assert abs(v - (1e-3 * a) ** 2 / 2) < 1e-12


def test_planted_mistakes_are_caught():
    """A transposed contrast or a conjugated kernel must break dense parity."""
    N, L, k = (3, 3, 2), (1, 1, 1), 1.3
    chi = _lossy_aniso_chi(N)
    u = _rand_field((3,) + N, 2)
    A = vie.dense_operator_matrix(N, L, (0, 0, 0), k, chi)
    expected = vie.unflatten_field(A @ vie.flatten_field(u), N)
    wrong1 = vie.em_matvec(chi.swapaxes(0, 1), u, vie.kernel_hat(N, L, k))
    wrong2 = vie.em_matvec(chi, u, np.conj(vie.kernel_hat(N, L, k)))
    for w in (wrong1, wrong2):
        assert np.linalg.norm(w - expected) / np.linalg.norm(expected) > 1e-3
