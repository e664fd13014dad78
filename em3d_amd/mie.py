"""Mie series of a dielectric sphere and the RCS comparison of the
far-field pipeline against it (src/em3d/mie.py:18-240 of the reference).

Host post-processing in NumPy/SciPy, as in the reference.  The series is
the textbook one (Bohren & Huffman, "Absorption and Scattering of Light by
Small Particles", ch. 4): Riccati-Bessel functions psi_n(z) = z j_n(z),
xi_n(z) = z h_n(z) with h_n = j_n + i y_n, scattering coefficients a_n, b_n,
angle functions pi_n, tau_n from their upward recurrence, amplitudes S1, S2,
terms up to n_max = round(x + 4 x^(1/3) + 2) (Wiscombe).  The comparison
figures (shape error of the max-normalised curves, peak ratio, maximum error
relative to the Mie peak) are the reference's, so results are comparable
with its tables.
"""
from __future__ import annotations

import numpy as np
from scipy.special import spherical_jn, spherical_yn

__all__ = ["mie_coefficients", "mie_cross_sections", "mie_rcs_plane", "compare_rcs_plane"]


def _check(a: float, eps_r: complex, k0: float) -> None:
    if a <= 0:
        raise ValueError(f"a must be > 0, got {a}")
    if k0 <= 0:
        raise ValueError(f"k0 must be > 0, got {k0}")
    if np.real(eps_r) <= 0 and np.imag(eps_r) == 0:
        raise ValueError(f"Re(eps_r) must be > 0 for lossless medium, got {eps_r}")


def n_terms(x: float) -> int:
    return max(1, round(float(x) + 4.0 * float(x) ** (1.0 / 3.0) + 2.0))


def _psi_xi(n, z):
    """psi_n, psi_n', xi_n, xi_n' at z for an array of orders n."""
    j, dj = spherical_jn(n, z), spherical_jn(n, z, derivative=True)
    y, dy = spherical_yn(n, z), spherical_yn(n, z, derivative=True)
    h, dh = j + 1j * y, dj + 1j * dy
    return z * j, j + z * dj, z * h, h + z * dh


def mie_coefficients(a: float, eps_r: complex, k0: float) -> dict:
    """a_n, b_n (scattered) and c_n, d_n (internal) for n = 1 .. n_max."""
    _check(a, eps_r, k0)
    x = k0 * a
    nmax = n_terms(x)
    if abs(complex(eps_r) - 1.0) <= 1e-14:
        z, o = np.zeros(nmax, np.complex128), np.ones(nmax, np.complex128)
        return {"a": z.copy(), "b": z.copy(), "c": o.copy(), "d": o.copy(), "n_max": nmax}
    m = np.sqrt(complex(eps_r))
    n = np.arange(1, nmax + 1)
    ps_m, dps_m, _, _ = _psi_xi(n, m * x)
    ps, dps, xi, dxi = _psi_xi(n, complex(x))
    da = m * ps_m * dxi - xi * dps_m
    db = ps_m * dxi - m * xi * dps_m
    an = (m * ps_m * dps - ps * dps_m) / da
    bn = (ps_m * dps - m * ps * dps_m) / db
    return {"a": an.astype(np.complex128), "b": bn.astype(np.complex128), "c": 1j * m / db, "d": 1j * m / da,
            "n_max": nmax}


def mie_cross_sections(a: float, eps_r: complex, k0: float) -> dict:
    c = mie_coefficients(a, eps_r, k0)
    n = np.arange(1, c["n_max"] + 1, dtype=np.float64)
    w = 2.0 * n + 1.0
    scat = float((2.0 * np.pi / k0 ** 2) * np.sum(w * (np.abs(c["a"]) ** 2 + np.abs(c["b"]) ** 2)))
    ext = float((2.0 * np.pi / k0 ** 2) * np.sum(w * np.real(c["a"] + c["b"])))
    return {"scat": scat, "ext": ext, "abs": ext - scat}


def angle_functions(nmax: int, mu):
    """pi_n(mu), tau_n(mu) for n = 1 .. nmax, shape (nmax, len(mu))."""
    mu = np.asarray(mu, dtype=np.float64)
    pi = np.zeros((nmax + 1, mu.size))
    tau = np.zeros((nmax + 1, mu.size))
    pi[1] = 1.0
    for n in range(1, nmax + 1):
        if n >= 2:
            pi[n] = ((2 * n - 1) * mu * pi[n - 1] - n * pi[n - 2]) / (n - 1)
        tau[n] = n * mu * pi[n] - (n + 1) * pi[n - 1]
    return pi[1:], tau[1:]


def _amplitudes(c, mu):
    n = np.arange(1, c["n_max"] + 1, dtype=np.float64)
    w = (2.0 * n + 1.0) / (n * (n + 1.0))
    pi, tau = angle_functions(c["n_max"], mu)
    S1 = (w * c["a"]) @ pi + (w * c["b"]) @ tau
    S2 = (w * c["a"]) @ tau + (w * c["b"]) @ pi
    return S1, S2


def mie_rcs_plane(a: float, eps_r: complex, k0: float, n_phi: int = 180, plane: str = "xy"):
    """(phi, sigma) of a sphere lit along +z with x polarisation, sigma = |S|^2 / k0^2
    over n_phi directions of the plane (the reference's plane conventions)."""
    _check(a, eps_r, k0)
    if plane not in ("xy", "xz", "yz"):
        raise ValueError(f"plane must be 'xy', 'xz', or 'yz', got {plane!r}")
    if n_phi < 1:
        raise ValueError(f"n_phi must be >= 1, got {n_phi}")
    c = mie_coefficients(a, eps_r, k0)
    phi = np.linspace(0.0, 2.0 * np.pi, n_phi, endpoint=False)
    if plane == "xy":
        S1, S2 = _amplitudes(c, np.zeros(n_phi))          # every direction at 90 degrees from +z
        sigma = (np.abs(S2) ** 2 * np.cos(phi) ** 2 + np.abs(S1) ** 2 * np.sin(phi) ** 2) / k0 ** 2
    else:
        S1, S2 = _amplitudes(c, np.cos(np.pi / 2.0 - phi))
        sigma = (np.abs(S2) ** 2 if plane == "xz" else np.abs(S1) ** 2) / k0 ** 2
    return phi, np.real(sigma).astype(np.float64)


def _peak_normalised(sigma):
    peak = float(np.max(sigma)) if sigma.size else 0.0
    return (sigma / peak if peak > 0.0 else np.zeros_like(sigma, dtype=np.float64)), peak


def compare_rcs_plane(u, problem, *, a: float, eps_r: complex, n_phi: int = 180, plane: str = "xy",
                      method: str = "direct", batch_size: int = 64, normalize: str = "max") -> dict:
    """Numerical RCS (farfield.rcs_plane) against the Mie curve: shape error of
    the max-normalised curves, ratio of the peaks, max error / Mie peak."""
    from . import farfield
    if normalize != "max":
        raise ValueError(f"normalize must be 'max', got {normalize!r}")
    if n_phi < 1:
        raise ValueError(f"n_phi must be >= 1, got {n_phi}")
    phi, num = farfield.rcs_plane(u, problem, n_phi=n_phi, plane=plane, method=method, batch_size=batch_size)
    _, mie = mie_rcs_plane(a, eps_r, problem.k0, n_phi=n_phi, plane=plane)
    num = np.asarray(num, dtype=np.float64)
    num_n, num_peak = _peak_normalised(num)
    mie_n, mie_peak = _peak_normalised(mie)
    den = float(np.linalg.norm(mie_n))
    if den > 0.0:
        shape_err = float(np.linalg.norm(num_n - mie_n) / den)
    else:
        shape_err = 0.0 if np.linalg.norm(num_n) == 0.0 else float("inf")
    if mie_peak > 0.0:
        scale_ratio = float(num_peak / mie_peak)
        abs_rel_err = float(np.max(np.abs(num - mie)) / mie_peak)
    else:
        scale_ratio = float("nan")
        abs_rel_err = 0.0 if num_peak == 0.0 else float("inf")
    return {"phi": phi, "sigma_num": num, "sigma_mie": mie, "sigma_num_norm": num_n, "sigma_mie_norm": mie_n,
            "shape_err": shape_err, "scale_ratio": scale_ratio, "abs_rel_err": abs_rel_err}
