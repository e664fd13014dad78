"""The fused Krylov steps of the device BiCGStab (p and s updates in the first
z pass, the dot products in the last), run on the host emulator of the HIP
passes, against NumPy formulas of the reference's steps
(src/em3d/solvers/bicgstab.py:96-116, 145-175) and the oracle's matvec.

The fast (register) z kernels take 2^k, 3 2^k and 5 2^k doubled lengths;
odd grid sizes take the generic kernels.  A right-hand side that is done keeps
its fields.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import vie
from tests.emu import emu


@pytest.fixture(scope="module")
def lib():
    return emu.load()


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _chi(kind, N, seed):
    d = _rand((3,) + N, seed) * 0.3
    chi = np.zeros((3, 3) + N, complex)
    for a in range(3):
        chi[a, a] = d[0] if kind == "iso" else d[a]
    return chi


CASES = [((8, 8, 8), "diag", "fast"), ((8, 4, 16), "iso", "fast"), ((12, 10, 24), "diag", "fast"),
         ((5, 6, 7), "diag", "generic"), ((4, 3, 2), "iso", "generic")]


def _coef(B, seed, first=(), done=()):
    rng = np.random.default_rng(seed)
    c = np.zeros(B, emu.KRY_COEF)
    for f in ("beta", "omega", "alpha"):
        c[f] = rng.uniform(-1, 1, (B, 2))
    c["first"] = [1 if b in first else 0 for b in range(B)]
    c["done"] = [1 if b in done else 0 for b in range(B)]
    return c


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
@pytest.mark.parametrize("N,kind,path", CASES)
def test_fused_p_update_and_rhat_dot(lib, N, kind, path, prec, tol):
    lib.emu_set_fast(1 if path == "fast" else 0)
    try:
        B, L, k = 3, (1.0, 0.9, 1.1), 3.0
        chi = _chi(kind, N, 1)
        op = emu.EmuOperator(lib, N, L, k, chi, 3, precision=prec, batch=B)
        p0, r, v, rh = (_rand((B, 3) + N, s) for s in (2, 3, 4, 5))
        coef = _coef(B, 6, first=(1,), done=(2,))
        p, y, _, dots = emu.apply_kry(op, p0, emu.KRY_P, r, v, emu.KRY_DOT, rh, coef)
        cdt = op.cdt
        Kh = vie.kernel_hat(N, L, k)
        for b in range(B):
            if b == 2:                                     # done: p untouched
                assert np.array_equal(p[b], p0[b].astype(cdt))
                continue
            be = complex(*coef["beta"][b]); om = complex(*coef["omega"][b])
            want = r[b] if b == 1 else r[b] + be * (p0[b] - om * v[b])
            assert _rel(p[b], want) < tol
            assert _rel(y[b], vie.em_matvec(chi, p[b].astype(complex), Kh)) < tol
            d = np.vdot(rh[b].astype(cdt).astype(complex), y[b].astype(complex))
            assert abs(complex(*dots[b, :2]) - d) <= 1e-12 * np.linalg.norm(rh[b]) * np.linalg.norm(y[b])
    finally:
        lib.emu_set_fast(1)


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
@pytest.mark.parametrize("N,kind,path", CASES)
def test_fused_s_update_and_t_dots(lib, N, kind, path, prec, tol):
    lib.emu_set_fast(1 if path == "fast" else 0)
    try:
        B, L, k = 2, (0.8, 1.0, 1.2), 4.0
        chi = _chi(kind, N, 7)
        op = emu.EmuOperator(lib, N, L, k, chi, 3, precision=prec, batch=B)
        s0, r, v = (_rand((B, 3) + N, sd) for sd in (8, 9, 10))
        coef = _coef(B, 11)
        s, t, ss, dots = emu.apply_kry(op, s0, emu.KRY_S, r, v, emu.KRY_DOT2, r, coef)
        Kh = vie.kernel_hat(N, L, k)
        for b in range(B):
            al = complex(*coef["alpha"][b])
            assert _rel(s[b], r[b] - al * v[b]) < tol
            sb, tb = s[b].astype(complex), t[b].astype(complex)
            assert abs(ss[b, 0] - np.vdot(sb, sb).real) <= 1e-12 * ss[b, 0]
            assert _rel(t[b], vie.em_matvec(chi, sb, Kh)) < tol
            tt, ts = np.vdot(tb, tb), np.vdot(tb, sb)
            assert abs(complex(dots[b, 0], dots[b, 1]) - tt) <= 1e-12 * abs(tt)
            assert abs(complex(dots[b, 2], dots[b, 3]) - ts) <= 1e-12 * np.linalg.norm(tb) * np.linalg.norm(sb)
    finally:
        lib.emu_set_fast(1)
