"""The GPU kernels' block phases, run thread by thread on the host, against the oracle.

tests/emu/emu_capi.cpp compiles the product's kernel headers with g++ and
replays every block of every pass (load, Stockham stages, spectral multiply,
store) on host memory, with the LDS tile poisoned by NaN.  This checks the
index arithmetic, pruning, octant folding and epilogues of the HIP kernels on
the CPU, before any GPU run.  It is not the GPU parity test: that is
tests/test_gpu_parity.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import inputs, vie
from tests.emu import emu


@pytest.fixture(scope="module")
def lib():
    return emu.load()


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


EM_CASES = [((4, 4, 4), (1.0, 1.0, 1.0), 1.0), ((4, 3, 2), (1.0, 0.8, 0.6), 2.1),
            ((3, 3, 3), (1.0, 1.0, 1.0), 1.7), ((5, 6, 7), (1.0, 1.2, 0.9), 3.0), ((8, 8, 8), (1, 1, 1), 4.0)]


@pytest.mark.parametrize("N,L,k", EM_CASES)
def test_em_full_contrast_matches_oracle(lib, N, L, k):
    chi = _rand((3, 3) + N, 3) * 0.3
    u = _rand((3,) + N, 4)
    op = emu.EmuOperator(lib, N, L, k, chi, 3)
    Kh = vie.kernel_hat(N, L, k)
    assert op.chi_kind == 3
    assert _rel(op.matvec(u), vie.em_matvec(chi, u, Kh)) < 1e-13
    assert _rel(op.rmatvec(u), vie.em_rmatvec(chi, u, Khat=Kh)) < 1e-13


@pytest.mark.parametrize("prec,tol", [("double", 1e-13), ("single", 1e-6)])
def test_em_iso_diag_batched(lib, prec, tol):
    N, L, k = (8, 8, 8), (1.0, 1.0, 1.0), 4.0
    iso = inputs.ellipsis_refraction(N, L, (0, 0, 0), eps_real=2.0, eps_imag=0.0, ell_center=(0, 0, 0),
                                     radius=(0.3, 0.3, 0.3))
    diag = iso.copy()
    diag[1, 1] *= 1.5 + 0.2j
    diag[2, 2] *= 0.7
    Kh = vie.kernel_hat(N, L, k)
    for chi, kind in ((iso, 1), (diag, 2)):
        op = emu.EmuOperator(lib, N, L, k, chi, 3, precision=prec, batch=3)
        assert op.chi_kind == kind
        u = _rand((3, 3) + N, 5)
        y, ya = op.matvec(u), op.rmatvec(u)
        for b in range(3):
            assert _rel(y[b], vie.em_matvec(chi, u[b], Kh)) < tol
            assert _rel(ya[b], vie.em_rmatvec(chi, u[b], Khat=Kh)) < tol


@pytest.mark.parametrize("N,L,k", [((3, 3, 2), (1.0, 0.8, 0.6), 1.2), ((6, 5, 4), (1, 1, 1), 3.0),
                                   ((8, 8, 8), (1, 1, 1), 10.0)])
def test_acoustic_matches_oracle(lib, N, L, k):
    chi = _rand(N, 6) * 0.3
    u = _rand(N, 7)
    op = emu.EmuOperator(lib, N, L, k, chi, 1)
    assert _rel(op.matvec(u), vie.acoustic_matvec(chi, u, vie.acoustic_kernel_hat(N, L, k))) < 1e-13
    assert _rel(op.rmatvec(u), vie.acoustic_rmatvec(chi, u, vie.acoustic_kernel_hat(N, L, k, adjoint=True))) < 1e-13


def test_unsupported_length_is_rejected(lib):
    assert not lib.emu_plan_create(11, 4, 4, 3, 1, 1)
    assert b"prime factor" in lib.emu_last_error()


def test_emulated_path_under_address_sanitizer(tmp_path):
    """Re-run one EM and one acoustic case in a subprocess with ASan loaded:
    every global access of the block phases must stay inside its buffer."""
    import os
    import subprocess
    import sys
    so = emu.build(str(tmp_path), extra=("-fsanitize=address", "-fno-omit-frame-pointer"))
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    script = f"""
import sys, numpy as np
sys.path.insert(0, {os.getcwd()!r})
from tests.emu import emu
from oracle import vie
lib = emu.load({so!r})
rng = np.random.default_rng(0)
N, L, k = (5, 4, 3), (1.0, 0.9, 0.7), 2.0
chi = rng.standard_normal((3, 3) + N) + 0j
u = rng.standard_normal((2, 3) + N) + 0j
op = emu.EmuOperator(lib, N, L, k, chi, 3, batch=2)
y = op.matvec(u); ya = op.rmatvec(u)
Kh = vie.kernel_hat(N, L, k)
assert np.allclose(y[1], vie.em_matvec(chi, u[1], Kh), atol=1e-12)
opa = emu.EmuOperator(lib, N, L, k, rng.standard_normal(N) + 0j, 1)
opa.matvec(rng.standard_normal(N) + 0j)
print("ok")
"""
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr[-3000:]


FAST_CASES = [((8, 8, 8), (1.0, 1.0, 1.0), 4.0), ((16, 8, 4), (1.0, 0.8, 0.6), 2.0),
              ((8, 16, 32), (1.0, 1.0, 1.0), 3.0), ((5, 8, 16), (1.0, 1.0, 1.0), 3.0),
              ((8, 4, 32), (1.0, 0.7, 1.3), 4.0), ((16, 6, 32), (1.0, 0.8, 0.6), 2.0),
              ((32, 8, 64), (1.0, 1.0, 1.0), 6.0), ((64, 4, 16), (1.0, 0.7, 1.3), 4.0),
              ((128, 2, 8), (1.0, 0.8, 0.6), 2.0),
              # 3 and 5 times a power of two: mixed-radix strided and z passes
              ((24, 40, 48), (1.0, 0.9, 1.1), 3.0), ((8, 24, 40), (1.0, 1.0, 1.0), 5.0),
              ((4, 96, 24), (0.6, 1.0, 0.8), 4.0), ((2, 8, 96), (0.5, 0.7, 1.0), 6.0),
              ((4, 4, 160), (0.5, 0.5, 1.2), 3.0), ((2, 192, 4), (0.4, 1.0, 0.5), 2.0),
              ((48, 4, 16), (1.0, 0.5, 0.6), 3.0), ((40, 6, 8), (1.0, 0.6, 0.4), 4.0),
              ((80, 2, 8), (1.0, 0.4, 0.3), 5.0), ((96, 3, 16), (1.0, 0.3, 0.5), 3.0),
              ((160, 2, 8), (1.0, 0.2, 0.2), 2.0), ((192, 2, 8), (1.0, 0.2, 0.2), 2.0),
              ((256, 2, 8), (1.0, 0.1, 0.1), 6.0)]   # 512-point x pass: 1024-thread blocks, ping-pong tiles


@pytest.mark.parametrize("prec,tol", [("double", 1e-13), ("single", 1e-6)])
@pytest.mark.parametrize("N,L,k", FAST_CASES)
def test_fast_kernels_match_oracle(lib, N, L, k, prec, tol):
    """Doubled axes of 2^k, 3 2^k or 5 2^k points take the register-resident
    two-stage kernels (e3_fast.hpp); other axes combine them with the generic
    passes."""
    lib.emu_set_fast(1)
    before = [lib.emu_fast_calls(i) for i in range(4)]
    chi = _rand((3, 3) + N, 8) * 0.3
    u = _rand((2, 3) + N, 9)
    op = emu.EmuOperator(lib, N, L, k, chi, 3, precision=prec, batch=2)
    Kh = vie.kernel_hat(N, L, k)
    y, ya = op.matvec(u), op.rmatvec(u)
    for b in range(2):
        assert _rel(y[b], vie.em_matvec(chi, u[b], Kh)) < tol
        assert _rel(ya[b], vie.em_rmatvec(chi, u[b], Khat=Kh)) < tol
    # these grids take the small-grid x and z tiles by default; the default tiles too
    lib.emu_plan_set_variant(op.h, 1, 0)
    lib.emu_plan_set_variant(op.h, 2, 0)
    y, ya = op.matvec(u), op.rmatvec(u)
    for b in range(2):
        assert _rel(y[b], vie.em_matvec(chi, u[b], Kh)) < tol
        assert _rel(ya[b], vie.em_rmatvec(chi, u[b], Khat=Kh)) < tol
    ca = _rand(N, 10) * 0.3
    ua = _rand((2,) + N, 11)
    opa = emu.EmuOperator(lib, N, L, k, ca, 1, precision=prec, batch=2)
    Ka, Kaa = vie.acoustic_kernel_hat(N, L, k), vie.acoustic_kernel_hat(N, L, k, adjoint=True)
    ya, yaa = opa.matvec(ua), opa.rmatvec(ua)
    for b in range(2):
        assert _rel(ya[b], vie.acoustic_matvec(ca, ua[b], Ka)) < tol
        assert _rel(yaa[b], vie.acoustic_rmatvec(ca, ua[b], Kaa)) < tol
    used = [lib.emu_fast_calls(i) - before[i] for i in range(4)]
    fast_len = {16, 32, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512}
    if 2 * N[2] in fast_len:
        assert used[2] > 0 and used[3] > 0      # z passes take the fast kernels
    if 2 * N[1] in fast_len:
        assert used[0] > 0                       # strided y passes on the fast kernel
    if N[2] % 8 == 0 and 2 * N[0] in fast_len:
        assert used[1] > 0                       # x pass on the fast kernel


def test_generic_and_fast_paths_agree(lib):
    N, L, k = (8, 8, 8), (1.0, 1.0, 1.0), 5.0
    chi = _rand((3, 3) + N, 12) * 0.3
    u = _rand((3,) + N, 13)
    lib.emu_set_fast(0)
    try:
        slow = emu.EmuOperator(lib, N, L, k, chi, 3).matvec(u)
    finally:
        lib.emu_set_fast(1)
    fast = emu.EmuOperator(lib, N, L, k, chi, 3).matvec(u)
    assert _rel(fast, slow) < 1e-14


VARIANT_CASES = [((128, 2, 8), 1, "single"), ((96, 2, 8), 1, "single"), ((64, 2, 8), 1, "single"),  # x: 256/192/128
                 ((2, 128, 8), 0, "single"), ((2, 96, 8), 0, "single"),          # strided passes at 256, 192
                 ((2, 2, 128), 2, "single"), ((2, 2, 96), 2, "single"),          # z passes at 256, 192
                 ((128, 2, 8), 1, "double"), ((2, 128, 8), 0, "double")]         # FP64 x and strided at 256


@pytest.mark.parametrize("N,kind,precision", VARIANT_CASES)
def test_every_kernel_variant_matches_oracle(lib, N, kind, precision):
    """Each variant the autotuner may pick (and the generic kernels, -1)
    gives the oracle's matvec and rmatvec."""
    L, k = (1.0, 0.6, 0.8), 3.0
    chi = _rand((3, 3) + N, 21) * 0.3
    u = _rand((1, 3) + N, 22)
    Kh = vie.kernel_hat(N, L, k)
    ref, refa = vie.em_matvec(chi, u[0], Kh), vie.em_rmatvec(chi, u[0], Khat=Kh)
    op = emu.EmuOperator(lib, N, L, k, chi, 3, precision=precision, batch=1)
    tol = 1e-6 if precision == "single" else 1e-12
    n = lib.emu_plan_variant_count(op.h, kind)
    assert n >= 2
    for v in range(-1, n):
        lib.emu_plan_set_variant(op.h, kind, v)
        assert _rel(op.matvec(u)[0], ref) < tol, v
        assert _rel(op.rmatvec(u)[0], refa) < tol, v
