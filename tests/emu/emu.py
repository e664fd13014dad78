"""ctypes wrapper of the host emulator of the HIP pass kernels (test only).

``load()`` compiles tests/emu/emu_capi.cpp with g++ into a temporary
directory (a few seconds) and returns the library.  The emulator runs the
product's kernel phases thread by thread on host memory; see emu_capi.cpp.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(out_dir: str | None = None, extra=()) -> str:
    out_dir = out_dir or tempfile.mkdtemp(prefix="em3d_emu_")
    so = os.path.join(out_dir, "libem3d_emu.so")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *extra,
           os.path.join(HERE, "emu_capi.cpp"), "-o", so]
    subprocess.run(cmd, check=True, capture_output=True)
    return so


def load(path: str | None = None):
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    lib = C.CDLL(path or build())
    lib.emu_plan_create.restype = C.c_void_p
    lib.emu_plan_create.argtypes = [C.c_int] * 6
    lib.emu_plan_destroy.argtypes = [C.c_void_p]
    lib.emu_plan_sizes.argtypes = [C.c_void_p] + [C.POINTER(C.c_longlong)] * 3
    lib.emu_kernel_build.argtypes = [C.c_void_p] + [C.c_double] * 4 + [C.c_void_p] * 2
    lib.emu_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_int]
    lib.emu_apply_kry.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p]
    lib.emu_apply_window.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]
    lib.emu_last_error.restype = C.c_char_p
    lib.emu_plan_variant_count.argtypes = [C.c_void_p, C.c_int]
    lib.emu_plan_set_variant.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.emu_set_fast.argtypes = [C.c_int]
    lib.emu_fast_calls.argtypes = [C.c_int]
    lib.emu_fast_calls.restype = C.c_longlong
    if path is None:
        _LIB = lib
    return lib


class EmuOperator:
    """Emulated HIP operator: khat built by the emulated kernel build."""

    def __init__(self, lib, N, L, k0, chi, ncomp, precision="double", batch=1, chi_kind=None):
        self.lib = lib
        self.N = tuple(int(n) for n in N)
        self.prec = 1 if precision == "double" else 0
        self.cdt = np.complex128 if self.prec else np.complex64
        self.ncomp = ncomp
        self.batch = batch
        self.h = lib.emu_plan_create(*self.N, ncomp, batch, self.prec)
        if not self.h:
            raise ValueError(lib.emu_last_error().decode())
        ke, we, be = C.c_longlong(), C.c_longlong(), C.c_longlong()
        lib.emu_plan_sizes(self.h, C.byref(ke), C.byref(we), C.byref(be))
        self.khat = np.zeros(ke.value, self.cdt)
        self.work = np.zeros(we.value, self.cdt)
        build = np.zeros(be.value, np.complex128)
        hx, hy, hz = (L[a] / N[a] for a in range(3))
        lib.emu_kernel_build(self.h, float(k0), hx, hy, hz, self.khat.ctypes.data, build.ctypes.data)
        self.chi_kind, self.chi = pack_chi(chi, ncomp, self.cdt, chi_kind)

    def _apply(self, u, adjoint):
        u = np.ascontiguousarray(u, dtype=self.cdt)
        y = np.empty_like(u)
        self.lib.emu_apply(self.h, self.khat.ctypes.data, self.chi.ctypes.data, self.chi_kind,
                           u.ctypes.data, y.ctypes.data, self.work.ctypes.data, int(adjoint))
        return y

    def matvec(self, u):
        return self._apply(u, False)

    def matvec_window(self, u, lo, n):
        """A u with the forward passes restricted to the box [lo, lo + n)."""
        u = np.ascontiguousarray(u, dtype=self.cdt)
        y = np.empty_like(u)
        win = np.array(list(lo) + list(n), dtype=np.int32)
        self.lib.emu_apply_window(self.h, self.khat.ctypes.data, self.chi.ctypes.data, self.chi_kind,
                                  u.ctypes.data, y.ctypes.data, self.work.ctypes.data, win.ctypes.data)
        return y

    def rmatvec(self, v):
        return self._apply(v, True)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.emu_plan_destroy(self.h)
            self.h = None


def pack_chi(chi, ncomp, cdt, kind=None):
    """Pack a reference-layout contrast into the library's chi layout."""
    chi = np.asarray(chi)
    if ncomp == 1:
        return 1, np.ascontiguousarray(chi, dtype=cdt)
    if kind is None:
        off = chi.copy()
        for a in range(3):
            off[a, a] = 0
        if not np.any(off):
            d = np.stack([chi[a, a] for a in range(3)])
            kind = 1 if (np.array_equal(d[0], d[1]) and np.array_equal(d[0], d[2])) else 2
        else:
            kind = 3
    if kind == 1:
        return 1, np.ascontiguousarray(chi[0, 0], dtype=cdt)
    if kind == 2:
        return 2, np.ascontiguousarray(np.stack([chi[a, a] for a in range(3)]), dtype=cdt)
    return 3, np.ascontiguousarray(chi, dtype=cdt)


KRY_NONE, KRY_P, KRY_S, KRY_DOT, KRY_DOT2 = 0, 1, 2, 3, 4
KRY_COEF = np.dtype([("beta", np.float64, 2), ("omega", np.float64, 2), ("alpha", np.float64, 2),
                     ("first", np.int32), ("done", np.int32)])


def apply_kry(op, u, op1, r, v, op5, rh, coef, stride=4096):
    """y = A u with the fused Krylov steps (u updated in place by op1);
    returns (y, per-RHS sums of the P1 partials, of the P5 partials)."""
    u = np.array(u, dtype=op.cdt, order="C", copy=True)     # updated in place by op1
    r = np.ascontiguousarray(r, dtype=op.cdt)
    v = np.ascontiguousarray(v, dtype=op.cdt)
    rh = np.ascontiguousarray(rh, dtype=op.cdt)
    coef = np.ascontiguousarray(coef, dtype=KRY_COEF)
    y = np.empty_like(u)
    B = op.batch
    p1 = np.full((B, stride, 4), np.nan)
    p5 = np.full((B, stride, 4), np.nan)
    counts = np.zeros(2, np.int32)
    op.lib.emu_apply_kry(op.h, op.khat.ctypes.data, op.chi.ctypes.data, op.chi_kind, u.ctypes.data, y.ctypes.data,
                         op.work.ctypes.data, op1, r.ctypes.data, v.ctypes.data, op5, rh.ctypes.data,
                         coef.ctypes.data, p1.ctypes.data, p5.ctypes.data, stride, counts.ctypes.data)
    s1 = p1[:, :counts[0]].sum(axis=1) if op1 == KRY_S else None
    s5 = p5[:, :counts[1]].sum(axis=1) if op5 in (KRY_DOT, KRY_DOT2) else None
    return u, y, s1, s5
