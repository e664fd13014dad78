"""ctypes binding of libem3d_hip.so (the C-ABI declared in include/em3d_hip.h).

The product never falls back to another implementation: if the library is not
built, or no GPU is visible, every compute entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libem3d_hip.so")

class BiCGStabParams(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("rtol", C.c_double), ("atol", C.c_double),
                ("divergence_guard", C.c_double), ("max_operator_actions", C.c_longlong), ("chunk", C.c_int),
                ("fused", C.c_int)]


class SolverResultC(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("matvec_count", C.c_int),
                ("history_len", C.c_int), ("true_final_residual", C.c_double)]


class FnoConfig(C.Structure):
    _fields_ = [("grid", C.c_int), ("channels", C.c_int), ("modes", C.c_int), ("layers", C.c_int),
                ("in_channels", C.c_int), ("out_channels", C.c_int), ("material_channels", C.c_int)]


class FnoModel(C.Structure):
    _fields_ = [("config", FnoConfig), ("factor", C.c_int), ("params", C.c_void_p), ("params_bf16", C.c_void_p),
                ("work", C.c_void_p), ("features", C.c_void_p), ("chi_features", C.c_void_p),
                ("chi_channels", C.c_int)]


STATUS = {0: "converged", 1: "nonfinite", 2: "max_operator_actions", 3: "breakdown_rho", 4: "breakdown_omega",
          5: "breakdown_alpha", 6: "breakdown_t", 7: "divergence_guard", 8: "max_iter"}

# Every symbol the header declares, with (restype, argtypes).
_VP, _I, _D, _LL, _SZ = C.c_void_p, C.c_int, C.c_double, C.c_longlong, C.c_size_t
SIGNATURES = {
    "em3d_version": (C.c_char_p, []),
    "em3d_last_error": (C.c_char_p, []),
    "em3d_plan_create": (_I, [_I, _I, _I, _I, _I, _I, C.POINTER(_VP)]),
    "em3d_plan_destroy": (_I, [_VP]),
    "em3d_plan_sizes": (_I, [_VP, C.POINTER(_SZ), C.POINTER(_SZ), C.POINTER(_SZ)]),
    "em3d_kernel_build": (_I, [_VP, _D, _D, _D, _D, _VP, _VP, _VP]),
    "em3d_matvec": (_I, [_VP, _VP, _VP, _I, _VP, _VP, _VP, _VP]),
    "em3d_rmatvec": (_I, [_VP, _VP, _VP, _I, _VP, _VP, _VP, _VP]),
    "em3d_matvec_noalloc": (_I, [_VP, _VP, _VP, _I, _VP, _VP, _VP]),
    "em3d_matvec_window": (_I, [_VP, _VP, _VP, _I, _VP, _VP, _VP, C.POINTER(_I), C.POINTER(_I), _VP]),
    "em3d_plan_set_timing": (_I, [_VP, _I]),
    "em3d_plan_timing": (_I, [_VP, C.POINTER(_D), _I, C.POINTER(_I), C.POINTER(_LL)]),
    "em3d_group_create": (_I, [_I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.POINTER(_VP)]),
    "em3d_group_matvec": (_I, [_VP, _VP]),
    "em3d_group_destroy": (_I, [_VP]),
    "em3d_plan_variant_count": (_I, [_VP, _I]),
    "em3d_plan_set_variants": (_I, [_VP, C.POINTER(C.c_int)]),
    "em3d_plan_get_variants": (_I, [_VP, C.POINTER(C.c_int)]),
    "em3d_plan_autotune": (_I, [_VP, _VP, _VP, _I, _VP, _VP, _VP, _I, C.POINTER(C.c_double), _VP]),
    "em3d_bicgstab_work_bytes": (_SZ, [_VP]),
    "em3d_bicgstab": (_I, [_VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, C.POINTER(BiCGStabParams), _VP, _VP, _I,
                           C.POINTER(SolverResultC), _VP]),
    "em3d_convert_axpy": (_I, [_LL, _I, _VP, _D, _D, _I, _VP, _I, _VP]),
    "em3d_support_bounds": (_I, [_I, _I, _I, _I, _VP, _I, _VP, _VP]),
    "em3d_support_gather": (_I, [_I, _LL, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "em3d_support_scatter": (_I, [_I, _LL, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "em3d_scatter_work_bytes": (_SZ, [_I, _I, _I, _I, _I, _I]),
    "em3d_scatter_integral": (_I, [_I, _I, _I, _VP, _VP, _VP, _I, _I, _I, _VP, _I, _VP, _I, _VP, _I, _D, _D, _VP,
                                   _VP, _VP]),
    "em3d_plane_waves": (_I, [_I, _I, _I, _VP, _VP, _VP, _I, _I, _VP, _D, _I, _VP, _VP]),
    "em3d_ellipsoid_contrast": (_I, [_I, _I, _I, _VP, _VP, _VP, _I, _I, _VP, _VP, _D, _I, _VP, _VP]),
    "em3d_eta_inclusions": (_I, [_I, _I, _I, _VP, _VP, _VP, _I, _I, _VP, _D, _D, _VP, _VP]),
    "em3d_fno_param_count": (_LL, [C.POINTER(FnoConfig)]),
    "em3d_fno_work_bytes": (_SZ, [C.POINTER(FnoConfig), _I]),
    "em3d_fno_init_work": (_I, [C.POINTER(FnoConfig), _I, _VP, _VP]),
    "em3d_fno_forward": (_I, [C.POINTER(FnoConfig), _VP, _VP, _VP, _VP, _VP, _I, _VP]),
    "em3d_fno_backward": (_I, [C.POINTER(FnoConfig), _VP, _VP, _VP, _VP, _VP, _I, _VP]),
    "em3d_fno_loss_work_bytes": (_SZ, [_I]),
    "em3d_fno_loss": (_I, [_I, _LL, _VP, _VP, _VP, _VP, _VP, _VP]),
    "em3d_fno_adam": (_I, [_LL, _VP, _VP, _VP, _VP, _VP, _D, _D, _D, _D, _D, _I, _VP]),
    "em3d_fno_to_bf16": (_I, [_LL, _VP, _VP, _VP]),
    "em3d_fno_features_bytes": (_SZ, [C.POINTER(FnoConfig), _I]),
    "em3d_fno_scale_work_bytes": (_SZ, [_I]),
    "em3d_fno_scale": (_I, [_I, _LL, _I, _VP, _VP, _VP, _VP]),
    "em3d_fno_restrict": (_I, [_I, _I, _I, _I, _I, _I, _I, _VP, _VP, _VP, _I, _VP, _VP]),
    "em3d_fno_chi_features": (_I, [_I, _I, _I, _I, _I, _VP, _I, _I, _D, _VP, _VP]),
    "em3d_fno_prolong": (_I, [_I, _I, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP]),
    "em3d_fno_prolong_adjoint": (_I, [_I, _I, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP]),
    "em3d_fno_precondition": (_I, [C.POINTER(FnoModel), _I, _I, _I, _I, _I, _I, _VP, _VP, _VP]),
    "em3d_bicgstab_pc_work_bytes": (_SZ, [_VP]),
    "em3d_bicgstab_pc": (_I, [_VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, C.POINTER(BiCGStabParams), _VP, _VP, _I,
                              C.POINTER(SolverResultC), C.POINTER(FnoModel), _VP]),
    "em3d_reduce_work_bytes": (_SZ, []),
    "em3d_dotc": (_I, [_I, _LL, _VP, _VP, _VP, _VP, _VP]),
    "em3d_axpbypcz": (_I, [_I, _LL, _D, _D, _VP, _D, _D, _VP, _D, _D, _VP, _VP, _VP]),
}

_LIB = None


class HipLibraryError(RuntimeError):
    """The HIP library is missing or a call into it failed."""


def lib():
    """Load (once) and return the library, with every signature declared."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HipLibraryError(
                f"{LIB_PATH} is not built; run `python -m em3d_amd.build` (hipcc, gfx950)")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().em3d_last_error().decode(errors="replace")
        raise HipLibraryError(f"{what}: {msg}" if what else msg)


def stream_handle(device=None) -> int:
    """HIP stream of torch's current stream on the given device."""
    import torch
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int:
    return t.data_ptr() if t is not None else None
