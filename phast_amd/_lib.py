"""ctypes binding of ``libphast_mi.so`` (the C ABI declared in ``include/phast_mi.h``).

The product path has no CPU fallback: importing a compute entry point without the
built library, or calling one with CPU tensors, raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libphast_mi.so"

PHAST_OK = 0
PHAST_ERR_ARG = -1
PHAST_ERR_HIP = -2
PHAST_ERR_NONPOSITIVE = -3
PHAST_ERR_NONFINITE = -4
PHAST_ERR_NOT_CONVERGED = -5
PHAST_ERR_UNSUPPORTED = -6

CG_RUNNING, CG_CONVERGED, CG_MAXITER, CG_NONPOSITIVE, CG_DIVERGED, CG_NONFINITE = range(6)
CG_STATUS_NAMES = {CG_RUNNING: "running", CG_CONVERGED: "converged", CG_MAXITER: "max_iter",
                   CG_NONPOSITIVE: "nonpositive", CG_DIVERGED: "diverged", CG_NONFINITE: "nonfinite"}


class PhastLibraryError(RuntimeError):
    """The HIP library is missing or a call failed inside the HIP runtime."""


class Q4Desc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("gradN", C.c_double * 32), ("N", C.c_double * 16), ("wdetJ", C.c_double * 4),
        ("lam", C.c_double), ("mu", C.c_double), ("kappa", C.c_double), ("nu", C.c_double),
        ("plane_stress", C.c_int32), ("split", C.c_int32),
        ("eta_residual", C.c_double), ("Gc", C.c_double), ("l0", C.c_double), ("dt", C.c_double),
        ("hx", C.c_double), ("hy", C.c_double),
        ("stress_split", C.c_int32),
    ]


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_D = C.c_double

_SIGNATURES = {
    "phast_version": ([], C.c_char_p),
    "phast_last_error": ([], C.c_char_p),
    "phast_q4_plan_create": ([C.POINTER(Q4Desc), _I32, _I32, _I32, C.POINTER(_P)], C.c_int),
    "phast_q4_plan_destroy": ([_P], C.c_int),
    "phast_q4_plan_sizes": ([_P, C.POINTER(_I64), C.POINTER(_I64)], C.c_int),
    "phast_q4_plan_offsets": ([_P, _P, _P], C.c_int),
    "phast_q4_lumped_mass": ([_P, _P, _P], C.c_int),
    "phast_q4_internal_force": ([_P, _P, _P, _P, _P], C.c_int),
    "phast_q4_strain": ([_P, _P, _P, _P], C.c_int),
    "phast_q4_driving_force": ([_P, _P, _P, _P, _P, _P], C.c_int),
    "phast_history_max": ([_I64, _P, _P, _P, _P], C.c_int),
    "phast_q4_elem_to_node": ([_P, _P, _P, _P, _P], C.c_int),
    "phast_q4_explicit_step": ([_P] * 17, C.c_int),
    "phast_q4_damping_force": ([_P, _P, _P, _D, _P, _P], C.c_int),
    "phast_q4_damage_rhs_diag": ([_P] * 6, C.c_int),
    "phast_q4_damage_apply": ([_P] * 6, C.c_int),
    "phast_q4_damage_solve": ([_P, _I32, _P, _P, _P, _P, _P, _P, _D, _I32, _P, _P, _P, _P, _P], C.c_int),
    "phast_q4_damage_solve_noalloc": ([_P, _I32, _P, _P, _P, _P, _P, _P, _D, _I32, _P, _P, _P, _P, _P], C.c_int),
    "phast_q4_batch_prepare": ([_P, _P], C.c_int),
    "phast_q4_batch_step": ([_P, _I32] + [_P] * 12 + [_D, _I32] + [_P] * 9, C.c_int),
    "phast_q4_batch_error": ([_P, C.POINTER(_I32), _I32], C.c_int),
    "phast_q4_batch_run": ([_P, _I32, _I32] + [_P] * 12 + [_D, _I32] + [_P] * 9, C.c_int),
    "phast_t3_plan_create": ([_I64, _I64, _P, _P, _D, _D, _D, _D, _D, _I32, _I32, _I32, _I32, _D, C.POINTER(_P)],
                             C.c_int),
    "phast_t3_plan_destroy": ([_P], C.c_int),
    "phast_q4u_plan_create": ([_I64, _I64, _P, _P, _D, _D, _D, _D, _D, _I32, _I32, _I32, C.POINTER(_P)], C.c_int),
    "phast_t3_mesh_info": ([_P, _P, _P, _P], C.c_int),
    "phast_t3_set_path": ([_P, _I32, C.POINTER(_I32), C.POINTER(_I32)], C.c_int),
    "phast_t3_internal_force": ([_P, _P, _P, _P, _P], C.c_int),
    "phast_t3_force_history": ([_P] * 8, C.c_int),
    "phast_t3_explicit_step": ([_P, _D] + [_P] * 16, C.c_int),
    "phast_t3_damping_force": ([_P, _P, _P, _D, _P, _P], C.c_int),
    "phast_t3_element_fields": ([_P] * 7, C.c_int),
    "phast_t3_elem_to_node": ([_P] * 5, C.c_int),
    "phast_t3_damage_apply": ([_P] * 6, C.c_int),
    "phast_t3_damage_rhs_diag": ([_P] * 6, C.c_int),
    "phast_t3_damage_solve": ([_P, _I32] + [_P] * 6 + [_D, _I32] + [_P] * 5, C.c_int),
    "phast_t3_set_fracture": ([_P, _D, _D, _D], C.c_int),
    "phast_t3_adjoint_solve": ([_P] * 7 + [_D, _I32] + [_P] * 4, C.c_int),
    "phast_t3_adjoint_terms": ([_P] * 9, C.c_int),
    "phast_hex8_plan_create": ([_I32, _I32, _I32, _D, _D, _D, _P, _I32, _P, _P, _P, _P, _P, _D, _D,
                                C.POINTER(_P)], C.c_int),
    "phast_hex8_plan_destroy": ([_P], C.c_int),
    "phast_hex8_set_tiles": ([_P, _I32, _I32, _I32], C.c_int),
    "phast_hex8_degradation": ([_P, _P, _P, _P], C.c_int),
    "phast_hex8_mech_action": ([_P, _P, _P, _P, _P], C.c_int),
    "phast_hex8_mech_action_d": ([_P, _P, _P, _P, _P], C.c_int),
    "phast_hex8_mech_diagonal": ([_P, _P, _P, _P], C.c_int),
    "phast_hex8_stress_sums": ([_P, _P, _P, _P, _P], C.c_int),
    "phast_hex8_driving_force": ([_P, _P, _P, _P], C.c_int),
    "phast_hex8_damage_action": ([_P, _P, _P, _D, _P, _P], C.c_int),
    "phast_hex8_damage_rhs": ([_P, _P, _D, _P, _P], C.c_int),
    "phast_hex8_damage_diagonal": ([_P, _P, _D, _P, _P], C.c_int),
    "phast_hex8_set_gc": ([_P, _P, _P], C.c_int),
    "phast_hex8_adj_coupling": ([_P] * 6, C.c_int),
    "phast_hex8_adj_driving": ([_P] * 5, C.c_int),
    "phast_hex8_qp_to_node": ([_P, _P, _D, _P, _P], C.c_int),
    "phast_hex8_qp_stress_apply": ([_P] * 4, C.c_int),
    "phast_hex8_field_snapshot": ([_P] * 7, C.c_int),
    "phast_hex8_adj_damage_history": ([_P, _P, _P, _D, _P, _P], C.c_int),
    "phast_hex8_adj_damage_gc": ([_P, _P, _P, _D, _P, _P], C.c_int),
}

_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """Load (once) and return the HIP library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise PhastLibraryError(
                f"{LIB_PATH} is missing: build it with `python -m phast_amd.build` "
                "(the HIP path has no CPU fallback)")
        lib_ = C.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib_, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = res
        _lib = lib_
    return _lib


def exported_symbols() -> list[str]:
    """Names of the declared entry points that the loaded library exports."""
    lib_ = lib()
    return [n for n in _SIGNATURES if hasattr(lib_, n)]


def check(rc: int, what: str = "") -> None:
    """Map a C return code onto the exception classes used by the reference."""
    if rc == PHAST_OK:
        return
    msg = lib().phast_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == PHAST_ERR_ARG:
        raise ValueError(text)
    if rc == PHAST_ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    raise PhastLibraryError(f"{text} (code {rc})")


def ptr(t: torch.Tensor | None) -> int | None:
    """Device pointer of a contiguous CUDA/HIP tensor (None passes a null pointer)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("the HIP path needs tensors on a ROCm device (got a CPU tensor)")
    if not t.is_contiguous():
        raise ValueError("the HIP path needs contiguous tensors")
    return t.data_ptr()


_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_CUR_DEVICE = getattr(torch._C, "_cuda_getDevice", None)


def stream(device: torch.device | None = None) -> int:
    """Raw hipStream_t of torch's current stream (on ``device``, default the current device).
    Read straight from torch's stream registry where torch offers it: this runs before every
    library call, and ``torch.cuda.current_stream`` builds a Stream object after several
    device queries."""
    if _RAW_STREAM is None or _CUR_DEVICE is None:
        return torch.cuda.current_stream(device).cuda_stream
    idx = None if device is None else torch.device(device).index
    return _RAW_STREAM(_CUR_DEVICE() if idx is None else idx)


def require_device() -> torch.device:
    """The ROCm device the product runs on; raise when none is visible."""
    if not torch.cuda.is_available():
        raise PhastLibraryError("no ROCm GPU is visible: the HIP path has no CPU fallback")
    lib()
    return torch.device("cuda", torch.cuda.current_device())
