"""ctypes binding of libheomq.so (the C ABI declared in include/heomq.h).

The library is loaded from this package directory (or the path HEOMQ_LIBRARY
names); a missing library is an error, never a fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_LIB = None
LIB_PATH = Path(__file__).resolve().parent / "libheomq.so"

OK, BAD_DIM, UNSUPPORTED, HIP_ERROR, NOT_CONVERGED, BREAKDOWN, STOPPED, BAD_ARG = 0, 1, 2, 3, 4, 8, 16, 32
TRACE_NONE, TRACE_LAST, TRACE_RDO = 0, 1, 2
HAM_SIAM, HAM_DQD = 0, 1
# doubles every solver writes into its stats array (heomq.h HEOMQ_STATS_*)
STATS_LEN, STATS_MIXED_LEN, STATS_CHEBY_LEN = 8, 64, 3


class HeomqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"heomq error {code}: {msg}")
        self.code = code


class CParams(C.Structure):
    _fields_ = [
        ("norbs", C.c_int32), ("nspin", C.c_int32), ("nalf", C.c_int32), ("nmats", C.c_int32),
        ("ntop", C.c_int32), ("lsdraw", C.c_int32), ("ham", C.c_int32), ("steady", C.c_int32),
        ("lscale", C.c_int32), ("reserved", C.c_int32),
        ("bandwidth", C.c_double * 4), ("linewidth", (C.c_double * 4) * 4),
        ("temperature", C.c_double * 4), ("engyshift", (C.c_double * 2) * 4), ("hpar", C.c_double * 16),
    ]


P = C.c_void_p
D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
I32 = C.POINTER(C.c_int32)
I64 = C.POINTER(C.c_int64)
I8 = C.POINTER(C.c_int8)

SIGNATURES = {
    "heomq_last_error": (C.c_char_p, []),
    "heomq_version": (C.c_int, []),
    "heomq_model_build": (C.c_int, [C.POINTER(CParams), C.POINTER(P)]),
    "heomq_model_free": (None, [P]),
    "heomq_model_info": (C.c_int, [P, I64, I64]),
    "heomq_model_coefs": (C.c_int, [P, D, D, D]),
    "heomq_model_operators": (C.c_int, [P, D, D]),
    "heomq_model_tier": (C.c_int, [P, C.c_int, I32]),
    "heomq_model_coupling": (C.c_int, [P, I32, I32]),
    "heomq_model_set_params": (C.c_int, [P, C.POINTER(CParams)]),
    "heomq_model_packed_layout": (C.c_int, [P, I64]),
    "heomq_model_from_reference": (C.c_int, [
        C.c_int32, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
        C.c_int64, C.c_int64, I32, I64, I64, I64, I64, I8, I8, I8, I64, I32, D, D, D, D, I32, I32, I32, I32, D, D, D,
        D, C.POINTER(P)]),
    "heomq_ctx_create": (C.c_int, [P, C.c_int, C.POINTER(P)]),
    "heomq_ctx_free": (None, [P]),
    "heomq_ctx_refresh": (C.c_int, [P, P]),
    "heomq_ctx_vector_len": (C.c_int64, [P]),
    "heomq_ctx_set_graphs": (C.c_int, [P, C.c_int]),
    "heomq_ctx_autotune": (C.c_int, [P, C.c_int, I32, D]),
    "heomq_ctx_set_precond": (C.c_int, [P, C.c_int]),
    "heomq_ctx_set_precond_shift": (C.c_int, [P, C.c_double]),
    "heomq_ctx_set_shadow_seed": (C.c_int, [P, C.c_uint64]),
    "heomq_apply": (C.c_int, [P, P, P, C.c_int, P]),
    "heomq_apply_transpose": (C.c_int, [P, P, P, C.c_int, P]),
    "heomq_apply_f32": (C.c_int, [P, P, P, C.c_int, P]),
    "heomq_apply_dparam": (C.c_int, [P, P, P, D, D, D, D, P]),
    "heomq_apply_packed": (C.c_int, [P, P, P, C.c_int, C.c_int, P]),
    "heomq_packed_info": (C.c_int, [P, I64]),
    "heomq_apply_packed_precond_noalloc": (C.c_int, [P, P, P, C.c_int, P]),
    "heomq_apply_packed_raw": (C.c_int, [P, P, P, C.c_int, C.c_int, P]),
    "heomq_apply_packed_raw_f32": (C.c_int, [P, P, P, C.c_int, C.c_int, P]),
    "heomq_apply_packed_precond": (C.c_int, [P, P, P, C.c_int, C.c_int, P]),
    "heomq_pack": (C.c_int, [P, P, P, P]),
    "heomq_unpack": (C.c_int, [P, P, P, P]),
    "heomq_calc_ax_tfqmr": (C.c_int, [P, C.c_int32, D, D]),
    "heomq_calc_atrx": (C.c_int, [P, C.c_int32, D, D]),
    "heomq_calcderiv": (C.c_int, [P, D, D]),
    "heomq_apply_bicg": (C.c_int, [P, P, P, P]),
    "heomq_apply_transpose_bicg": (C.c_int, [P, P, P, P]),
    "heomq_calc_ax_bicg": (C.c_int, [P, C.c_int32, D, D]),
    "heomq_calc_atrx_bicg": (C.c_int, [P, C.c_int32, D, D]),
    "heomq_solve_bicg": (C.c_int, [P, P, C.c_double, C.c_int, D, P]),
    "heomq_solve_bicg_host": (C.c_int, [P, D, C.c_double, C.c_int, D]),
    "heomq_spa_pattern": (C.c_int, [P, C.c_int64, I32, I64, I32, I32]),
    "heomq_spa_build_pattern": (C.c_int, [P, C.c_double, C.c_uint64, C.c_int64, I64, I32, I64, I32, I32]),
    "heomq_spa_solve_steady": (C.c_int, [P, P, C.c_double, C.c_int, D, P]),
    "heomq_apply_spa": (C.c_int, [P, P, P, P]),
    "heomq_calc_ax_tfqmr_spa": (C.c_int, [P, C.c_int64, D, D]),
    "heomq_calcderiv_spa_at": (C.c_int, [P, C.c_double, P, P, P]),
    "heomq_calcderiv_spa_tt": (C.c_int, [P, C.c_double, D, D]),
    "heomq_model_jacobi_info": (C.c_int, [P, D, D, D]),
    "heomq_jacobi_info": (C.c_int, [P, D]),
    "heomq_jacobi_step": (C.c_int, [P, P, P, C.c_double, P]),
    "heomq_calc_ax_jacobi": (C.c_int, [P, C.c_int32, D, D]),
    "heomq_solve_jacobi": (C.c_int, [P, P, C.c_double, C.c_int, C.c_double, D, P]),
    "heomq_solve_jacobi_host": (C.c_int, [P, D, C.c_double, C.c_int, D]),
    "heomq_solve_steady": (C.c_int, [P, P, P, C.c_double, C.c_int, C.c_int, C.c_int, D, P]),
    "heomq_solve_steady_mixed": (C.c_int, [P, P, P, C.c_double, C.c_double, C.c_int, C.c_int, D, P]),
    "heomq_solve_steady_host": (C.c_int, [P, D, C.c_double, C.c_int, D]),
    "heomq_batch_create": (C.c_int, [C.POINTER(CParams), C.c_int, C.c_int, C.POINTER(P)]),
    "heomq_batch_free": (None, [P]),
    "heomq_batch_info": (C.c_int, [P, I64]),
    "heomq_batch_apply_packed": (C.c_int, [P, P, P, C.c_int, P]),
    "heomq_batch_solve": (C.c_int, [P, C.c_double, C.c_int, D, D, D, D, P]),
    "heomq_sweep_create": (C.c_int, [C.POINTER(CParams), C.c_int, C.c_int, C.POINTER(P)]),
    "heomq_sweep_free": (None, [P]),
    "heomq_sweep_info": (C.c_int, [P, I64]),
    "heomq_sweep_problems": (C.c_int, [P, I32, I64, I64]),
    "heomq_sweep_apply_packed": (C.c_int, [P, P, P, C.c_int, P]),
    "heomq_sweep_solve": (C.c_int, [P, C.c_double, C.c_int, D, D, D, D, P]),
    "heomq_rk4": (C.c_int, [P, P, C.c_double, C.c_int, P]),
    "heomq_rk4_at": (C.c_int, [P, P, C.c_double, C.c_double, C.c_int, P]),
    "heomq_ctx_set_field": (C.c_int, [P, C.c_int, D, D, C.c_double]),
    "heomq_calcderiv_at": (C.c_int, [P, C.c_double, P, P, P]),
    "heomq_calcderiv_tt": (C.c_int, [P, C.c_double, D, D]),
    "heomq_ctx_set_system_field": (C.c_int, [P, C.c_int, D, I32, D]),
    "heomq_ctx_field_shifts": (C.c_int, [P, C.c_double, D]),
    "heomq_ctx_set_presmooth": (C.c_int, [P, C.c_int]),
    "heomq_ctx_set_precond_fusion": (C.c_int, [P, C.c_int]),
    "heomq_ctx_set_trun_der": (C.c_int, [P, C.c_int]),
    "heomq_ctx_set_apply_form": (C.c_int, [P, C.c_int, C.c_int]),
    "heomq_ctx_get_apply_form": (C.c_int, [P, I32]),
    "heomq_chebyshev": (C.c_int, [P, P, C.c_double, C.c_int, D, P]),
    "heomq_rk4_packed": (C.c_int, [P, P, C.c_double, C.c_double, C.c_int, P]),
    "heomq_chebyshev_packed": (C.c_int, [P, P, C.c_double, C.c_int, D, P]),
    "heomq_observables": (C.c_int, [P, P, D, D, P]),
}


def lib():
    """Load libheomq.so once; raise if it was not built.  HEOMQ_LIBRARY names
    another build of the same library (the host-sanitizer build of the CPU
    checks)."""
    global _LIB
    if _LIB is None:
        path = Path(os.environ.get("HEOMQ_LIBRARY") or LIB_PATH)
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it first (python __graft_entry__.py build)")
        h = C.CDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


def check(rc: int):
    if rc != OK:
        msg = lib().heomq_last_error().decode(errors="replace")
        raise HeomqError(rc, msg)
    return rc
