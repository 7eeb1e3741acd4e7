"""ctypes binding of libsuetes_sisl.so (the C ABI declared in include/suetes_sisl.h).

The library is loaded lazily on first use; a missing library or a call on a
non-HIP tensor raises immediately.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsuetes_sisl.so")

ERRORS = {-1: "bad argument", -2: "HIP launch failed", -3: "non-finite residual",
          -4: "input exceeds the fixed scratch of an allocation-free entry point"}

EXPORTED = [
    "sisl_krylov_size", "sisl_tricubic_f64", "sisl_tricubic_f32", "sisl_scan_blocks",
    "sisl_tricubic_adj_plan_f64", "sisl_tricubic_adj_plan_f32", "sisl_tricubic_adj_f64",
    "sisl_tricubic_adj_f32", "sisl_departure_f64", "sisl_departure_f32",
    "sisl_helmholtz_apply_f64", "sisl_helmholtz_apply_f32", "sisl_helmholtz_apply_T_f64",
    "sisl_helmholtz_apply_T_f32", "sisl_precond_apply_f64", "sisl_precond_apply_f32",
    "sisl_precond_apply_T_f64", "sisl_precond_apply_T_f32", "sisl_gmres_workspace_bytes_f64",
    "sisl_gmres_workspace_bytes_mixed", "sisl_gmres_f64", "sisl_gmres_mixed", "sisl_dot_f64",
    "sisl_timer_enable", "sisl_timer_read", "sisl_autotune_operator", "sisl_autotune_scratch_bytes",
    "sisl_set_operator_shape", "sisl_tricubic_batched_f64", "sisl_tricubic_batched_f32",
    "sisl_departure_batched_f64", "sisl_departure_batched_f32", "sisl_gmres_batched_workspace_bytes",
    "sisl_gmres_batched_setup", "sisl_gmres_batched", "sisl_stagger_f64", "sisl_stagger_f32",
    "sisl_stagger_T_f64", "sisl_stagger_T_f32", "sisl_tricubic_adj_frac_f64",
    "sisl_tricubic_adj_frac_f32", "sisl_tricubic_adj_rows_f64", "sisl_tricubic_adj_rows_f32",
    "sisl_ffsl_1d_f64", "sisl_ffsl_1d_f32", "sisl_ffsl_1d_T_f64", "sisl_ffsl_1d_T_f32",
    "sisl_set_graphs", "sisl_ffsl_1d_f64_noalloc", "sisl_hyperdiffusion_workspace_f64",
    "sisl_hyperdiffusion_f64", "sisl_explicit_f64", "sisl_predictor_f64", "sisl_final_f64",
    "sisl_acoustic_workspace_f64", "sisl_acoustic_loop_f64", "sisl_slow_tendencies_f64",
    "sisl_tricubic_limited_adj_split_f64", "sisl_tricubic_limited_adj_split_f32",
    "sisl_tricubic_limited_adj_corners_f64", "sisl_tricubic_limited_adj_corners_f32",
    "sisl_helmholtz_apply_fused_f32", "sisl_helmholtz_apply_h16_f32", "sisl_probe_stream", "sisl_probe_operator_streams_f32",
    "sisl_tricubic2_f64", "sisl_tricubic2_f32", "sisl_tricubic_mixed", "sisl_tricubic2_mixed",
    "sisl_gmres_batched_outcome", "sisl_get_operator_shape", "sisl_probe_operator_streams_h16_f32",
    "sisl_explicit_adj_workspace_f64", "sisl_explicit_adj_f64", "sisl_predictor_adj_workspace_f64",
    "sisl_predictor_adj_f64", "sisl_final_adj_f64",
]

# timed kernels (enum SislTimedKernel)
TIMED = {"helmholtz": 0, "precond": 1, "krylov_dots": 2, "krylov_update": 3,
         "krylov_normalize": 4, "tricubic": 5, "helmholtz_T": 6, "precond_T": 7, "tricubic_adj": 8,
         "helmholtz_h16": 9, "helmholtz_T_h16": 10}


class SislGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("periodic_x", C.c_int32), ("periodic_y", C.c_int32)]


OPERATOR_FIELDS = ["cua", "zxu", "cx", "cva", "zev", "cy", "dw", "cw", "cz", "kx", "ky", "rdzw",
                   "cpm", "mudx", "mvdy", "afu", "afv", "mmx", "mmy", "idw", "gw", "lpi", "lower",
                   "invden", "cprime", "dzw0", "dzwn"]


class SislOperator(C.Structure):
    _fields_ = [("g", SislGrid), ("alpha", C.c_double)] + [(f, C.c_void_p) for f in OPERATOR_FIELDS]


class SislGmresParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("atol", C.c_double), ("restart", C.c_int32),
                ("maxiter", C.c_int32), ("transpose", C.c_int32), ("early_exit", C.c_int32),
                ("inner_rtol", C.c_double), ("max_outer", C.c_int32), ("basis_fp16", C.c_int32),
                ("coef_bf16", C.c_int32), ("inner_single_cgs", C.c_int32),
                ("inner_basis_operand", C.c_int32), ("op_refine", C.c_void_p)]


class SislGmresStats(C.Structure):
    _fields_ = [("cycles", C.c_int32), ("arnoldi_steps", C.c_int32), ("residual", C.c_double),
                ("ptol", C.c_double), ("b_norm", C.c_double)]


_lib = None


def lib():
    """Load the library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                               "(python -m suetes_amd._build) first")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        sig = {
            "sisl_krylov_size": (i64, [C.POINTER(SislGrid)]),
            "sisl_scan_blocks": (i32, [i64]),
            "sisl_gmres_workspace_bytes_f64": (i64, [C.POINTER(SislGrid), i32]),
            "sisl_gmres_workspace_bytes_mixed": (i64, [C.POINTER(SislGrid), i32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        for p in ("f64", "f32"):
            getattr(L, f"sisl_tricubic_{p}").argtypes = [i32, i32, i32, vp, vp, vp, i32, vp]
            getattr(L, f"sisl_tricubic_adj_plan_{p}").argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp]
            getattr(L, f"sisl_tricubic_adj_{p}").argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, vp]
            getattr(L, f"sisl_tricubic_adj_frac_{p}").argtypes = [i32, i32, i32, vp, vp, vp, vp]
            getattr(L, f"sisl_tricubic_adj_rows_{p}").argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp]
            getattr(L, f"sisl_tricubic_limited_adj_split_{p}").argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, vp]
            getattr(L, f"sisl_tricubic_limited_adj_corners_{p}").argtypes = [i32, i32, i32, vp, vp, vp, vp, vp]
            getattr(L, f"sisl_departure_{p}").argtypes = [C.POINTER(SislGrid), i32, vp, vp, vp, vp,
                                                          d, d, d, i32, vp, vp, vp]
            getattr(L, f"sisl_helmholtz_apply_{p}").argtypes = [C.POINTER(SislOperator), vp, vp, vp]
            getattr(L, f"sisl_helmholtz_apply_T_{p}").argtypes = [C.POINTER(SislOperator), vp, vp, vp, vp]
            getattr(L, f"sisl_precond_apply_{p}").argtypes = [C.POINTER(SislOperator), vp, vp, vp]
            getattr(L, f"sisl_precond_apply_T_{p}").argtypes = [C.POINTER(SislOperator), vp, vp, vp, vp]
        L.sisl_gmres_f64.argtypes = [C.POINTER(SislOperator), C.POINTER(SislGmresParams), vp, vp, vp,
                                     C.POINTER(SislGmresStats), vp]
        L.sisl_gmres_mixed.argtypes = [C.POINTER(SislOperator), C.POINTER(SislOperator),
                                       C.POINTER(SislGmresParams), vp, vp, vp,
                                       C.POINTER(SislGmresStats), vp]
        L.sisl_dot_f64.argtypes = [vp, vp, i64, vp, vp, vp]
        L.sisl_autotune_operator.argtypes = [C.POINTER(SislOperator), i32, vp, vp, vp, C.POINTER(i32),
                                             C.POINTER(i32), vp]
        L.sisl_autotune_scratch_bytes.restype = i64
        L.sisl_autotune_scratch_bytes.argtypes = [C.POINTER(SislGrid), i32]
        L.sisl_set_operator_shape.argtypes = [C.POINTER(SislGrid), i32, i32, i32]
        L.sisl_get_operator_shape.argtypes = [C.POINTER(SislGrid), i32, i32, C.POINTER(i32), C.POINTER(i32)]
        pp = C.POINTER(vp)
        for p in ("f64", "f32"):
            getattr(L, f"sisl_tricubic_batched_{p}").argtypes = [i32, C.POINTER(i32), pp, pp, pp, i32, vp]
            getattr(L, f"sisl_departure_batched_{p}").argtypes = [i32, C.POINTER(SislGrid), i32, pp, pp, pp, pp,
                                                                  C.POINTER(d), C.POINTER(d), d, i32, pp, pp, vp]
        L.sisl_gmres_batched_workspace_bytes.restype = i64
        L.sisl_gmres_batched_workspace_bytes.argtypes = [C.POINTER(SislGrid), i32, i32]
        L.sisl_gmres_batched_setup.argtypes = [C.POINTER(SislOperator), C.POINTER(SislOperator), i32, i32, vp, vp]
        L.sisl_gmres_batched.argtypes = [i32, C.POINTER(SislGmresParams), pp, pp, vp, C.POINTER(SislGmresStats), vp]
        L.sisl_gmres_batched_outcome.argtypes = [i32, vp, C.POINTER(SislGmresStats), vp]
        for p in ("f64", "f32"):
            getattr(L, f"sisl_stagger_{p}").argtypes = [vp, vp, i64, i32, i64, i32, i32, d, vp, i64, vp]
            getattr(L, f"sisl_stagger_T_{p}").argtypes = [vp, vp, i64, i32, i64, i32, i32, i32, d, vp, i64, vp]
        for p in ("f64", "f32"):
            getattr(L, f"sisl_ffsl_1d_{p}").argtypes = [vp, vp, vp, vp, i64, i32, i64, i32, vp]
            getattr(L, f"sisl_ffsl_1d_T_{p}").argtypes = [vp, vp, vp, vp, vp, vp, i64, i32, i64, i32, vp]
        L.sisl_ffsl_1d_f64_noalloc.argtypes = [vp, vp, vp, i64, i32, i64, i32, vp]
        for name in ("sisl_tricubic2_f64", "sisl_tricubic2_f32", "sisl_tricubic2_mixed"):
            getattr(L, name).argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, i32, vp]
        L.sisl_tricubic_mixed.argtypes = [i32, i32, i32, vp, vp, vp, i32, vp]
        L.sisl_helmholtz_apply_fused_f32.argtypes = [C.POINTER(SislOperator), vp, vp, vp, vp, d, i32, i32, vp]
        L.sisl_helmholtz_apply_h16_f32.argtypes = [C.POINTER(SislOperator), vp, vp, vp, i32, i32, vp]
        L.sisl_probe_stream.argtypes = [vp, vp, i64, i32, vp]
        L.sisl_probe_operator_streams_f32.argtypes = [C.POINTER(SislOperator), vp, vp, vp, i32, vp]
        L.sisl_probe_operator_streams_h16_f32.argtypes = [C.POINTER(SislOperator), vp, vp, i32, vp]
        L.sisl_set_graphs.argtypes = [i32]
        L.sisl_timer_enable.argtypes = [i32]
        L.sisl_timer_read.argtypes = [i32, i32, C.POINTER(d), C.POINTER(i64), C.POINTER(i64),
                                      C.POINTER(d)]
        _lib = L
    return _lib


def timer_enable(on: bool):
    check(lib().sisl_timer_enable(int(bool(on))), "sisl_timer_enable")


def timer_read(kernel: str, fp32: bool):
    """(total_ms, launches, working_launches, algorithmic_bytes) of one timed
    kernel since the last timer_enable(True)."""
    ms, n, nw, by = C.c_double(), C.c_int64(), C.c_int64(), C.c_double()
    check(lib().sisl_timer_read(TIMED[kernel], int(bool(fp32)), C.byref(ms), C.byref(n), C.byref(nw),
                                C.byref(by)), "sisl_timer_read")
    return ms.value, n.value, nw.value, by.value


def ptr_array(tensors):
    """ctypes array of the tensors' device pointers."""
    return (C.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])


def check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {ERRORS.get(rc, rc)}")


def stream_ptr(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_device(*tensors):
    """The kernels run only on HIP device memory."""
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise RuntimeError("suetes_amd kernels need tensors on a HIP device (cuda:N); "
                               "there is no CPU path")
        if not t.is_contiguous():
            raise RuntimeError("suetes_amd kernels need contiguous tensors in the internal layout")


def ptr(t) -> int:
    return t.data_ptr() if t is not None else 0
