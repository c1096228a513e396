"""ctypes binding of libbpf_amd.so (the C ABI declared in include/bpf.h).

The library is required: importing the package raises if it is not built,
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbpf_amd.so")

BPF_OK = 0
BPF_E_INVALID, BPF_E_HIP, BPF_E_SINGULAR, BPF_E_NOMEM, BPF_E_UNSUPPORTED, BPF_E_BREAKDOWN = 1, 2, 3, 4, 5, 6
BPF_F64, BPF_F32 = 0, 1
BPF_COL_MAJOR, BPF_ROW_MAJOR = 0, 1
BPF_PREC_ILU0, BPF_PREC_IC0, BPF_PREC_BLOCK_JACOBI = 0, 1, 2
BPF_STATUS_UNKNOWN, BPF_STATUS_SOLVED, BPF_STATUS_TIRED, BPF_STATUS_OVERTIMED = 0, 1, 2, 3
BPF_PRECISION_FP64, BPF_PRECISION_MIXED, BPF_PRECISION_MIXED32 = 0, 1, 2
ORDERINGS = {"AMD": 0, "MDQ": 1, "METIS": 2, "RCM": 3, "NATURAL": 4}


class SolverParams(C.Structure):
    _fields_ = [
        ("memory", C.c_int), ("restart", C.c_int), ("itmax", C.c_int),
        ("atol", C.c_double), ("rtol", C.c_double), ("timemax", C.c_double),
        ("verbose", C.c_int), ("history", C.c_int),
        ("berr_tol", C.c_double), ("precision", C.c_int), ("block_width", C.c_int),
        ("max_refine", C.c_int), ("inner_rtol", C.c_double), ("reorthogonalization", C.c_int),
        ("poly_degree", C.c_int),
        ("bmgs_blocks", C.c_int),
        ("inner_its", C.c_int),
        ("inner_width", C.c_int),
    ]


class SolverStats(C.Structure):
    _fields_ = [
        ("niter", C.c_int), ("solved", C.c_int), ("status", C.c_int), ("refine_steps", C.c_int),
        ("timer", C.c_double), ("nresiduals", C.c_int), ("max_berr", C.c_double),
    ]


class BpfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[bpf error {code}] {msg}")
        self.code = code


class SingularException(BpfError):
    """Raised for zero / non-positive pivots (LinearAlgebra.SingularException)."""


vp = C.c_void_p
ip = C.POINTER(C.c_int32)
dp = C.POINTER(C.c_double)
i = C.c_int
d = C.c_double

# name: (argtypes)
SIGNATURES = {
    "bpf_last_error": [],
    "bpf_version": [],
    "bpf_set_option": [i, C.c_longlong],
    "bpf_get_option": [i, C.POINTER(C.c_longlong)],
    "bpf_device_count": [C.POINTER(C.c_int)],
    "bpf_ctx_create": [i, C.POINTER(vp)],
    "bpf_ctx_destroy": [vp],
    "bpf_ctx_synchronize": [vp],
    "bpf_malloc": [vp, C.POINTER(vp), C.c_size_t],
    "bpf_free": [vp, vp],
    "bpf_memcpy_h2d": [vp, vp, vp, C.c_size_t],
    "bpf_memcpy_d2h": [vp, vp, vp, C.c_size_t],
    "bpf_memcpy_d2d": [vp, vp, vp, C.c_size_t],
    "bpf_timer_start": [vp],
    "bpf_timer_stop": [vp, dp],
    "bpf_profile_enable": [vp, i],
    "bpf_profile_read": [vp, i, C.POINTER(C.c_longlong), dp, dp, dp],
    "bpf_csr_create": [vp, i, i, ip, ip, dp, C.POINTER(vp)],
    "bpf_csr_set_values": [vp, dp],
    "bpf_csr_set_values_device": [vp, vp],
    "bpf_csr_info": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), dp],
    "bpf_csr_destroy": [vp],
    "bpf_spmm": [vp, vp, i, i, i, vp, i, vp, i, d, d],
    "bpf_prec_create": [vp, vp, i, C.POINTER(vp)],
    "bpf_prec_factor": [vp, vp],
    "bpf_prec_apply": [vp, i, i, i, vp, i, vp, i],
    "bpf_prec_get_factor": [vp, dp],
    "bpf_block_jacobi_create": [vp, vp, i, C.POINTER(vp)],
    "bpf_block_jacobi_partition": [vp, ip],
    "bpf_prec_levels": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "bpf_prec_destroy": [vp],
    "bpf_solver_params_default": [C.POINTER(SolverParams)],
    "bpf_block_gmres": [vp, vp, vp, i, vp, i, vp, i, C.POINTER(SolverParams), C.POINTER(SolverStats), dp, i],
    "bpf_block_gmres_x0": [vp, vp, vp, i, vp, i, vp, i, vp, i, C.POINTER(SolverParams), C.POINTER(SolverStats), dp, i],
    "bpf_block_bicgstab": [vp, vp, vp, i, vp, i, vp, i, vp, i, C.POINTER(SolverParams), C.POINTER(SolverStats), dp, i],
    "bpf_block_gmres_ex": [vp, vp, vp, vp, i, i, vp, i, vp, i, vp, i, C.POINTER(SolverParams), C.POINTER(SolverStats), dp, i],
    "bpf_block_bicgstab_t": [vp, vp, vp, i, i, vp, i, vp, i, vp, i, C.POINTER(SolverParams), C.POINTER(SolverStats), dp, i],
    "bpf_backward_error": [vp, vp, i, vp, i, vp, i, dp],
    "bpf_backward_error_noalloc": [vp, vp, i, vp, i, vp, i, vp],
    "bpf_hessenberg_eigenvalues": [dp, i, dp, dp],
    "bpf_householder": [vp, i, i, vp, i, vp, i, ip],
    "bpf_block_hessenberg_lsq": [vp, i, i, dp, i, dp, i, dp, i, dp],
    "bpf_batch_create": [vp, i, ip, ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(dp), ip, C.POINTER(vp)],
    "bpf_batch_solve": [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(SolverParams), C.POINTER(SolverStats), ip],
    "bpf_batch_inner_widths": [vp, ip],
    "bpf_batch_destroy": [vp],
    "bpf_rf_create": [vp, vp, i, d, C.POINTER(vp)],
    "bpf_rf_refactor": [vp, vp],
    "bpf_rf_solve": [vp, i, i, vp, i, vp, i],
    "bpf_rf_solve_refined": [vp, vp, i, vp, i, vp, i, i, d, dp, C.POINTER(C.c_int)],
    "bpf_rf_info": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "bpf_rf_get_factor": [vp, ip, ip, ip, ip, dp],
    "bpf_rf_destroy": [vp],
    "bpf_rf_batch_create": [vp, vp, i, i, d, C.POINTER(vp)],
    "bpf_rf_batch_refactor": [vp, vp, i],
    "bpf_rf_batch_solve": [vp, vp, i, vp, i],
}

_lib = None


def load():
    """Loads libbpf_amd.so and declares every entry point; raises if missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libbpf_amd.so is not built ({LIB_PATH}); run `python build.py` in the package "
            "directory or __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name in ("bpf_last_error", "bpf_version") else C.c_int
    _lib = L
    return L


def check(rc: int):
    if rc != BPF_OK:
        msg = _lib.bpf_last_error().decode(errors="replace")
        if rc == BPF_E_SINGULAR:
            raise SingularException(rc, msg)
        raise BpfError(rc, msg)
