"""ctypes declarations of libdtc_amd.so (include/dtc_capi.h).

The library is the HIP product; if it is missing this module raises instead
of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libdtc_amd.so")
HEADER = os.path.join(os.path.dirname(PKG_DIR), "include", "dtc_capi.h")

DTC_OK = 0
DTC_ERR_EMPTY_MESH = 1
DTC_ERR_FEW_CONSTRAINTS = 2
DTC_ERR_OOM = 3
DTC_ERR_NO_CONVERGENCE = 4
DTC_ERR_BAD_ARGUMENT = 5
DTC_ERR_DEVICE = 6
DTC_ERR_UNSUPPORTED = 7
DTC_FP64 = 0
DTC_FP32_REFINE = 1
DTC_PC_JACOBI = 0
DTC_PC_MG = 1
DTC_KSTAT_MV = 0
DTC_KSTAT_CHEB = 1
DTC_KSTAT_BRES = 2
DTC_KSTAT_RESID = 3
KSTAT_NAMES = {DTC_KSTAT_MV: "mv", DTC_KSTAT_CHEB: "cheb", DTC_KSTAT_BRES: "bres", DTC_KSTAT_RESID: "resid"}

I64 = C.c_int64
P = C.c_void_p
D3 = C.c_double * 3
I3 = C.c_int64 * 3


class DomainDesc(C.Structure):
    _fields_ = [
        ("phi", P), ("phi_kind", C.c_int), ("micro_nen", C.c_int), ("macro_order", C.c_int), ("_pad", C.c_int),
        ("n", I3), ("llimit", I64), ("xa_n", I3), ("delta", D3), ("size_domain", C.c_double),
        ("E", C.c_double), ("nu", C.c_double), ("rve_strain", C.c_double), ("modulus", P),
    ]


class SolveOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("rotation", C.c_int), ("rel_tol", C.c_double), ("inner_tol", C.c_double),
        ("max_iter", I64), ("rot_step_deg", C.c_double), ("num_zero", C.c_double),
        ("precond", C.c_int), ("mg_degree", C.c_int),
    ]


class DomainResult(C.Structure):
    _fields_ = [
        ("status", C.c_int), ("precision_used", C.c_int), ("n_elems", I64), ("n_nodes", I64), ("n_bnodes", I64),
        ("iterations", I64), ("outer_iterations", I64), ("relres_max", C.c_double),
        ("ff", C.c_double * 3600), ("K_macro", C.c_double * 3600), ("C", C.c_double * 36), ("EE", C.c_double * 36),
        ("density", C.c_double), ("rot_alpha", C.c_double * 2), ("rot_n", C.c_double * 6),
        ("rot_aa", C.c_double * 18), ("rot_EE", C.c_double * 72), ("rot_jj", I64 * 2), ("rot_crit", C.c_double * 2),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "dtc_last_error": (C.c_char_p, []),
    "dtc_version": (C.c_int, []),
    "dtc_ddc_params": (C.c_int, [P, P, P, P, P, P]),
    "dtc_domain_bounds": (C.c_int, [I64, P, P, P, P, P]),
    "dtc_extract_block": (C.c_int, [P, C.c_int, P, P, P, P]),
    "dtc_mesh_count": (C.c_int, [P, C.c_int, P, I64, C.c_int, P, P]),
    "dtc_mesh_build": (C.c_int, [P, C.c_int, P, I64, C.c_int, P, P, I64, I64, P, P, P, P]),
    "dtc_boundaries": (C.c_int, [P, I64, P, P, P, C.c_int, P, P, P]),
    "dtc_element_stiffness": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, P]),
    "dtc_solve_domain": (C.c_int, [P, P, P, P, P]),
    "dtc_solve_batch": (C.c_int, [P, I64, P, P]),
    "dtc_batch_create": (C.c_int, [P, I64, P, P]),
    "dtc_batch_run": (C.c_int, [P, P]),
    "dtc_batch_grad": (C.c_int, [P, I64, P, P]),
    "dtc_batch_stats": (C.c_int, [P, P, P, P, P, P]),
    "dtc_batch_timing": (C.c_int, [P, C.c_int]),
    "dtc_batch_kernel_stats": (C.c_int, [P, C.c_int, P, P, P]),
    "dtc_batch_kernel_flops": (C.c_int, [P, C.c_int, P]),
    "dtc_batch_destroy": (C.c_int, [P]),
    "dtc_rotation_search": (C.c_int, [P, C.c_int, C.c_double, C.c_double, P, P, P, P, P, P]),
    "dtc_rotation_search_noalloc": (C.c_int, [P, C.c_int, C.c_double, C.c_double, P, P, P, P, P, P]),
    "dtc_rotation_noalloc_max_points": (I64, []),
}

_lib = None


class DTCError(RuntimeError):
    def __init__(self, fn: str, rc: int, msg: str):
        super().__init__(f"{fn} failed with code {rc}: {msg}")
        self.rc = rc


def load():
    """Load libdtc_amd.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          f"(make -C m-dtc-direct-tensor-computation_amd)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(fn: str, rc: int, allowed=(DTC_OK,)):
    if rc not in allowed:
        msg = load().dtc_last_error().decode(errors="replace")
        raise DTCError(fn, rc, msg)
    return rc
