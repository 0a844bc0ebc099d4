"""ctypes binding of libslepc_amd.so (the C ABI declared in include/slepc_amd.h).

The shared library is built in-tree (``slepc_amd/libslepc_amd.so``) by
``slepc_amd.build.build()``.  Loading fails loudly when it is missing: there is
no CPU fallback for any operation of this package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SLEPC_AMD_LIBRARY") or os.path.join(_HERE, "libslepc_amd.so")
_LIB = None

FP64, FP32 = 0, 1
LARGEST_MAGNITUDE, SMALLEST_MAGNITUDE, LARGEST_REAL, SMALLEST_REAL = 1, 2, 3, 4
ERROR_ABSOLUTE, ERROR_RELATIVE, ERROR_BACKWARD = 0, 1, 2


class SlepcAmdError(RuntimeError):
    pass


class EpsStats(C.Structure):
    _fields_ = [("matvecs", C.c_int64), ("gs_passes", C.c_int64), ("matvecs32", C.c_int64),
                ("gs_passes32", C.c_int64), ("its", C.c_int), ("its32", C.c_int), ("nconv", C.c_int),
                ("reason", C.c_int), ("seconds", C.c_double), ("seconds32", C.c_double),
                ("seconds_ds", C.c_double), ("bytes_spmv", C.c_double), ("bytes_bv", C.c_double),
                ("bytes_restart", C.c_double), ("true_residuals", C.c_int64),
                ("ksp_its", C.c_int64)]


_I32P = np.ctypeslib.ndpointer(np.int32, flags="C")
_F64P = np.ctypeslib.ndpointer(np.float64, flags="C")
_F32P = np.ctypeslib.ndpointer(np.float32, flags="C")
_VP = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "sa_init": (C.c_int, [C.c_int]),
    "sa_last_error": (C.c_char_p, []),
    "sa_stream": (C.c_void_p, []),
    "sa_synchronize": (C.c_int, []),
    "sa_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64]),
    "sa_free": (C.c_int, [_VP]),
    "sa_memcpy_h2d": (C.c_int, [_VP, _VP, C.c_int64]),
    "sa_memcpy_d2h": (C.c_int, [_VP, _VP, C.c_int64]),
    "sa_timer_start": (C.c_int, []),
    "sa_timer_stop": (C.c_int, [C.POINTER(C.c_double)]),
    "sa_mat_create_csr": (C.c_int, [C.c_int, C.c_int64, _I32P, _I32P, _F64P, C.c_int, C.POINTER(_VP)]),
    "sa_set_spmv_stream_bytes": (C.c_int64, [C.c_int64]),
    "sa_set_concurrent_solves": (C.c_int, [C.c_int]),
    "sa_mat_build_balance": (C.c_int, [_VP, C.c_int, C.c_int, C.c_double, _VP]),
    "sa_mat_destroy": (C.c_int, [_VP]),
    "sa_mat_get_info": (C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "sa_mat_mult": (C.c_int, [_VP, _VP, _VP]),
    "sa_mat_residual_norms": (C.c_int, [_VP, C.c_int, _VP, C.c_int64, _VP, _VP, _VP]),
    "sa_csr_assemble": (C.c_int64, [C.c_int, C.c_int64, _I32P, _I32P, _F64P, _VP, _VP, _VP]),
    "sa_gen_lap2d": (C.c_int64, [C.c_int, C.c_int, C.c_double, C.c_uint64, _VP, _VP, _VP]),
    "sa_gen_convdiff3d": (C.c_int64, [C.c_int, C.c_int, C.c_int, _VP, _VP, _VP]),
    "sa_gen_stencil27": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_uint64, _VP, _VP, _VP]),
    "sa_gen_fv2d": (C.c_int64, [C.c_int, C.c_int, _F64P, C.POINTER(C.c_double), _VP, _VP, _VP]),
    "sa_gen_markov": (C.c_int64, [C.c_int, _VP, _VP, _VP]),
    "sa_fv2d_theta": (None, [C.c_uint64, C.c_int, _F64P]),
    "sa_mat_load_petsc_binary": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int64), _VP, _VP, _VP]),
    "sa_vec_write_petsc_binary": (C.c_int, [C.c_char_p, C.c_int64, _VP, C.c_int]),
    "sa_vec_load_petsc_binary": (C.c_int, [C.c_char_p, C.c_int64, C.POINTER(C.c_int64), _VP]),
    "sa_format_eigenvalues": (C.c_int64, [C.c_int, _VP, _VP, C.c_int, C.c_char_p, _VP, C.c_int64]),
    "sa_rand_uniform": (None, [C.c_uint64, C.c_int64, _F64P]),
    "sa_u01": (C.c_double, [C.c_uint64, C.c_uint64]),
    "sa_sub_seed": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "sa_bv_create": (C.c_int, [C.c_int64, C.c_int, C.c_int, C.POINTER(_VP)]),
    "sa_bv_destroy": (C.c_int, [_VP]),
    "sa_bv_get_info": (C.c_int, [_VP, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int)]),
    "sa_bv_get_array": (C.c_int, [_VP, C.POINTER(_VP)]),
    "sa_bv_set_active_columns": (C.c_int, [_VP, C.c_int, C.c_int]),
    "sa_bv_get_active_columns": (C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sa_bv_set_column": (C.c_int, [_VP, C.c_int, _F64P]),
    "sa_bv_get_column": (C.c_int, [_VP, C.c_int, _F64P]),
    "sa_bv_dotvec": (C.c_int, [_VP, _VP, _VP]),
    "sa_bv_multvec": (C.c_int, [_VP, C.c_double, C.c_double, _VP, _VP]),
    "sa_bv_multinplace": (C.c_int, [_VP, _F64P, C.c_int, C.c_int, C.c_int]),
    "sa_bv_scale_column": (C.c_int, [_VP, C.c_int, C.c_double]),
    "sa_bv_norm_column": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_double)]),
    "sa_bv_norm": (C.c_int, [_VP, C.POINTER(C.c_double)]),
    "sa_bv_copy_column": (C.c_int, [_VP, C.c_int, C.c_int]),
    "sa_bv_mat_mult_column": (C.c_int, [_VP, _VP, C.c_int]),
    "sa_bv_orthonormalize_column": (C.c_int, [_VP, C.c_int, C.c_double, C.POINTER(C.c_double),
                                              C.POINTER(C.c_int)]),
    "sa_bv_get_coefficients": (C.c_int, [_VP, C.c_int, C.c_int, _F64P]),
    "sa_bv_dot": (C.c_int, [_VP, _VP, _VP, C.c_int]),
    "sa_bv_matmult": (C.c_int, [_VP, _VP, _VP]),
    "sa_bv_orthogonalize": (C.c_int, [_VP, C.c_int, _VP, C.c_int]),
    "sa_bv_mat_arnoldi": (C.c_int, [_VP, _VP, C.c_int, C.POINTER(C.c_int), C.c_double,
                                    C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "sa_eps_create": (C.c_int, [C.POINTER(_VP)]),
    "sa_eps_destroy": (C.c_int, [_VP]),
    "sa_eps_set_operators": (C.c_int, [_VP, _VP]),
    "sa_eps_set_operators_fp32": (C.c_int, [_VP, _VP]),
    "sa_eps_set_problem_type": (C.c_int, [_VP, C.c_int]),
    "sa_eps_set_type": (C.c_int, [_VP, C.c_int]),
    "sa_eps_set_which": (C.c_int, [_VP, C.c_int]),
    "sa_eps_set_target": (C.c_int, [_VP, C.c_double]),
    "sa_eps_set_extraction": (C.c_int, [_VP, C.c_int]),
    "sa_eps_set_two_sided": (C.c_int, [_VP, C.c_int]),
    "sa_eps_set_balance": (C.c_int, [_VP, C.c_int, C.c_int, C.c_double]),
    "sa_eps_set_true_residual": (C.c_int, [_VP, C.c_int]),
    "sa_eps_get_true_residual": (C.c_int, [_VP, C.POINTER(C.c_int)]),
    "sa_eps_krylovschur_set_restart": (C.c_int, [_VP, C.c_double]),
    "sa_eps_krylovschur_get_restart": (C.c_int, [_VP, C.POINTER(C.c_double)]),
    "sa_eps_krylovschur_set_locking": (C.c_int, [_VP, C.c_int]),
    "sa_eps_krylovschur_get_locking": (C.c_int, [_VP, C.POINTER(C.c_int)]),
    "sa_eps_arnoldi_set_delayed": (C.c_int, [_VP, C.c_int]),
    "sa_eps_arnoldi_get_delayed": (C.c_int, [_VP, C.POINTER(C.c_int)]),
    "sa_eps_lanczos_set_reorthog": (C.c_int, [_VP, C.c_int]),
    "sa_eps_lanczos_get_reorthog": (C.c_int, [_VP, C.POINTER(C.c_int)]),
    "sa_eps_set_st": (C.c_int, [_VP, C.c_int, C.c_int, C.c_double]),
    "sa_eps_set_st_ksp": (C.c_int, [_VP, C.c_int, C.c_double, C.c_int]),
    "sa_eps_set_st_ksp_type": (C.c_int, [_VP, C.c_int, C.c_int]),
    "sa_eps_get_left_eigenvector": (C.c_int, [_VP, C.c_int, _VP, _VP]),
    "sa_eps_set_dimensions": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int]),
    "sa_eps_set_tolerances": (C.c_int, [_VP, C.c_double, C.c_int]),
    "sa_eps_set_initial_space": (C.c_int, [_VP, C.c_int, _F64P]),
    "sa_eps_set_deflation_space": (C.c_int, [_VP, C.c_int, _F64P]),
    "sa_eps_set_mixed": (C.c_int, [_VP, C.c_int, C.c_double]),
    "sa_eps_solve": (C.c_int, [_VP]),
    "sa_eps_get_converged": (C.c_int, [_VP, C.POINTER(C.c_int)]),
    "sa_eps_get_iteration_number": (C.c_int, [_VP, C.POINTER(C.c_int)]),
    "sa_eps_get_converged_reason": (C.c_int, [_VP, C.POINTER(C.c_int)]),
    "sa_eps_monitor_set": (C.c_int, [_VP, _VP, _VP]),
    "sa_eps_set_orthogonalization": (C.c_int, [_VP, C.c_int, C.c_double]),
    "sa_bv_set_orthogonalization": (C.c_int, [_VP, C.c_int]),
    "sa_bv_orthogonalize_some_column": (C.c_int, [_VP, C.c_int, _VP, C.c_double, _VP, C.POINTER(C.c_double),
                                                  C.POINTER(C.c_int)]),
    "sa_eps_monitor_cancel": (C.c_int, [_VP]),
    "sa_eps_set_stopping_test_function": (C.c_int, [_VP, _VP, _VP]),
    "sa_eps_set_convergence_test": (C.c_int, [_VP, C.c_int]),
    "sa_eps_set_convergence_test_function": (C.c_int, [_VP, _VP, _VP]),
    "sa_eps_get_matrix_norms": (C.c_int, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "sa_eps_set_eigenvalue_comparison": (C.c_int, [_VP, _VP, _VP]),
    "sa_eps_get_eigenpair": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), _VP, _VP]),
    "sa_eps_get_error_estimate": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_double)]),
    "sa_eps_compute_error": (C.c_int, [_VP, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "sa_eps_get_stats": (C.c_int, [_VP, C.POINTER(EpsStats)]),
}


def lib():
    """Loads libslepc_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise SlepcAmdError("libslepc_amd.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        msg = lib().sa_last_error()
        raise SlepcAmdError("slepc_amd error %d: %s" % (rc, msg.decode() if msg else "?"))
    return rc
