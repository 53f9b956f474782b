"""ctypes binding of libg2o_hip.so (include/g2o_hip.h).

The shared library is built in-tree (g2o_amd/csrc/libg2o_hip.so) by
``__graft_entry__.build()`` / ``make -C g2o_amd/csrc``. There is no fallback:
if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as ct
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "csrc", "libg2o_hip.so")

OK, ERR_INVALID, ERR_NOT_PD, ERR_NAN, ERR_OOM, ERR_HIP, ERR_UNSUPPORTED, ERR_NO_DEVICE, ERR_IO = range(9)
FP64, FP32 = 0, 1
SCHUR_EXPLICIT, SCHUR_IMPLICIT = 0, 1
LEVENBERG, GAUSS_NEWTON, DOGLEG = 0, 1, 2

# exported symbols, in the order include/g2o_hip.h declares them
EXPORTS = [
    "g2ohip_default_options", "g2ohip_create", "g2ohip_destroy", "g2ohip_last_error", "g2ohip_set_options",
    "g2ohip_get_options", "g2ohip_ba_set_structure", "g2ohip_ba_structure_info", "g2ohip_ba_set_measurements",
    "g2ohip_ba_set_information",
    "g2ohip_ba_set_state", "g2ohip_ba_get_state", "g2ohip_compute_active_errors", "g2ohip_get_errors",
    "g2ohip_build_system", "g2ohip_set_lambda", "g2ohip_restore_diagonal", "g2ohip_solve", "g2ohip_get_x",
    "g2ohip_get_b", "g2ohip_multiply_hessian", "g2ohip_update", "g2ohip_push", "g2ohip_pop", "g2ohip_discard_top", "g2ohip_lambda_init",
    "g2ohip_compute_scale", "g2ohip_optimize", "g2ohip_set_algorithm", "g2ohip_set_force_stop_flag", "g2ohip_compute_marginals", "g2ohip_ba_get_hessian",
    "g2ohip_ba_get_schur", "g2ohip_last_solve_info", "g2ohip_lm_trials", "g2ohip_set_profiling",
    "g2ohip_get_profile",
    "g2ohip_linear_solve_pcg", "g2ohip_pg_set_structure", "g2ohip_pg_structure_info", "g2ohip_pg_set_measurements",
    "g2ohip_pg_set_state", "g2ohip_pg_get_state", "g2ohip_pg_get_system", "g2ohip_bal_set_structure",
    "g2ohip_bal_structure_info", "g2ohip_bal_set_measurements", "g2ohip_bal_set_state", "g2ohip_bal_get_state",
    "g2ohip_bal_get_hessian", "g2ohip_bal_get_schur", "g2ohip_graph_create",
    "g2ohip_graph_destroy", "g2ohip_graph_last_error", "g2ohip_graph_log", "g2ohip_graph_load", "g2ohip_graph_info",
    "g2ohip_graph_get_ba", "g2ohip_graph_get_ba_params", "g2ohip_graph_get_pg", "g2ohip_graph_set_ba_state", "g2ohip_graph_set_pg_state",
    "g2ohip_graph_add_ba", "g2ohip_graph_add_pg", "g2ohip_graph_save", "g2ohip_graph_load_bal", "g2ohip_graph_get_bal",
    "g2ohip_ba_adjoint", "g2ohip_ba_newton", "g2ohip_linear_solve_pcg_device", "g2ohip_replay_pcg",
]
K_LINEARIZE, K_SCHUR, K_PCG, K_REFINE, K_CHI2, K_UPDATE, K_SCHUR_WINDOWS, K_PCG_SPMV = range(8)
KERNEL_NAMES = ["linearize", "schur", "pcg", "refine", "chi2", "update", "schur_windows", "pcg_spmv"]


class Options(ct.Structure):
    _fields_ = [("precision", ct.c_int), ("schur_mode", ct.c_int), ("pcg_tolerance", ct.c_double),
                ("pcg_max_iter", ct.c_int), ("refine_steps", ct.c_int), ("max_trials", ct.c_int),
                ("user_lambda_init", ct.c_double), ("tau", ct.c_double)]


class IterStats(ct.Structure):
    _fields_ = [("chi2_before", ct.c_double), ("chi2", ct.c_double), ("lambda_", ct.c_double),
                ("levenberg_iterations", ct.c_int), ("iterations_linear_solver", ct.c_int),
                ("refine_steps", ct.c_int), ("status", ct.c_int), ("time_residuals", ct.c_double),
                ("time_quadratic_form", ct.c_double), ("time_schur_complement", ct.c_double),
                ("time_linear_solver", ct.c_double), ("time_update", ct.c_double),
                ("time_iteration", ct.c_double), ("fp64_fallback", ct.c_int)]


_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)
_up = ct.POINTER(ct.c_ubyte)
_vp = ct.c_void_p

_lib = None


def lib():
    """Load the library (once). Raises OSError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with `make -C g2o_amd/csrc` "
                      "or __graft_entry__.build()")
    L = ct.CDLL(LIB_PATH)
    sig = {
        "g2ohip_default_options": (None, [ct.POINTER(Options)]),
        "g2ohip_create": (ct.c_int, [ct.c_int, ct.POINTER(_vp)]),
        "g2ohip_destroy": (None, [_vp]),
        "g2ohip_last_error": (ct.c_char_p, [_vp]),
        "g2ohip_set_options": (ct.c_int, [_vp, ct.POINTER(Options)]),
        "g2ohip_get_options": (ct.c_int, [_vp, ct.POINTER(Options)]),
        "g2ohip_ba_set_structure": (ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int, _ip, _ip, _up, _up]),
        "g2ohip_ba_structure_info": (ct.c_int, [_vp, ct.POINTER(ct.c_longlong)]),
        "g2ohip_ba_set_measurements": (ct.c_int, [_vp, _dp, _dp, _dp, ct.c_double]),
        "g2ohip_ba_set_information": (ct.c_int, [_vp, _dp]),
        "g2ohip_ba_set_state": (ct.c_int, [_vp, _dp, _dp]),
        "g2ohip_ba_get_state": (ct.c_int, [_vp, _dp, _dp]),
        "g2ohip_compute_active_errors": (ct.c_int, [_vp, _dp]),
        "g2ohip_get_errors": (ct.c_int, [_vp, _dp]),
        "g2ohip_build_system": (ct.c_int, [_vp]),
        "g2ohip_set_lambda": (ct.c_int, [_vp, ct.c_double]),
        "g2ohip_restore_diagonal": (ct.c_int, [_vp]),
        "g2ohip_solve": (ct.c_int, [_vp, _ip]),
        "g2ohip_get_x": (ct.c_int, [_vp, _dp]),
        "g2ohip_get_b": (ct.c_int, [_vp, _dp]),
        "g2ohip_multiply_hessian": (ct.c_int, [_vp, _dp, _dp]),
        "g2ohip_update": (ct.c_int, [_vp]),
        "g2ohip_push": (ct.c_int, [_vp]),
        "g2ohip_pop": (ct.c_int, [_vp]),
        "g2ohip_discard_top": (ct.c_int, [_vp]),
        "g2ohip_lambda_init": (ct.c_int, [_vp, _dp]),
        "g2ohip_compute_scale": (ct.c_int, [_vp, _dp]),
        "g2ohip_optimize": (ct.c_int, [_vp, ct.c_int, ct.POINTER(IterStats), _ip]),
        "g2ohip_set_algorithm": (ct.c_int, [_vp, ct.c_int]),
        "g2ohip_set_force_stop_flag": (ct.c_int, [_vp, _ip]),
        "g2ohip_compute_marginals": (ct.c_int, [_vp, ct.c_int, _ip, _ip, _dp, ct.c_double]),
        "g2ohip_ba_get_hessian": (ct.c_int, [_vp, _dp, _dp, _dp]),
        "g2ohip_ba_get_schur": (ct.c_int, [_vp, _ip, _ip, _dp, _dp]),
        "g2ohip_last_solve_info": (ct.c_int, [_vp, _dp]),
        "g2ohip_lm_trials": (ct.c_int, [_vp, _dp, ct.c_int, _ip]),
        "g2ohip_set_profiling": (ct.c_int, [_vp, ct.c_int]),
        "g2ohip_get_profile": (ct.c_int, [_vp, ct.c_int, _dp, _ip]),
        "g2ohip_linear_solve_pcg": (ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int, _ip, _ip, _dp, _dp, _dp,
                                               ct.c_double, ct.c_int, _dp, _ip]),
        "g2ohip_replay_pcg": (ct.c_int, [_vp, ct.c_int, _dp, _dp]),
        "g2ohip_linear_solve_pcg_device": (ct.c_int, [_vp, ct.c_int, ct.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                                      ct.c_double, ct.c_int, _dp, _ip]),
        "g2ohip_ba_adjoint": (ct.c_int, [_vp, _dp, _dp, _dp, _dp, ct.c_double, ct.c_int, _ip]),
        "g2ohip_ba_newton": (ct.c_int, [_vp, ct.c_int, ct.c_double, ct.c_double, _ip]),
        "g2ohip_pg_set_structure": (ct.c_int, [_vp, ct.c_int, ct.c_int, _ip, _ip, _up]),
        "g2ohip_pg_structure_info": (ct.c_int, [_vp, ct.POINTER(ct.c_longlong)]),
        "g2ohip_pg_set_measurements": (ct.c_int, [_vp, _dp, _dp]),
        "g2ohip_pg_set_state": (ct.c_int, [_vp, _dp]),
        "g2ohip_pg_get_state": (ct.c_int, [_vp, _dp]),
        "g2ohip_pg_get_system": (ct.c_int, [_vp, _ip, _ip, _dp, _dp]),
        "g2ohip_bal_set_structure": (ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int, _ip, _ip, _up, _up]),
        "g2ohip_bal_structure_info": (ct.c_int, [_vp, ct.POINTER(ct.c_longlong)]),
        "g2ohip_bal_set_measurements": (ct.c_int, [_vp, _dp, _dp, ct.c_double]),
        "g2ohip_bal_set_state": (ct.c_int, [_vp, _dp, _dp]),
        "g2ohip_bal_get_state": (ct.c_int, [_vp, _dp, _dp]),
        "g2ohip_bal_get_hessian": (ct.c_int, [_vp, _dp, _dp, _dp]),
        "g2ohip_bal_get_schur": (ct.c_int, [_vp, _ip, _ip, _dp, _dp]),
        "g2ohip_graph_create": (ct.c_int, [ct.POINTER(_vp)]),
        "g2ohip_graph_destroy": (None, [_vp]),
        "g2ohip_graph_last_error": (ct.c_char_p, [_vp]),
        "g2ohip_graph_log": (ct.c_char_p, [_vp]),
        "g2ohip_graph_load": (ct.c_int, [_vp, ct.c_char_p, ct.c_int]),
        "g2ohip_graph_info": (ct.c_int, [_vp, ct.POINTER(ct.c_longlong)]),
        "g2ohip_graph_get_ba": (ct.c_int, [_vp, _ip, _dp, _up, _ip, _dp, _up, _ip, _ip, _dp, _dp]),
        "g2ohip_graph_get_ba_params": (ct.c_int, [_vp, _ip, _dp, _ip]),
        "g2ohip_graph_get_pg": (ct.c_int, [_vp, _ip, _dp, _up, _ip, _ip, _dp, _dp]),
        "g2ohip_graph_set_ba_state": (ct.c_int, [_vp, _dp, _dp]),
        "g2ohip_graph_set_pg_state": (ct.c_int, [_vp, _dp]),
        "g2ohip_graph_add_ba": (ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int, _ip, _dp, _up, _ip, _dp, _up, _ip, _ip,
                                           _dp, _dp]),
        "g2ohip_graph_add_pg": (ct.c_int, [_vp, ct.c_int, ct.c_int, _ip, _dp, _up, _ip, _ip, _dp, _dp]),
        "g2ohip_graph_save": (ct.c_int, [_vp, ct.c_char_p, ct.c_int]),
        "g2ohip_graph_load_bal": (ct.c_int, [_vp, ct.c_char_p, ct.c_int]),
        "g2ohip_graph_get_bal": (ct.c_int, [_vp, ct.POINTER(ct.c_longlong), _dp, _dp, _ip, _ip, _dp]),
    }
    for name in EXPORTS:
        fn = getattr(L, name)
        fn.restype, fn.argtypes = sig[name]
    _lib = L
    return L
