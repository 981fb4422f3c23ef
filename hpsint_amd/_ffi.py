"""ctypes binding of libhpsint_amd.so (the C ABI in include/hpsint_amd.h).

The library is the product: there is no fallback.  Loading fails loudly
when it has not been built (``python -m hpsint_amd.build``).
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
# HPSINT_AMD_LIBRARY selects another build of the same library (measurements)
LIB_PATH = os.environ.get("HPSINT_AMD_LIBRARY", os.path.join(HERE, "libhpsint_amd.so"))
HEADER = os.path.join(os.path.dirname(HERE), "include", "hpsint_amd.h")

HPS_OK = 0
HPS_ERR_INVALID_ARGUMENT = 1
HPS_ERR_INVALID_N_COMP = 2
HPS_ERR_UNSUPPORTED = 3
HPS_ERR_OUT_OF_MEMORY = 4
HPS_ERR_HIP = 5
HPS_ERR_NOT_INITIALIZED = 6
HPS_ERR_NEWTON_NO_CONVERGENCE = 7
HPS_ERR_LINEAR_NO_CONVERGENCE = 8

HPS_MOBILITY_SCALAR = 0
HPS_MOBILITY_TENSORIAL = 1
HPS_PRECON_POINT_JACOBI = 0
HPS_PRECON_NODE_BLOCK_JACOBI = 1
HPS_PRECON_BLOCK_CH_ILU = 2


class HpsParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("fe_degree", ctypes.c_int),
        ("n_cells", ctypes.c_int * 3),
        ("cell_size", ctypes.c_double * 3),
        ("n_components", ctypes.c_int),
        ("mobility", ctypes.c_int),
        ("A", ctypes.c_double),
        ("B", ctypes.c_double),
        ("kappa_c", ctypes.c_double),
        ("kappa_p", ctypes.c_double),
        ("Mvol", ctypes.c_double),
        ("Mvap", ctypes.c_double),
        ("Msurf", ctypes.c_double),
        ("Mgb", ctypes.c_double),
        ("L", ctypes.c_double),
    ]


class HpsNewtonParams(ctypes.Structure):
    _fields_ = [
        ("nl_abs_tol", ctypes.c_double),
        ("nl_rel_tol", ctypes.c_double),
        ("nl_max_iter", ctypes.c_int),
        ("l_abs_tol", ctypes.c_double),
        ("l_rel_tol", ctypes.c_double),
        ("l_max_iter", ctypes.c_int),
        ("l_restart", ctypes.c_int),
        ("precond", ctypes.c_int),
        ("use_damping", ctypes.c_int),
        ("threshold_newton_iter", ctypes.c_int),
        ("threshold_linear_iter", ctypes.c_int),
        ("inner_fp32", ctypes.c_int),
    ]


class HpsNewtonStats(ctypes.Structure):
    _fields_ = [
        ("newton_iterations", ctypes.c_int),
        ("linear_iterations", ctypes.c_int),
        ("residual_evaluations", ctypes.c_int),
        ("precond_updates", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("res0", ctypes.c_double),
        ("res", ctypes.c_double),
    ]


class HpsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hpsint_amd error %d: %s" % (code, msg))
        self.code = code


class InvalidNumberOfComponents(HpsError):
    """ExcInvalidNumberOfComponents (S/instantiation.h:61-68)"""


class NewtonDidNotConverge(HpsError):
    """ExcNewtonDidNotConverge (I/lac/solvers_nonlinear.h:183-207)"""


class LinearSolverNoConvergence(HpsError):
    """SolverControl::NoConvergence raised by GMRES"""


_LIB = None

vp = ctypes.c_void_p
i64 = ctypes.c_int64
dbl = ctypes.c_double
cint = ctypes.c_int
pd = ctypes.POINTER(ctypes.c_double)

_SIGNATURES = {
    "hps_create": [ctypes.POINTER(HpsParams), ctypes.POINTER(vp)],
    "hps_destroy": [vp],
    "hps_set_stream": [vp, vp],
    "hps_sync": [vp],
    "hps_sizes": [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)],
    "hps_set_time_weights": [vp, dbl, dbl, pd, cint],
    "hps_set_linearization": [vp, vp],
    "hps_residual": [vp, cint, vp, vp, ctypes.POINTER(vp)],
    "hps_set_constrained_nodes": [vp, ctypes.POINTER(i64), i64],
    "hps_set_grain_cut_off": [vp, dbl],
    "hps_set_component_mask": [vp, vp],
    "hps_get_component_mask": [vp, vp],
    "hps_vmult_f64": [vp, vp, vp],
    "hps_vmult_f32": [vp, vp, vp],
    "hps_precond_update": [vp, cint],
    "hps_precond_update_f32": [vp, cint],
    "hps_precond_apply_f64": [vp, vp, vp],
    "hps_precond_apply_f32": [vp, vp, vp],
    "hps_precond_get_diagonal": [vp, vp],
    "hps_precond_get_blocks": [vp, vp],
    "hps_dot": [vp, i64, vp, vp, pd],
    "hps_norm2": [vp, i64, vp, pd],
    "hps_block_norms": [vp, vp, pd],
    "hps_axpy": [vp, i64, dbl, vp, vp],
    "hps_sadd": [vp, i64, dbl, dbl, vp, vp],
    "hps_scale": [vp, i64, dbl, vp],
    "hps_copy": [vp, i64, vp, vp],
    "hps_fill": [vp, i64, dbl, vp],
    "hps_multi_dot": [vp, i64, cint, vp, vp, pd],
    "hps_multi_dot_noalloc": [vp, i64, cint, vp, vp, vp],
    "hps_multi_axpy_norm": [vp, i64, cint, vp, pd, vp, pd],
    "hps_convert_f64_to_f32": [vp, i64, vp, vp],
    "hps_norm2_f32": [vp, i64, vp, pd],
    "hps_multi_dot_f32": [vp, i64, cint, vp, vp, pd],
    "hps_multi_axpy_norm_f32": [vp, i64, cint, vp, pd, vp, pd],
    "hps_scale_f32": [vp, i64, dbl, vp],
    "hps_axpy_multi_f32": [vp, i64, cint, pd, vp, vp],
    "hps_axpy_multi": [vp, i64, cint, pd, vp, vp],
    "hps_convert_f32_to_f64": [vp, i64, vp, vp],
    "hps_from_blocks": [vp, vp, ctypes.POINTER(vp)],
    "hps_to_blocks": [vp, ctypes.POINTER(vp), vp],
    "hps_last_kernel_ms": [vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
    "hps_enable_timing": [vp, cint],
    "hps_kernel_ms_mean": [vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                           ctypes.POINTER(cint)],
    "hps_newton_create": [vp, ctypes.POINTER(HpsNewtonParams), ctypes.POINTER(vp)],
    "hps_newton_destroy": [vp],
    "hps_newton_clear": [vp],
    "hps_newton_solve": [vp, vp, ctypes.POINTER(vp), ctypes.POINTER(HpsNewtonStats)],
    "hps_newton_solve_ref": [vp, vp, ctypes.POINTER(vp), dbl, ctypes.POINTER(HpsNewtonStats)],
    "hps_mlp_gemm": [vp, cint, cint, cint, cint, cint, cint, vp, cint, vp, cint, vp, cint, vp,
                     vp, cint],
    "hps_mlp_sum_splits": [vp, i64, cint, vp, vp],
    "hps_mlp_colsum": [vp, cint, cint, vp, vp],
    "hps_mlp_loss": [vp, i64, vp, vp, vp, vp, cint, ctypes.POINTER(ctypes.c_float)],
    "hps_mlp_adam": [vp, cint, cint, vp, vp, vp, vp, ctypes.c_float, ctypes.c_float,
                     ctypes.c_float, ctypes.c_float, cint, vp],
    "hps_mlp_features": [vp, vp, vp, dbl, dbl, dbl, vp, vp, vp],
    "hps_mlp_predict": [vp, vp, vp, vp, vp, vp, vp, vp],
    "hps_mlp_targets": [vp, vp, vp, vp, vp, vp, vp, vp],
}


def header_symbols():
    """Names of the entry points declared in include/hpsint_amd.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(hps_\w+)\s*\(", text, re.M)))


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libhpsint_amd.so is not built (run `python -m hpsint_amd.build`); "
                "there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = cint
        L.hps_last_error.restype = ctypes.c_char_p
        L.hps_last_error.argtypes = []
        L.hps_newton_default_params.restype = None
        L.hps_newton_default_params.argtypes = [ctypes.POINTER(HpsNewtonParams)]
        _LIB = L
    return _LIB


def check(code):
    if code != HPS_OK:
        msg = lib().hps_last_error().decode()
        if code == HPS_ERR_INVALID_N_COMP:
            raise InvalidNumberOfComponents(code, msg)
        if code == HPS_ERR_NEWTON_NO_CONVERGENCE:
            raise NewtonDidNotConverge(code, msg)
        if code == HPS_ERR_LINEAR_NO_CONVERGENCE:
            raise LinearSolverNoConvergence(code, msg)
        raise HpsError(code, msg)
