"""ctypes binding of libscihpc_hip.so (the C ABI in include/scihpc_hip.h).

Loading fails loudly when the library is missing: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libscihpc_hip.so")

OK, NOT_CONVERGED, BAD_ARG, HIP_ERROR, FELL_BACK, IO_ERROR = 0, 1, 2, 3, 4, 5
MG_NEUMANN, MG_DIRICHLET = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
BC_CODES = {"neumann": 0, "dirichlet": 1, "periodic": 2, "no_slip": 3, "slip": 4}


class ScihpcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"scihpc status {code}: {msg}")
        self.code = code


class SolveOpts(C.Structure):
    _fields_ = [("rel_tol", C.c_double), ("max_cycles", C.c_int), ("nu1", C.c_int),
                ("nu2", C.c_int), ("precision", C.c_int), ("use_learned", C.c_int),
                ("zero_guess", C.c_int), ("accel", C.c_int), ("learned_correction", C.c_int),
                ("lagged_check", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("learned_accepted", C.c_int),
                ("learned_rejected", C.c_int), ("rel_residual", C.c_double), ("ms", C.c_double),
                ("learned_residual", C.c_double), ("learned_skipped", C.c_int),
                ("correction_accepted", C.c_int), ("correction_rejected", C.c_int),
                ("correction_residual", C.c_double)]


class FpcParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("dt_old", C.c_double), ("dt_old2", C.c_double),
                ("rho12", C.c_double), ("ppe_tol", C.c_double), ("max_cycles", C.c_int), ("cycle", C.c_int)]


class FlowParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("reynolds", C.c_double), ("froude", C.c_double), ("weber", C.c_double),
                ("ls_width", C.c_double), ("density_ratio", C.c_double), ("viscosity_ratio", C.c_double),
                ("positive_ref", C.c_int)]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _nullable(base):
    """ndpointer that also accepts None (passed as NULL)."""
    class _P(base):
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return base.from_param(obj)
    return _P


_fpn = _nullable(_fp)
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i, _d = C.c_int, C.c_double
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)

# name -> argtypes (all return int status unless listed in _RET)
SIGNATURES = {
    "scihpc_mg_create": [_i, _i, _i, _i, _d, _d, _d, _i, C.POINTER(_vp)],
    "scihpc_mg_destroy": [_vp],
    "scihpc_mg_num_levels": [_vp, _pi],
    "scihpc_mg_level_shape": [_vp, _i, _pi, _pi, _pi, _pi],
    "scihpc_mg_set_density": [_vp, _dp],
    "scihpc_mg_set_density_padded": [_vp, _dp],
    "scihpc_mg_set_rhs": [_vp, _i, _dp],
    "scihpc_mg_get_rhs": [_vp, _i, _dp],
    "scihpc_mg_set_sol": [_vp, _i, _dp],
    "scihpc_mg_get_sol": [_vp, _i, _dp],
    "scihpc_mg_get_res": [_vp, _i, _dp],
    "scihpc_mg_set_rhs_device": [_vp, _vp],
    "scihpc_mg_get_sol_device": [_vp, _vp],
    "scihpc_mg_set_nocompat": [_vp, _i, _i],
    "scihpc_mg_residual": [_vp, _i, _pd],
    "scihpc_mg_apply": [_vp, _i, _dp, _dp],
    "scihpc_mg_compat": [_vp, _i, _i],
    "scihpc_mg_relax": [_vp, _i, _i],
    "scihpc_mg_restriction": [_vp, _i],
    "scihpc_mg_prolongation": [_vp, _i],
    "scihpc_mg_coarse_solve": [_vp],
    "scihpc_mg_vcycle_ref": [_vp],
    "scihpc_mg_fullcycle_ref": [_vp],
    "scihpc_mg_vcycle": [_vp, _i, _i],
    "scihpc_mg_solve": [_vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)],
    "scihpc_mg_time_kernel": [_vp, _i, _i, _pd, _pd],
    "scihpc_mg_probe_kernel": [_vp, _i, _i],
    "scihpc_mg_probe_result": [_vp, _pi, _pd, _pd, _pd],
    "scihpc_mg_attach_learned": [_vp, _vp, _i],
    "scihpc_mg_learned_features": [_vp, _i, _fp, _fpn, _fpn],
    "scihpc_unet_create": [_i, _i, _i, C.c_ulonglong, C.POINTER(_vp)],
    "scihpc_unet_create3": [_i, _i, _i, _i, C.c_ulonglong, C.POINTER(_vp)],
    "scihpc_unet_set_loss_smoothing": [_vp, _i, C.c_float],
    "scihpc_unet_destroy": [_vp],
    "scihpc_unet_num_params": [_vp, C.POINTER(C.c_longlong)],
    "scihpc_unet_layer": [_vp, _i, _pi, _pi, _pi, C.POINTER(C.c_longlong)],
    "scihpc_unet_get_params": [_vp, _fp],
    "scihpc_unet_set_params": [_vp, _fp],
    "scihpc_unet_get_grads": [_vp, _fp],
    "scihpc_unet_set_out_scale": [_vp, C.c_float],
    "scihpc_unet_get_out_scale": [_vp, C.POINTER(C.c_float)],
    "scihpc_unet_set_samples": [_vp, _i, _fp, _fp, _fpn],
    "scihpc_unet_forward": [_vp, _i, _ip, _fpn],
    "scihpc_unet_backward": [_vp, _pd],
    "scihpc_unet_train_step": [_vp, _i, _ip, C.c_float, _pd],
    "scihpc_unet_sync": [_vp],
    "scihpc_unet_time": [_vp, _i, _i, _i, _pd, _pd],
    "scihpc_conv3x3_forward": [_u16p, _i, _i, _i, _i, _u16p, _fpn, _i, _i, _u16p],
    "scihpc_conv3x3x3_forward": [_u16p, _i, _i, _i, _i, _i, _u16p, _fpn, _i, _i, _u16p],
    "scihpc_proj_create": [_i, _i, _i, _i, _d, _d, _d, _ip, C.POINTER(_vp)],
    "scihpc_proj_destroy": [_vp],
    "scihpc_proj_fpc_step": [_vp, _vp, C.POINTER(FpcParams)] + [_dp] * 8 + [_pi],
    "scihpc_proj_step": [_vp, _vp, _d, C.POINTER(SolveOpts)] + [_dp] * 5 + [C.POINTER(SolveStats)],
    "scihpc_proj_momentum_step": [_vp, C.POINTER(FlowParams), _dp] + [_dp] * 6 + [_vp],
    "scihpc_proj_set_sources": [_vp, _dp, _dp, _dp],
    "scihpc_proj_get_sources": [_vp, _dp, _dp, _dp],
    "scihpc_proj_node_velocity": [_vp] + [_dp] * 6,
    "scihpc_proj_find_sources": [_vp, C.POINTER(FlowParams), _dp] + [_dp] * 6 + [_vp],
    "scihpc_proj_intermediate_velocity": [_vp, C.POINTER(FlowParams)] + [_dp] * 3,
    "scihpc_proj_time_momentum": [_vp, C.POINTER(FlowParams), _i, _pd],
    "scihpc_ls_create": [_i, _i, _i, _i, _d, _d, _d, _ip, _dp, C.POINTER(_vp)],
    "scihpc_ls_destroy": [_vp],
    "scihpc_ls_set_phi": [_vp, _dp],
    "scihpc_ls_get_phi": [_vp, _dp],
    "scihpc_ls_set_velocity": [_vp, _dp, _dp, _dp],
    "scihpc_ls_weno5_derivatives": [_vp, _dp, _dp, _dp, _dp],
    "scihpc_ls_tvd_rk3": [_vp, _d, _i],
    "scihpc_ls_apply_bc": [_vp, _dp, _i],
    "scihpc_ls_density": [_vp, _d, _d, _i, _dp],
    "scihpc_ls_time_step": [_vp, _d, _i, _pd, _pd],
    "scihpc_ls_stabilize_gradient": [_vp, _dp, _pd],
    "scihpc_ls_set_scheme": [_vp, _i],
    "scihpc_ls_set_line_solver": [_vp, _i],
    "scihpc_ls_compact_derivatives": [_vp, _dp, _i] + [_dp] * 6,
    "scihpc_ls_mass": [_vp, _d, _d, _i, _pd],
    "scihpc_ls_mpls": [_vp, _d, _d, _d, _i, _d, _i],
    "scihpc_ls_find_sign": [_vp, _d],
    "scihpc_ls_find_delta": [_vp, _d],
    "scihpc_ls_set_field": [_vp, _i, _dp],
    "scihpc_ls_get_field": [_vp, _i, _dp],
    "scihpc_ls_redistance": [_vp, _d, _d, _i],
    "scihpc_vtk_create": [_i, _i, _i, _i, _dp, _dp, C.c_char_p, C.POINTER(_vp)],
    "scihpc_vtk_destroy": [_vp],
    "scihpc_vtk_begin": [_vp, C.c_uint],
    "scihpc_vtk_add_scalar": [_vp, C.c_char_p, _vp, _i],
    "scihpc_vtk_add_vector": [_vp, C.c_char_p, _vp, _vp, _vp, _i],
    "scihpc_vtk_close": [_vp],
    "scihpc_vtk_time_pack": [_vp, _i, _i, _pd],
}
_RET = {"scihpc_last_error": C.c_char_p}

_lib = None


def lib():
    """Load libscihpc_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing; build it with `python -m scihpc_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = C.c_int
        L.scihpc_last_error.restype = C.c_char_p
        L.scihpc_last_error.argtypes = []
        _lib = L
    return _lib


def check(rc: int, allow=(OK,)) -> int:
    if rc not in allow:
        msg = lib().scihpc_last_error().decode(errors="replace")
        raise ScihpcError(rc, msg)
    return rc
