"""ctypes binding of libmflops.so (the C ABI declared in include/mflops/mf.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libmflops.so")

MF_OK, MF_ERR_ARG, MF_ERR_SETUP, MF_ERR_DEVICE, MF_ERR_NOCONV = 0, 1, 2, 3, 4
MF_FP64, MF_MIXED, MF_MIXED_BF16X3, MF_MIXED_FP16X2 = 0, 1, 2, 3

# every exported symbol with (restype, argtypes)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class SolveReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("residual", C.c_double), ("residual0", C.c_double)]


SIGNATURES = {
    "mf_last_error": (C.c_char_p, []),
    "mf_version": (C.c_int, []),
    "mf_der_coeffs_1d": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp]),
    "mf_solver_coeffs_1d": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mf_der_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "mf_der_destroy": (C.c_int, [_vp]),
    "mf_der_lines": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_longlong, _vp]),
    "mf_der3d": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "mf_der3d_host": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int]),
    "mf_der3d_batch": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int,
                                 C.c_int, _vp]),
    "mf_solver3d_create": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int,
                                     C.POINTER(_vp)]),
    "mf_solver3d_destroy": (C.c_int, [_vp]),
    "mf_solver3d_solve": (C.c_int, [_vp, _vp, C.c_double, _vp, _vp, _vp, _vp, C.c_int,
                                    C.c_double, C.c_int, C.POINTER(SolveReport), _vp]),
    "mf_solver3d_solve_host": (C.c_int, [_vp, _vp, C.c_double, _vp, _vp, _vp, _vp, C.c_int,
                                         C.c_double, C.c_int, C.POINTER(SolveReport)]),
    "mf_solver3d_dims": (C.c_int, [_vp, _ip, _ip, _ip]),
    "mf_influence_impulses": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_longlong,
                                        C.c_int, _vp, _vp]),
    "mf_md_create": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_double, C.POINTER(_vp)]),
    "mf_md_destroy": (C.c_int, [_vp]),
    "mf_md_info": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mf_md_matvec": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mf_md_solve": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_int, _vp, _vp, _vp]),
    "mf_md_set_guess": (C.c_int, [_vp, C.c_int]),
    "mf_md_write_matrix": (C.c_int, [_vp, C.c_char_p]),
    "mf_md_create_cached": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_char_p,
                                      C.POINTER(C.c_int), C.POINTER(_vp)]),
    "mf_md_create_ex": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_char_p,
                                  _vp, C.POINTER(_vp)]),
    "mf_md_set_null_space": (C.c_int, [_vp, C.c_int]),
    "mf_md_assembly_stats": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mf_tune_infl_gemm": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_double)]),
    "mf_solver3d_profile": (C.c_int, [_vp, C.c_int]),
    "mf_solver3d_profile_read": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int]),
    "mf_solver3d_time_transforms": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, _vp]),
    "mf_ns_create": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _vp, C.c_int, C.c_double, C.c_double,
                               C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "mf_ns_destroy": (C.c_int, [_vp]),
    "mf_ns_set_bc": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
    "mf_ns_set_forcing": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "mf_ns_set_flow_rate": (C.c_int, [_vp, C.c_int, C.c_double, _vp]),
    "mf_ns_set_state": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "mf_ns_step": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, _vp]),
    "mf_ns_get": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mf_ns_info": (C.c_int, [_vp, _dp, _dp, _vp]),
    "mf_ns_set_les": (C.c_int, [_vp, C.c_int, C.c_double, _vp]),
    "mf_ns_set_scheme": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
    "mf_ns_set_sgs_coupling": (C.c_int, [_vp, C.c_int]),
    "mf_der_lines_host": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_longlong]),
    "mf_sgs_eddy_targets": (C.c_int, [_vp, C.c_longlong, C.c_double, _vp, _vp]),
    "mf_ns_subiterations": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mf_ns_restart_write": (C.c_int, [_vp, C.c_char_p, _vp, C.c_int, _vp]),
    "mf_ns_restart_read": (C.c_int, [_vp, C.c_char_p, _vp]),
    "mf_io_write_field": (C.c_int, [C.c_char_p, C.c_char_p, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                    C.c_int]),
    "mf_io_field_info": (C.c_int, [C.c_char_p, C.c_char_p, _vp, _vp, C.c_int, _vp]),
    "mf_io_read_field": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, _vp]),
    "mf_ns_les_nu": (C.c_int, [_vp, _vp, _vp]),
    "mf_filter_create": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int, C.c_int,
                                   C.c_int, C.POINTER(_vp)]),
    "mf_filter_destroy": (C.c_int, [_vp]),
    "mf_filter_matrix": (C.c_int, [_vp, _vp]),
    "mf_filter3d": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mf_filter3d_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "mf_les_nu": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int, C.c_int, C.c_double,
                            _vp, _vp]),
    "mf_mlp_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "mf_mlp_destroy": (C.c_int, [_vp]),
    "mf_mlp_params": (C.c_int, [_vp, _vp]),
    "mf_mlp_forward": (C.c_int, [_vp, _vp, _vp, C.c_longlong, _vp]),
    "mf_mlp_loss_grad": (C.c_int, [_vp, _vp, _vp, C.c_longlong, _vp, _vp, _vp]),
    "mf_mlp_adam_step": (C.c_int, [_vp, _vp, _vp, C.c_longlong, C.c_float, _vp, _vp]),
    "mf_sgs_targets": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double,
                                 _vp, _vp]),
    "mf_sgs_features": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                  _vp, _vp, C.c_double, C.c_double, _vp, _vp, _vp]),
    "mf_ns_set_sgs": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_double, _vp]),
}

PROF_CLASSES = ("fold", "transform_fp64", "transform_fp32", "residual_fp64",
                "multiply_through", "faces")

_lib = None


class MFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("mflops error %d: %s" % (code, msg))
        self.code = code


def load() -> C.CDLL:
    """Load the HIP library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmflops.so not built: run `python mflops3d-md_amd/build.py` "
                           "(expected at %s)" % LIB_PATH)
    # torch (which holds the device fields) brings a HIP runtime of its own:
    # it has to initialise before this library's first HIP call, which would
    # otherwise leave torch without a device
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != MF_OK:
        msg = load().mf_last_error()
        raise MFError(code, msg.decode() if msg else "")
