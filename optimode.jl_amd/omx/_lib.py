"""ctypes binding of libomx.so (the C ABI declared in include/omx.h).

The library is loaded from this package directory.  There is no fallback: if the
shared library is missing or fails to load, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libomx.so")

OMX_OK = 0
OMX_ERR_NO_DEVICE = 1
OMX_ERR_BAD_SHAPE = 2
OMX_ERR_NOT_CONVERGED = 3
OMX_ERR_CHOLESKY = 4
OMX_ERR_INVALID_ARG = 5
OMX_ERR_HIP = 6
OMX_ERR_INTERNAL = 7
OMX_ERR_UNSUPPORTED = 8
OMX_PREC_FP32 = 0
OMX_PREC_FP64 = 1

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class OmxInfo(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_int32),
        ("n_newton", ctypes.c_int32),
        ("lobpcg_iters_fp32", ctypes.c_int32),
        ("lobpcg_iters_fp64", ctypes.c_int32),
        ("n_matvec", ctypes.c_int64),
        ("n_precon", ctypes.c_int64),
        ("max_residual", ctypes.c_double),
        ("device_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class OmxProblem(ctypes.Structure):
    """``omx_problem`` of include/omx.h."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("Dx", ctypes.c_double), ("Dy", ctypes.c_double),
        ("epsinv", ctypes.POINTER(ctypes.c_double)),
        ("omega", ctypes.c_double),
        ("nev", ctypes.c_int32), ("max_eigsolves", ctypes.c_int32), ("maxiter", ctypes.c_int32),
        ("k_tol", ctypes.c_double), ("eig_tol", ctypes.c_double),
        ("kmags_out", ctypes.POINTER(ctypes.c_double)),
        ("evecs_out", ctypes.POINTER(ctypes.c_double)),
        ("info", OmxInfo),
        ("status", ctypes.c_int32),
    ]


# (name, restype, argtypes) for every entry point of include/omx.h
SIGNATURES = [
    ("omx_version", ctypes.c_char_p, []),
    ("omx_device_count", ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    ("omx_status_string", ctypes.c_char_p, [ctypes.c_int]),
    ("omx_create", ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, _dp]),
    ("omx_create_3d", ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]),
    ("omx_period_derivatives", ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
    ("omx_destroy", ctypes.c_int, [_vp]),
    ("omx_set_epsinv", ctypes.c_int, [_vp, _dp]),
    ("omx_last_error", ctypes.c_char_p, [_vp]),
    ("omx_set_option", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_double]),
    ("omx_apply_helmholtz", ctypes.c_int, [_vp, ctypes.c_double, _dp, _dp, ctypes.c_int, ctypes.c_int]),
    ("omx_overlap_pullback", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int]),
    ("omx_mode_pullback", ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp]),
    ("omx_apply_helmholtz_noalloc", ctypes.c_int, [_vp, ctypes.c_double, _dp, _dp, ctypes.c_int, ctypes.c_int]),
    ("omx_autotune", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp]),
    ("omx_apply_preconditioner", ctypes.c_int, [_vp, ctypes.c_double, _dp, _dp, ctypes.c_int, ctypes.c_int]),
    ("omx_hmkh", ctypes.c_int, [_vp, ctypes.c_double, _dp, ctypes.c_int, _dp, ctypes.c_int]),
    ("omx_hmh", ctypes.c_int, [_vp, ctypes.c_double, _dp, _dp, ctypes.c_int, _dp]),
    ("omx_group_index", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp]),
    ("omx_ng_gvd", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("omx_build_modes", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp]),
    ("omx_overlap_matrix", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    ("omx_get_mode_fields", ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    ("omx_solve_omega2", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                        ctypes.POINTER(OmxInfo)]),
    ("omx_solve_k", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, ctypes.POINTER(OmxInfo)]),
    ("omx_get_evecs", ctypes.c_int, [_vp, _dp]),
    ("omx_solve_k_adjoint", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp,
                                           ctypes.POINTER(ctypes.c_uint8), _dp, _dp]),
    ("omx_solve_k_periodic_adjoint", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp,
                                                    ctypes.POINTER(ctypes.c_uint8), _dp, _dp, _dp]),
    ("omx_get_kernel_stats", ctypes.c_int, [_vp, ctypes.c_char_p, _dp, ctypes.POINTER(ctypes.c_int64), _dp]),
    ("omx_solve_k_batch", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(OmxProblem), ctypes.c_int,
                                         ctypes.c_int]),
    ("omx_sliceinv_3x3", ctypes.c_int, [ctypes.c_int, ctypes.c_int64, _dp, _dp]),
    ("omx_sliceinv_3x3_pullback", ctypes.c_int, [ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp]),
    ("omx_warmnet_create", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                          ctypes.POINTER(_vp)]),
    ("omx_warmnet_destroy", ctypes.c_int, [_vp]),
    ("omx_warmnet_num_params", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    ("omx_warmnet_get_params", ctypes.c_int, [_vp, _fp]),
    ("omx_warmnet_set_params", ctypes.c_int, [_vp, _fp]),
    ("omx_warmnet_forward", ctypes.c_int, [_vp, ctypes.c_int, _fp, _fp, _fp]),
    ("omx_warmnet_loss_grad", ctypes.c_int, [_vp, ctypes.c_int, _fp, _fp, _fp, ctypes.c_float, _dp, _fp]),
    ("omx_warmnet_train", ctypes.c_int, [_vp, ctypes.c_int, _fp, _fp, _fp, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_float, ctypes.c_uint64, _dp]),
    ("omx_warmnet_set_precision", ctypes.c_int, [_vp, ctypes.c_int]),
    ("omx_warmnet_train_resume", ctypes.c_int, [_vp, ctypes.c_int, _fp, _fp, _fp, ctypes.c_float, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_float, ctypes.c_uint64, _dp]),
    ("omx_warmnet_guess", ctypes.c_int, [_vp, _vp, ctypes.c_double, _fp, _dp]),
    ("omx_solve_k_learned", ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                           ctypes.POINTER(OmxInfo)]),
    ("omx_smooth_epsinv", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp, _dp, _dp, ctypes.c_int,
                                         _dp]),
]

_lib = None


def lib():
    """Load libomx.so once; raises OSError when it is missing (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"libomx.so not found at {LIB_PATH}; build it with `python optimode.jl_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class OmxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[omx status {status}] {msg}")
        self.status = status


def check(status, handle=None):
    if status != OMX_OK:
        L = lib()
        msg = L.omx_last_error(handle).decode() if handle is not None else L.omx_last_error(None).decode()
        if not msg:
            msg = L.omx_status_string(status).decode()
        raise OmxError(status, msg)


def dptr(a):
    """Pointer to a C-contiguous float64 / complex128 numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] or a.ndim == 1
    return a.ctypes.data_as(_dp)


def as_c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def fptr(a):
    """Pointer to a C-contiguous float32 numpy array (or None)."""
    if a is None:
        return None
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float32 array")
    return a.ctypes.data_as(_fp)
