"""ctypes binding of libstokes_hip.so (the C ABI of include/stokes_to.h).

The library is built in-tree by stokes_to_amd.build.  Loading fails loudly when it
is missing: there is no CPU fallback for any operation of this package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(HERE, "libstokes_hip.so")

ST_OK, ST_EINVAL, ST_EHIP, ST_ENOCONV, ST_ESINGULAR, ST_ESTATE, ST_ENAN = 0, -1, -2, -3, -4, -5, -6
ST_FP64, ST_FP32, ST_BF16 = 0, 1, 2
ST_INIT_ZERO, ST_INIT_KEEP, ST_INIT_HOST, ST_INIT_LEARNED = 0, 1, 2, 3
ST_OPT_OC, ST_OPT_MMA = 0, 1

ERRNAMES = {ST_EINVAL: "EINVAL", ST_EHIP: "EHIP", ST_ENOCONV: "ENOCONV", ST_ESINGULAR: "ESINGULAR",
            ST_ESTATE: "ESTATE", ST_ENAN: "ENAN"}


class StParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("n", ctypes.c_int),
        ("alpha_max", ctypes.c_double), ("alpha_min", ctypes.c_double), ("q", ctypes.c_double),
        ("power", ctypes.c_double), ("alpha_scale", ctypes.c_double),
        ("precision", ctypes.c_int), ("rtol", ctypes.c_double), ("max_iter", ctypes.c_int),
        ("restart", ctypes.c_int), ("nu_pre", ctypes.c_int), ("nu_post", ctypes.c_int),
        ("omega", ctypes.c_double), ("closed", ctypes.c_int), ("filter_radius", ctypes.c_double),
        ("optimizer", ctypes.c_int), ("fused_sweep", ctypes.c_int), ("warm_order", ctypes.c_int),
    ]


class StStats(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int), ("converged", ctypes.c_int), ("relres", ctypes.c_double),
        ("relres0", ctypes.c_double), ("ms", ctypes.c_double), ("setup_ms", ctypes.c_double),
        ("init_ms", ctypes.c_double), ("learned_used", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every entry point of include/stokes_to.h with its signature
D = ctypes.POINTER(ctypes.c_double)
U8 = ctypes.POINTER(ctypes.c_uint8)
C = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_double
SIGNATURES = {
    "st_default_params": (None, [ctypes.POINTER(StParams), I, I]),
    "st_create": (I, [ctypes.POINTER(C), ctypes.POINTER(StParams)]),
    "st_destroy": (None, [C]),
    "st_last_error": (ctypes.c_char_p, [C]),
    "st_get_params": (I, [C, ctypes.POINTER(StParams)]),
    "st_num_dofs": (ctypes.c_long, [C]),
    "st_num_cells": (ctypes.c_long, [C]),
    "st_set_geometry": (I, [C, U8, I]),
    "st_set_penalty": (I, [C, D]),
    "st_set_design": (I, [C, D]),
    "st_get_design": (I, [C, D]),
    "st_set_design_device": (I, [C, D]),
    "st_profile": (I, [C, I]),
    "st_profile_read": (I, [C, I, D, ctypes.POINTER(ctypes.c_long), D]),
    "st_set_rhs": (I, [C, D, D]),
    "st_solve": (I, [C, I, D, D, ctypes.POINTER(StStats)]),
    "st_get_solution": (I, [C, D]),
    "st_reset_solution": (I, [C]),
    "st_get_gradient": (I, [C, D]),
    "st_net_eval": (I, [C, D, D]),
    "st_apply": (I, [C, D, D]),
    "st_apply_device": (I, [C, C, C]),
    "st_objective": (I, [C, D, D]),
    "st_energy_grad": (I, [C, D, D]),
    "st_objective_noalloc": (I, [I, I, C, C, C, C, C]),
    "st_sensitivity": (I, [C, D, D, D, D]),
    "st_filter": (I, [C, D, D, I]),
    "st_oc_update": (I, [C, D, D, F, F, F, F, D, D]),
    "st_mma_update": (I, [C, D, D, F, D, D, D, ctypes.POINTER(ctypes.c_int)]),
    "st_mma_reset": (I, [C]),
    "st_to_step": (I, [C, I, F, F, F, F, D, D, D, ctypes.POINTER(StStats), ctypes.POINTER(StStats)]),
    "st_record_sample": (I, [C]),
    "st_num_samples": (I, [C]),
    "st_net_train": (I, [C, I, F, D]),
    "st_net_num_params": (ctypes.c_long, [C]),
    "st_net_get_params": (I, [C, ctypes.POINTER(ctypes.c_float)]),
    "st_net_set_params": (I, [C, ctypes.POINTER(ctypes.c_float)]),
    "st_net_grad": (I, [C, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), D,
                        ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "st_net_conv_test": (I, [I, I, I, I, I, I] + [ctypes.POINTER(ctypes.c_float)] * 7),
    "st_mg_levels": (I, [C]),
    "st_mg_level_n": (I, [C, I]),
    "st_mg_vcycle": (I, [C, D, D]),
    "st_mg_level_desc": (I, [C, I, U8, D]),
    "st_mg_smooth": (I, [C, I, I, D, D, D]),
    "st_mg_smooth_mp": (I, [C, I, I, D, D, D]),
    "st_mg_restrict": (I, [C, I, D, D]),
    "st_mg_prolong": (I, [C, I, D, D, D]),
}

_libs = {}


def load(path=None):
    path = path or DEFAULT_LIB
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `python -m stokes_to_amd.build` "
                           "(the HIP library is required, there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = lib
    return lib


class StokesError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRNAMES.get(code, code)}: {msg}")
        self.code = code


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(D)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def u8ptr(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(U8)
