"""ctypes binding of libhfm.so (declarations in include/hfm.h).

The library is built in-tree next to this file by heatflow-mpm_amd/build.py.  There is no
fallback: if the library is missing or a call fails, a RuntimeError says why.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

LIB_NAME = "libhfm.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

HFM_OK = 0
ERRORS = {-1: "bad argument", -2: "HIP error", -3: "out of device memory", -4: "non-finite value", -5: "bad state"}

# enum hfm_field (include/hfm.h)
FIELD_NAMES = ["T", "T2", "TW", "U", "V", "W", "P", "RHS", "ALPHA_S", "ALPHA_P", "TSUM", "KAPPA",
               "AX", "AY", "AZ", "QX", "QY", "QZ", "KX", "KY", "KZ", "MX", "MY", "MZ", "PDIAG", "HRHS", "HDIAG"]
FIELD = {name: i for i, name in enumerate(FIELD_NAMES)}

PREC_FP32 = 0
PREC_MIXED = 1
FLAG_TIME_KERNELS = 1
FLAG_FUSED_CG = 2
FLAG_MG_PCG = 4
FLAG_MACCORMACK = 8
FLAG_PERSISTENT_CG = 16
KT_PCG_DIR, KT_PCG_UPD, KT_PCG_FUSED, KT_PCG_PERSIST = 0, 1, 2, 3


class Params(C.Structure):
    """struct hfm_params (include/hfm.h)."""
    _fields_ = [("n", C.c_int32), ("buoyancy", C.c_int32), ("cg_max_iter", C.c_int32), ("cg_check_every", C.c_int32),
                ("heat_max_iter", C.c_int32), ("heat_check_every", C.c_int32), ("precision", C.c_int32),
                ("flags", C.c_int32),
                ("h", C.c_double), ("rho", C.c_double), ("t_amb", C.c_double), ("gravity", C.c_double),
                ("alpha_f_min", C.c_double), ("cg_tol", C.c_double), ("heat_tol", C.c_double),
                ("kappa_air", C.c_double), ("kappa_mat", C.c_double), ("vol", C.c_double), ("k_particle", C.c_double),
                ("t_sol", C.c_double), ("t_liq", C.c_double), ("yield_solid", C.c_double), ("mu_solid", C.c_double),
                ("mu_liquid", C.c_double), ("t_max", C.c_double), ("dir_planes", C.c_int32)]


class Stats(C.Structure):
    """struct hfm_stats (include/hfm.h)."""
    _fields_ = [("p_iters", C.c_int32), ("p_iters_exact", C.c_int32), ("h_iters", C.c_int32),
                ("h_iters_exact", C.c_int32), ("p_converged", C.c_int32), ("h_converged", C.c_int32),
                ("substeps", C.c_int64), ("p_rr", C.c_double), ("p_atol", C.c_double), ("h_rr", C.c_double),
                ("h_atol", C.c_double), ("rhs_mean", C.c_double), ("p_iters_total", C.c_int64),
                ("h_iters_total", C.c_int64), ("p_outer", C.c_int32), ("h_outer", C.c_int32),
                ("source_dT_sum", C.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_VP = C.c_void_p

# name -> (restype, argtypes); every symbol declared in include/hfm.h
SIGNATURES = {
    "hfm_scene_create": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(_VP)]),
    "hfm_scene_destroy": (C.c_int, [_VP]),
    "hfm_set_static": (C.c_int, [_VP, _F32P, _F32P, _F32P, _F32P, _F32P, _F32P, _F32P]),
    "hfm_set_fixed": (C.c_int, [_VP, _I32P, _F32P]),
    "hfm_set_source": (C.c_int, [_VP, _I32P, C.c_double, C.c_double]),
    "hfm_set_particles": (C.c_int, [_VP, C.c_int64, _VP, _VP, _VP]),
    "hfm_update_particles": (C.c_int, [_VP, _VP, _VP]),
    "hfm_substep": (C.c_int, [_VP, C.c_double]),
    "hfm_run": (C.c_int, [_VP, C.c_double, C.c_int32, C.POINTER(C.c_double)]),
    "hfm_get_field": (C.c_int, [_VP, C.c_int32, _F32P]),
    "hfm_set_field": (C.c_int, [_VP, C.c_int32, _F32P]),
    "hfm_field_size": (C.c_int64, [_VP, C.c_int32]),
    "hfm_get_cell_velocity": (C.c_int, [_VP, _F32P]),
    "hfm_get_particles": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "hfm_get_particle_bins": (C.c_int, [_VP, _I32P, _I32P, _F32P]),
    "hfm_diagnostics": (C.c_int, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hfm_get_stats": (C.c_int, [_VP, C.POINTER(Stats)]),
    "hfm_kernel_time": (C.c_int, [_VP, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "hfm_kernel_time_reset": (C.c_int, [_VP]),
    "hfm_set_kernel_timing": (C.c_int, [_VP, C.c_int32]),
    "hfm_synchronize": (C.c_int, [_VP]),
    "hfm_scene_bytes": (C.c_int64, [_VP]),
    "hfm_batch_create": (C.c_int, [C.POINTER(_VP), C.c_int32, C.POINTER(_VP)]),
    "hfm_batch_destroy": (C.c_int, [_VP]),
    "hfm_batch_set_scenes": (C.c_int, [_VP, C.POINTER(_VP), C.c_int32]),
    "hfm_batch_substep": (C.c_int, [_VP, C.c_double]),
    "hfm_batch_run": (C.c_int, [_VP, C.c_double, C.c_int32, C.POINTER(C.c_double)]),
    "hfm_batch_synchronize": (C.c_int, [_VP]),
    "hfm_batch_set_kernel_timing": (C.c_int, [_VP, C.c_int32]),
    "hfm_batch_kernel_time": (C.c_int, [_VP, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "hfm_batch_kernel_time_reset": (C.c_int, [_VP]),
    "hfm_device_count": (C.c_int, []),
    "hfm_mesh_inside": (C.c_int, [_VP, C.c_int64, _VP, C.c_int64, C.c_float, _VP, _VP, C.c_int]),
    "hfm_mesh_inside_noalloc": (C.c_int, [_VP, C.c_int64, _VP, C.c_int64, C.c_float, _VP, _VP, C.c_int]),
    "hfm_last_error": (C.c_char_p, []),
}

_lib = None
_lock = threading.Lock()


def load(path: str | None = None) -> C.CDLL:
    """Load libhfm.so once; raise RuntimeError if it is not built."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is not built: run `python heatflow-mpm_amd/build.py` (or __graft_entry__.build()); "
                               "there is no CPU fallback")
        lib = C.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def check(rc: int, what: str) -> None:
    if rc != HFM_OK:
        msg = _lib.hfm_last_error() if _lib is not None else b""
        raise RuntimeError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg.decode(errors='replace') if msg else ''}")


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
