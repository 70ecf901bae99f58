"""ctypes binding of libbiofem_hip.so (include/biofem_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C biofem-ia_amd``).
Loading fails loudly if it is missing: there is no CPU fallback for the force path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libbiofem_hip.so")

BF_OK = 0
BF_ERR_ARG = -1
BF_ERR_DEVICE = -2
BF_ERR_NOMEM = -3
BF_ERR_ELEMENT = -4
BF_ERR_STATE = -5
BF_ERR_UNSUPPORTED = -6
BF_ERR_IO = -7
ERR_NAMES = {
    BF_ERR_ARG: "BF_ERR_ARG",
    BF_ERR_DEVICE: "BF_ERR_DEVICE",
    BF_ERR_NOMEM: "BF_ERR_NOMEM",
    BF_ERR_ELEMENT: "BF_ERR_ELEMENT",
    BF_ERR_STATE: "BF_ERR_STATE",
    BF_ERR_UNSUPPORTED: "BF_ERR_UNSUPPORTED",
    BF_ERR_IO: "BF_ERR_IO",
}
BF_ELEM_CC3333 = 6
BF_ELEM_NONFINITE = 7
FLAG_CC3333 = 1
FLAG_NONFINITE = 2
FLAG_CLAMP = 4
FLAG_FALLBACK = 8
BF_PREC_FP64 = 0
BF_PREC_MIXED = 1
BF_PREC_FP64_FAST = 2
SENTINEL = 1000000
MAX_NEN = 18

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class BfMeshDesc(C.Structure):
    _fields_ = [
        ("n_v", C.c_int), ("n_nodes", C.c_int), ("n_elmt", C.c_int), ("n_elmt_base", C.c_int),
        ("patch_stride", C.c_int), ("patch", _ip), ("ire", _ip), ("val", _ip),
        ("nv1", _ip), ("nv2", _ip), ("nv3", _ip),
        ("x0", _dp), ("y0", _dp), ("z0", _dp), ("dA0", _dp), ("n_fib", _dp),
        ("gamma_scale", _dp), ("activation_delay", _dp),
        ("n_ghost_rules", C.c_int), ("ghost_rules", _ip),
    ]


class BfParams(C.Structure):
    _fields_ = [
        ("n_qp", C.c_int), ("c33_iters", C.c_int), ("cart_fib_act", C.c_int),
        ("manufactured", C.c_int), ("h0", C.c_double), ("mu", C.c_double), ("K", C.c_double),
        ("gamma", C.c_double), ("ramp_type", C.c_int), ("gamma_T", C.c_double),
        ("tau_up", C.c_double), ("tau_down", C.c_double), ("apd", C.c_double),
        ("pulse_t0", C.c_double), ("pulse_sigma", C.c_double),
        ("manufactured_gamma0", C.c_double), ("manufactured_T", C.c_double),
        ("ti", C.c_int), ("tisteps", C.c_int), ("dt", C.c_double),
    ]


class BfVtkFrame(C.Structure):
    _fields_ = [
        ("n_v", C.c_int), ("n_nodes", C.c_int), ("n_elmt", C.c_int),
        ("x", _dp), ("y", _dp), ("z", _dp), ("x0", _dp), ("y0", _dp), ("z0", _dp),
        ("nv1", _ip), ("nv2", _ip), ("nv3", _ip), ("contact", _ip), ("m", _dp), ("IE", _dp),
        ("Fint", _dp), ("Fext", _dp), ("Fdyn", _dp), ("Fcnt", _dp), ("u", _dp), ("xdd", _dp),
        ("dA", _dp), ("dA0", _dp), ("n_fib", _dp), ("gamma", _dp),
        ("muscle_activation", C.c_int), ("manufactured", C.c_int), ("lv_geom_process", C.c_int),
    ]


# name -> (restype, argtypes); every symbol include/biofem_hip.h and biofem_io.h declare
SIGNATURES = {
    "bf_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "bf_ctx_destroy": (C.c_int, [_vp]),
    "bf_mesh_upload": (C.c_int, [_vp, C.POINTER(BfMeshDesc)]),
    "bf_params_set": (C.c_int, [_vp, C.POINTER(BfParams)]),
    "bf_force_eval": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_int, C.c_int]),
    "bf_force_eval_device": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "bf_force_eval_device_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "bf_ghost_update": (C.c_int, [_vp, _dp, _dp, _dp]),
    "bf_ghost_update_device": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "bf_c33_download": (C.c_int, [_vp, _dp]),
    "bf_update_thickness": (C.c_int, [_vp, _dp]),
    "bf_elem_flags_download": (C.c_int, [_vp, C.POINTER(C.c_ubyte)]),
    "bf_last_error": (C.c_int, [_vp, _ip, _ip, _ip]),
    "bf_last_fallback_count": (C.c_int, [_vp, _ip]),
    "bf_last_handed_qps": (C.c_int, [_vp, C.POINTER(C.c_ubyte), _ip]),
    "bf_set_fallback": (C.c_int, [_vp, C.c_int]),
    "bf_last_timing": (C.c_int, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "bf_timing_enable": (C.c_int, [_vp, C.c_int]),
    "bf_timing_read": (C.c_int, [_vp, C.POINTER(C.c_longlong), _dp, _dp]),
    "bf_timing_read_split": (C.c_int, [_vp, C.POINTER(C.c_longlong), _dp, _dp, _dp]),
    "bf_timing_read_solve": (C.c_int, [_vp, C.POINTER(C.c_longlong), _dp]),
    "bf_state_init": (C.c_int, [_vp, _dp, C.POINTER(C.c_ubyte), _dp, C.c_double, C.c_double]),
    "bf_step": (C.c_int, [_vp, C.c_int, C.c_int]),
    "bf_state_download": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "bf_batch_step": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "bf_coords_upload": (C.c_int, [_vp, _dp, _dp, _dp]),
    "bf_fint_download": (C.c_int, [_vp, _dp]),
    "bf_batch_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "bf_batch_destroy": (C.c_int, [_vp]),
    "bf_batch_set_cases": (C.c_int, [_vp, C.POINTER(_vp), C.c_int]),
    "bf_batch_refresh": (C.c_int, [_vp]),
    "bf_batch_eval": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "bf_batch_sync": (C.c_int, [_vp]),
    "bf_batch_last_ms": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "bf_batch_timing_enable": (C.c_int, [_vp, C.c_int]),
    "bf_batch_last_fallback_count": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "bf_batch_info": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bf_batch_case_status": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bf_batch_step_async": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_longlong]),
    "bf_batch_step_wait": (C.c_int, [_vp, C.c_longlong, C.POINTER(C.c_float),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bf_io_dimension": (C.c_int, [C.c_char_p, C.c_int, _ip, _ip, _ip, _ip, _ip]),
    "bf_io_input": (C.c_int, [C.c_char_p, C.c_int, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _ip,
                              _ip, _ip, _ip]),
    "bf_io_write_lists": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _ip,
                                    _ip, _ip, C.c_int, _ip, _ip, C.c_int]),
    "bf_io_write_vtk": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(BfVtkFrame)]),
    "bf_io_write_state": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                    C.c_int]),
    "bf_io_read_state": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                   _ip]),
    "bf_batch_set_launch": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bf_batch_get_launch": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bf_batch_tune": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "bf_device_buffers": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "bf_sync": (C.c_int, [_vp]),
    "bf_topo_area0": (C.c_int, [C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _dp]),
    "bf_topo_irrver": (C.c_int, [C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, C.c_int, _ip, _ip, _ip, _ip]),
    "bf_topo_patch": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, C.c_int, _ip]),
    "bf_topo_ghost_ring": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip]),
    "bf_loop_tables": (C.c_int, [C.c_int, _dp]),
    "bf_version": (C.c_char_p, []),
}

_lib = None


def lib() -> C.CDLL:
    """Load libbiofem_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C biofem-ia_amd); the force path has no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class BfError(RuntimeError):
    def __init__(self, fn: str, rc: int):
        super().__init__(f"{fn} returned {ERR_NAMES.get(rc, rc)}")
        self.rc = rc


def check(fn: str, rc: int) -> None:
    if rc != BF_OK:
        raise BfError(fn, rc)


def ip(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)
