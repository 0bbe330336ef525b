"""ctypes binding of libfemeko_hip.so (the C ABI declared in include/femeko_hip.h).

The library is loaded from the package's own lib/ directory. There is no
fallback: if it is missing or fails to load, every call raises
FemekoHIPError.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libfemeko_hip.so")

FK_OK = 0
FK_ERR_INVALID_ARG = 1
FK_ERR_HIP = 2
FK_ERR_DEGENERATE = 3
FK_ERR_INDEX = 4
FK_ERR_NONFINITE = 5
FK_ERR_NOT_CONVERGED = 6
FK_ERR_SINGULAR = 7
FK_ERR_OOM = 8
FK_ERR_STATE = 9
FK_ERR_UNSUPPORTED = 10
FK_ERR_DOMAIN = 11

FK_PREC_F64 = 1
FK_PREC_F32_REFINED = 2
FK_PREC_F32_RESTART = 3
FK_ASM_COLOURED = 0
FK_ASM_ATOMIC = 1
FK_ASM_GATHER = 2
FK_REORDER_AUTO = 0
FK_REORDER_NONE = 1
FK_REORDER_RCM = 2


class FemekoHIPError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


_c = ctypes
_dp = _c.POINTER(_c.c_double)
_fp = _c.POINTER(_c.c_float)
_ip = _c.POINTER(_c.c_int32)
_up = _c.POINTER(_c.c_uint8)
_i64p = _c.POINTER(_c.c_int64)
_vp = _c.c_void_p
_vpp = _c.POINTER(_c.c_void_p)

# name -> (restype, argtypes); every function of include/femeko_hip.h
SIGNATURES = {
    "fk_version": (_c.c_int, []),
    "fk_status_string": (_c.c_char_p, [_c.c_int]),
    "fk_context_create": (_c.c_int, [_c.c_int, _vpp]),
    "fk_context_destroy": (_c.c_int, [_vp]),
    "fk_context_last_error": (_c.c_char_p, [_vp]),
    "fk_context_synchronize": (_c.c_int, [_vp]),
    "fk_mesh_create": (_c.c_int, [_vp, _dp, _c.c_int64, _ip, _c.c_int64, _c.c_int32, _dp, _vpp]),
    "fk_mesh_create_ex": (_c.c_int, [_vp, _dp, _c.c_int64, _ip, _c.c_int64, _c.c_int32, _dp, _c.c_int32, _vpp]),
    "fk_mesh_numbering": (_c.c_int, [_vp, _ip, _dp, _dp]),
    "fk_rcm_order": (_c.c_int, [_ip, _c.c_int64, _c.c_int64, _c.c_int32, _ip]),
    "fk_mesh_destroy": (_c.c_int, [_vp]),
    "fk_mesh_info": (_c.c_int, [_vp, _i64p, _ip, _i64p]),
    "fk_mesh_volumes": (_c.c_int, [_vp, _dp]),
    "fk_pattern": (_c.c_int, [_vp, _c.c_int32, _ip, _ip]),
    "fk_set_coefficient": (_c.c_int, [_vp, _dp]),
    "fk_assemble": (_c.c_int, [_vp, _c.c_int32]),
    "fk_values_generation": (_c.c_int, [_vp, _i64p]),
    "fk_values": (_c.c_int, [_vp, _dp]),
    "fk_values_f32": (_c.c_int, [_vp, _fp]),
    "fk_stiffness_matrix": (_c.c_int, [_vp, _dp, _c.c_int32, _ip, _ip, _dp]),
    "fk_local_stiffness": (_c.c_int, [_vp, _dp, _dp]),
    "fk_tangential_stiffness": (_c.c_int, [_vp, _dp, _dp, _dp]),
    "fk_lagrange": (_c.c_int, [_vp, _dp]),
    "fk_set_dirichlet": (_c.c_int, [_vp, _up, _dp]),
    "fk_solve": (_c.c_int, [_vp, _c.c_int32, _c.c_double, _c.c_int64, _c.c_double, _c.c_int32, _c.c_int32,
                            _i64p, _ip, _dp]),
    "fk_solve_rhs": (_c.c_int, [_vp, _dp, _c.c_double, _c.c_int64, _dp, _i64p, _dp]),
    "fk_potential": (_c.c_int, [_vp, _dp]),
    "fk_set_potential": (_c.c_int, [_vp, _dp]),
    "fk_field": (_c.c_int, [_vp, _dp, _dp, _dp]),
    "fk_field_noalloc": (_c.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "fk_boundary_integral": (_c.c_int, [_vp, _c.c_int64, _ip, _dp, _dp, _dp, _ip, _c.c_int32, _c.c_int32, _dp]),
    "fk_scalar_potential": (_c.c_int, [_dp, _dp, _ip, _ip, _dp, _c.c_int32, _c.c_int32, _c.c_int32, _dp, _dp, _dp,
                                       _c.c_int32]),
    "fk_scalar_potential_last_error": (_c.c_char_p, []),
    "fk_solve_neumann": (_c.c_int, [_vp, _dp, _c.c_double, _c.c_int64, _dp, _dp, _i64p, _dp]),
    "fk_nonlinear_froehlich": (_c.c_int, [_vp, _up, _c.c_double, _c.c_double, _c.c_double, _c.c_double,
                                          _c.c_int32, _c.c_double, _ip, _i64p, _dp]),
    "fk_set_inner_precision": (_c.c_int, [_vp, _c.c_int32]),
    "fk_nonlinear_spline": (_c.c_int, [_vp, _up, _dp, _c.c_int32, _dp, _dp, _c.c_int32, _dp, _c.c_double,
                                       _c.c_double, _c.c_int32, _c.c_double, _ip, _i64p, _dp]),
    "fk_demag_field": (_c.c_int, [_vp, _up, _dp, _dp, _c.c_double, _c.c_int64, _c.c_int32, _dp, _i64p, _dp]),
    "fk_newton_saddle": (_c.c_int, [_vp, _up, _dp, _c.c_double, _c.c_double, _dp, _c.c_int32, _dp, _dp, _c.c_int32,
                                    _dp, _c.c_double, _c.c_double, _c.c_int32, _c.c_double, _ip, _i64p, _dp, _dp]),
    "fk_line_search": (_c.c_int, [_vp, _dp, _dp, _dp, _c.c_double, _dp]),
    "fk_get_coefficient": (_c.c_int, [_vp, _dp]),
    "fk_adjoint_field_target": (_c.c_int, [_vp, _up, _dp, _c.c_double, _dp, _dp, _i64p]),
    "fk_batch_create": (_c.c_int, [_vp, _c.c_int32, _i64p, _i64p, _c.POINTER(_dp), _c.POINTER(_ip), _c.c_int32,
                                   _c.c_int32, _vpp]),
    "fk_batch_destroy": (_c.c_int, [_vp]),
    "fk_batch_info": (_c.c_int, [_vp, _i64p, _i64p, _ip, _ip]),
    "fk_batch_set_coefficient": (_c.c_int, [_vp, _c.POINTER(_dp)]),
    "fk_batch_assemble": (_c.c_int, [_vp]),
    "fk_batch_set_dirichlet": (_c.c_int, [_vp, _c.POINTER(_up), _c.POINTER(_dp)]),
    "fk_batch_solve": (_c.c_int, [_vp, _c.c_double, _c.c_double, _c.c_int32, _c.c_int64, _i64p, _ip, _dp]),
    "fk_batch_potential": (_c.c_int, [_vp, _c.c_int32, _dp]),
    "fk_batch_autotune": (_c.c_int, [_vp, _ip, _ip, _dp]),
    "fk_batch_set_precision": (_c.c_int, [_vp, _c.c_int32]),
    "fk_batch_set_format": (_c.c_int, [_vp, _c.c_int32]),
    "fk_batch_solver_format": (_c.c_int, [_vp, _c.POINTER(_c.c_int32)]),
    "fk_batch_set_launch": (_c.c_int, [_vp, _c.c_int32, _c.c_int32]),
    "fk_batch_launch": (_c.c_int, [_vp, _ip, _ip, _ip, _ip]),
    "fk_set_timing": (_c.c_int, [_vp, _c.c_int32]),
    "fk_spmv_timing": (_c.c_int, [_vp, _i64p, _dp]),
    "fk_copy_probe": (_c.c_int, [_vp, _c.c_int64, _c.c_int32, _dp]),
    "fk_spmv_burst": (_c.c_int, [_vp, _c.c_int32, _c.c_int32, _c.POINTER(_c.c_double)]),
    "fk_set_lanes_per_row": (_c.c_int, [_vp, _c.c_int32]),
    "fk_set_spmv_format": (_c.c_int, [_vp, _c.c_int32]),
    "fk_sell_layout": (_c.c_int, [_vp, _ip, _ip, _c.POINTER(_c.c_int64)]),
    "fk_set_sell_shape": (_c.c_int, [_vp, _c.c_int32, _c.c_int32]),
    "fk_sell_info": (_c.c_int, [_vp, _c.c_void_p, _c.c_void_p, _c.c_void_p]),
    "fk_solver_format": (_c.c_int, [_vp, _c.c_void_p, _c.c_void_p, _c.c_void_p]),
}

_lib = None


def load():
    """Load the library (once). Raises FemekoHIPError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FemekoHIPError(FK_ERR_STATE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    try:
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - depends on the machine
        raise FemekoHIPError(FK_ERR_HIP, f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def fptr(a):
    return a.ctypes.data_as(_fp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def uptr(a):
    return a.ctypes.data_as(_up)
