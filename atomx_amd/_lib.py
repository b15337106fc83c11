"""ctypes binding of libatomx_slater.so (the C ABI declared in include/atomx_slater.h).

The product path has no CPU fallback: importing a compute entry point without
the built library, or calling one without a HIP device, raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

from ._build import LIB_PATH

AX_OK = 0
AX_ERR_VALUE = 1
AX_ERR_TYPE = 2
AX_ERR_MEMORY = 3
AX_ERR_NONFINITE = 4
AX_ERR_HIP = 5

AX_PREC_FP64 = 0
AX_PREC_SPLIT = 1

PRECISIONS = {"fp64": AX_PREC_FP64, "split": AX_PREC_SPLIT}

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes); every symbol declared in include/atomx_slater.h
SIGNATURES = {
    "ax_last_error": (ctypes.c_char_p, []),
    "ax_version": (_int, []),
    "ax_split_slices": (_int, []),
    "ax_split_ktile": (_int, []),
    "ax_split_products": (_int, []),
    "ax_device_count": (_int, [ctypes.POINTER(_int)]),
    "ax_yk_partial_matrices": (_int, [_int, _dp, _dp]),
    "ax_grid_rpow": (_int, [_vp, _i64, _int, _vp, _vp, _vp]),
    "ax_pair_density": (_int, [_vp, _vp, _i64, _i64, _int, _vp, _vp, _vp, _vp, _i64, _vp]),
    "ax_yk_batch": (_int, [_vp, _i64, _vp, _vp, _vp, _i64, _int, _int, _int, _vp, _vp, _int,
                           _vp, _vp, _vp, _vp, _i64, _vp]),
    "ax_rk_dot": (_int, [_vp, _i64, _vp, _i64, _i64, _int, _vp, _vp, _vp, _vp]),
    "ax_bspline_eval": (_int, [_vp, _i64, _int, _vp, _int, _int, _int, _int, _vp, _i64, _vp]),
    "ax_assemble_integral": (_int, [_int, _int, _int, _int, _int, _vp, _vp, _vp, _int, _vp,
                                    _vp]),
    "ax_symmetrize": (_int, [_vp, _int, _int, _vp]),
    "ax_exchange_matrix_yk_batch": (_int, [_vp, _i64, _vp, _vp, _vp, _vp, _int, _int, _int,
                                           _int, _int, _int, _vp, _vp, _vp, _vp]),
    "ax_exchange_matrix_yk": (_int, [_vp, _vp, _vp, _vp, _int, _int, _int, _int, _int, _vp,
                                     _vp]),
    "ax_slater_rk_direct": (_int, [_vp, _vp, _vp, _vp, _int, _int, _int, _int, _vp, _vp, _vp,
                                   _vp, ctypes.c_double, _vp, _vp, _vp]),
    "ax_rk_gemm_workspace_bytes": (_i64, [_int, _i64, _int]),
    "ax_rk_gemm": (_int, [_vp, _i64, _vp, _vp, _i64, _int, _i64, _int, _vp, _vp, _vp]),
    "ax_rk_gemm_noalloc": (_int, [_vp, _i64, _vp, _vp, _i64, _int, _i64, _int, _vp, _vp]),
    "ax_fock_plan_create": (_int, [_int, _ip, _ip, _dp, _dp, _dp, _vp, _vp, _vp, _vp, _vp,
                                   _int, _int, ctypes.POINTER(_vp)]),
    "ax_fock_plan_run": (_int, [_vp, _vp]),
    "ax_fock_plan_bytes": (_i64, [_vp]),
    "ax_fock_plan_rk_equivalents": (_i64, [_vp]),
    "ax_fock_plan_destroy": (_int, [_vp]),
    "ax_rk_tensor_create": (_int, [_dp, _dp, _int, _int, _int, _ip, _int, _int,
                                   ctypes.POINTER(_vp)]),
    "ax_rk_tensor_create_pairs": (_int, [_dp, _dp, _int, _int, _int, _int, _ip, _ip, _int,
                                         _int, ctypes.POINTER(_vp)]),
    "ax_rk_tensor_destroy": (_int, [_vp]),
    "ax_rk_tensor_create_many": (_int, [_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _ip, _ip,
                                        _ip, ctypes.POINTER(_vp), _ip, _ip, _vp,
                                        ctypes.POINTER(_vp)]),
    "ax_server_create": (_int, [_i64, _int, _int, ctypes.POINTER(_vp)]),
    "ax_server_destroy": (_int, [_vp]),
    "ax_server_submit": (_int, [_vp, _vp, _vp, _int, _int, _int, _vp, _int, _int, _vp, _vp,
                                _i64, ctypes.POINTER(_i64)]),
    "ax_server_drop_state": (_int, [_vp, _i64]),
    "ax_server_poll": (_int, [_vp, _i64, ctypes.POINTER(ctypes.c_int32)]),
    "ax_server_wait": (_int, [_vp, _i64]),
    "ax_server_result": (_int, [_vp, _i64, ctypes.POINTER(_vp), ctypes.POINTER(_i64),
                                ctypes.POINTER(ctypes.c_int32)]),
    "ax_server_release": (_int, [_vp, _i64]),
    "ax_server_stats": (_int, [_vp, ctypes.POINTER(_i64), _int]),
    "ax_rk_tensor_describe": (_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(ctypes.c_int32),
                                     _ip, ctypes.POINTER(_i64)]),
    "ax_rk_tensor_estimate": (_i64, [_int, _int, _int, _ip, _int, _int]),
    "ax_rk_tensor_size": (_i64, [_vp]),
    "ax_rk_tensor_pairs": (_int, [_vp, _int]),
    "ax_rk_tensor_offset": (_i64, [_vp, _int]),
    "ax_rk_tensor_footprint": (_i64, [_vp]),
    "ax_rk_tensor_slices": (_int, [_vp]),
    "ax_rk_tensor_run": (_int, [_vp, _vp, _vp, _vp, _vp]),
    "ax_rk_tensor_run_stage": (_int, [_vp, _int, _vp, _vp, _vp, _vp]),
    "ax_rk_tensor_run_batch": (_int, [ctypes.POINTER(_vp), _int, ctypes.POINTER(_vp),
                                      ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "ax_rk_tensor_set_schedule": (_int, [_vp, _int, _int]),
    "ax_rk_tensor_get_schedule": (_int, [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    "ax_rk_tensor_tune": (_int, [_vp, _vp, _vp, _vp, _int, _vp, ctypes.POINTER(ctypes.c_float)]),
    "ax_pool_cached_bytes": (_i64, []),
    "ax_pool_trim": (_int, []),
    "ax_gauss_legendre_01": (_int, [_int, _dp, _dp]),
    "ax_cell_moments": (_int, [_vp, _vp, _vp, _vp, _int, _int, _int, _int, _int, _int, _vp, _vp]),
    "ax_cell_triang": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _int, _int, _int, _int, _int, _int,
                              _vp, _vp, _vp, _vp]),
    "ax_cell_pair_integrals": (_int, [_vp, _vp, _vp, _vp, _vp, _int, _int, _int, _int, _int, _vp,
                                      _vp]),
    "ax_cell_global_moments": (_int, [_vp, _int, _int, _int, _vp, _vp]),
    "ax_density_sym": (_int, [_vp, _i64, _vp, _vp, _int, _int, _int, _int, _int, _vp, _vp]),
    "ax_combine_rows": (_int, [_vp, _i64, _int, _vp, _int, _vp, _vp]),
    "ax_exchange_density": (_int, [_vp, _vp, _i64, _int, _int, _vp, _int, _vp, _vp]),
    "ax_convol_case1": (_int, [_vp, _int, _int, _int, _vp, _i64, _int, _vp, _i64, _vp]),
    "ax_convol_case2": (_int, [_vp, _int, _int, _int, _int, _vp, _i64, _int, _vp, _i64, _vp]),
    "ax_convol_case4_acc": (_int, [_vp, _int, _int, _int, _vp, _i64, _int, _int, _vp, _vp, _vp]),
    "ax_banded_to_full_add_at": (_int, [_vp, _vp, _vp, _i64, _int, _int, _int, ctypes.c_double,
                                        _vp]),
    "ax_convol_case4": (_int, [_vp, _int, _int, _int, _vp, _i64, _int, _int, _vp, _i64, _vp]),
    "ax_convol_semi_workspace": (_i64, [_int, _int, _int]),
    "ax_convol_case1_semi": (_int, [_vp, _vp, _vp, _vp, _vp, _int, _int, _int, _vp, _i64, _int,
                                    _vp, _vp, _i64, _vp]),
    "ax_convol_case4_semi": (_int, [_vp, _vp, _vp, _vp, _vp, _int, _int, _int, _vp, _i64, _int,
                                    _int, _vp, _i64, _vp]),
    "ax_sum_amb_banded": (_int, [_vp, _i64, _int, _vp, _i64, _int, _int, _int, _int, _vp, _vp]),
    "ax_banded_to_full_add": (_int, [_vp, _i64, _vp, _i64, _int, _int, _int, ctypes.c_double, _vp]),
    "ax_axpy": (_int, [_vp, _vp, _i64, ctypes.c_double, _vp]),
    "ax_mul_rows": (_int, [_vp, _vp, _vp, _int, _i64, _vp]),
    "ax_bilinear_stage2": (_int, [_vp, _int, _int, _int, _vp, _int, _vp, _vp]),
    "ax_fock_finalize": (_int, [_vp, _vp, _vp, _vp, _int, ctypes.c_double, _int, _vp]),
    "ax_sum_amb_ptr": (_int, [_vp, _vp, _int, _int, _int, _int, _vp, _vp]),
    "ax_spinor_eri_packed": (_int, [_int, _int, _int, _int, _vp, _vp, _vp, _vp, _vp, _int, _int,
                                    _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "ax_spinor_eri_list": (_int, [_int, _int, _int, _int, _vp, _vp, _vp, _vp, _vp, _i64, _vp,
                                  _vp, _vp, _vp]),
    "ax_bilinear_stage2_batch": (_int, [_int, _int, _int, _int, _vp, _vp]),
    "ax_mbpt2_workspace_doubles": (_i64, [_int, _int]),
    "ax_mbpt2_pair_sum": (_int, [_int, _int, _int, _int, _int, _vp, _vp, _vp, _i64, _i64, _vp,
                                 _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ax_feat_sample": (_int, [_vp, _int, _int, _int, _int, _vp, _i64, _int, _int, _vp, _i64,
                              _i64, _vp]),
    "ax_feat_apply": (_int, [_vp, _i64, _int, _i64, _vp, _vp, _vp, _vp, _int, _int, _vp, _vp,
                             _vp, _vp, _vp, _i64, _vp]),
    "ax_feat_apply_workspace": (_i64, [_int, _int]),
    "ax_feat_products": (_int, [_vp, _i64, _i64, _int, _vp, _vp, _int, _vp, _i64, _vp]),
    "ax_feat_dots": (_int, [_vp, _i64, _vp, _i64, _i64, _int, _vp, _vp, _int, _vp, _vp]),
    "ax_cpc_moments": (_int, [_vp, _i64, _int, _vp, _i64, _int, _vp, _vp, _vp, _int, _int,
                              _int, _vp, _vp, _vp]),
    "ax_cpc_scan": (_int, [_vp, _vp, _int, _i64, _int, _vp, _vp, _vp]),
    "ax_cpc_workspace_bytes": (_i64, [_int, _int, _int, _int, _int, _i64, _i64, _int, _int]),
    "ax_cpc_apply": (_int, [_vp, _int, _int, _vp, _i64, _int, _vp, _i64, _vp, _i64, _vp, _i64,
                            _int, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _int, _int,
                            _vp, _vp, _vp]),
}

_lock = threading.Lock()
_lib = None


class NativeLibraryMissing(RuntimeError):
    """libatomx_slater.so has not been built (run __graft_entry__.build())."""


def load():
    """Loads the shared library (no GPU needed) and binds every symbol."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise NativeLibraryMissing(
                    f"{LIB_PATH} not found: build the HIP extension first "
                    "(python -c 'import __graft_entry__ as g; g.build()')")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def register(name, restype, argtypes):
    """Binds an additional symbol of the library (used by sibling modules)."""
    lib = load()
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    SIGNATURES[name] = (restype, argtypes)
    return fn


def last_error() -> str:
    msg = load().ax_last_error()
    return msg.decode() if msg else ""


_EXC = {
    AX_ERR_VALUE: ValueError,
    AX_ERR_TYPE: TypeError,
    AX_ERR_MEMORY: MemoryError,
    AX_ERR_NONFINITE: FloatingPointError,
    AX_ERR_HIP: RuntimeError,
}


def check(status: int, what: str = "") -> None:
    if status != AX_OK:
        exc = _EXC.get(status, RuntimeError)
        raise exc(f"{what}: {last_error()}" if what else last_error())


def call(name: str, *args):
    """Calls an ax_* entry point that returns a status, raising on failure."""
    fn = getattr(load(), name)
    check(fn(*args), name)
