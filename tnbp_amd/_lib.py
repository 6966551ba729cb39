"""ctypes binding of libhbp.so (the C-ABI declared in include/hbp.h).

The library is required: importing the product path without a built
``libhbp.so`` raises, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhbp.so")

HBP_MAX_ORDER = 16
HBP_C128, HBP_F64, HBP_C64, HBP_F32 = 0, 1, 2, 3
HBP_PREC_FP64, HBP_PREC_MIXED = 0, 1
HBP_ABSORB_AUTO, HBP_ABSORB_LAZY, HBP_ABSORB_EAGER = 0, 1, 2
HBP_SWEEP_AUTO, HBP_SWEEP_PER_SITE, HBP_SWEEP_GROUPED, HBP_SWEEP_MERGED = 0, 1, 2, 3
HBP_MODE_FUSED, HBP_MODE_PLAIN = 0, 1
HBP_SCHEDULE_JACOBI, HBP_SCHEDULE_CHECKERBOARD = 0, 1
HBP_BATCH_ALONE, HBP_BATCH_SMALL, HBP_BATCH_GRAPH = 0, 1, 2

STATUS = {
    1: ValueError,       # HBP_ERR_INVALID_ARGUMENT (std::invalid_argument)
    2: RuntimeError,     # HBP_ERR_LOGIC (std::logic_error)
    3: MemoryError,      # HBP_ERR_BAD_ALLOC
    4: IndexError,       # HBP_ERR_OUT_OF_RANGE
    5: RuntimeError,     # HBP_ERR_HIP
    6: NotImplementedError,  # HBP_ERR_UNSUPPORTED
}

_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

# name -> argtypes (restype is always int = hbp_status unless listed in _RES)
SIGNATURES = {
    "hbp_ctx_create": [C.c_int, C.POINTER(_vp)],
    "hbp_ctx_destroy": [_vp],
    "hbp_ctx_synchronize": [_vp],
    "hbp_ctx_stream": [_vp, C.POINTER(_vp)],
    "hbp_ctx_memory": [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "hbp_ctx_trim": [_vp],
    "hbp_autotune_report": [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)],
    "hbp_device_memory": [C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "hbp_ctx_graph_stats": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "hbp_ctx_profile": [_vp, C.c_int],
    "hbp_ctx_profile_read": [_vp, C.c_int, C.POINTER(C.c_int64), _dp, _dp, _dp],
    "hbp_tensor_create": [_vp, C.c_int, C.c_int, _i32p, C.POINTER(_vp)],
    "hbp_tensor_destroy": [_vp, _vp],
    "hbp_tensor_info": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), _i32p, C.POINTER(C.c_size_t)],
    "hbp_tensor_upload": [_vp, _vp, _vp],
    "hbp_tensor_download": [_vp, _vp, _vp],
    "hbp_tensor_copy": [_vp, _vp, _vp],
    "hbp_tensor_reshape": [_vp, _vp, C.c_int, _i32p],
    "hbp_transpose": [_vp, _vp, _i32p, _vp],
    "hbp_conj": [_vp, _vp],
    "hbp_fill": [_vp, _vp, C.c_double, C.c_double],
    "hbp_contract": [_vp, _vp, _i32p, _vp, _i32p, _vp, _i32p, C.c_int],
    "hbp_norm": [_vp, _vp, _dp],
    "hbp_normalize": [_vp, _vp, _dp],
    "hbp_scale": [_vp, _vp, C.c_double, C.c_double],
    "hbp_linear_combine": [_vp, C.c_int, C.POINTER(_vp), _dp, _vp],
    "hbp_trace": [_vp, _vp, C.c_int, _i32p, _vp],
    "hbp_diag": [_vp, _vp, _vp],
    "hbp_get_elem": [_vp, _vp, _i32p, _dp, _dp],
    "hbp_svd": [_vp, _vp, C.c_int, _vp, _vp, _vp],
    "hbp_trunc_svd": [_vp, _vp, C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, _dp],
    "hbp_svd_noalloc": [_vp, _vp, C.c_int, C.c_int, _vp, _dp, _vp],
    "hbp_tensor_data": [_vp, C.POINTER(_vp)],
    "hbp_tensor_version": [_vp, C.POINTER(C.c_uint64)],
    "hbp_peps_bind_sites": [_vp, _vp, C.POINTER(_vp)],
    "hbp_peps_load_messages": [_vp, _vp, C.POINTER(_vp)],
    "hbp_peps_store_messages": [_vp, _vp, C.c_int, _i32p, C.POINTER(_vp)],
    "hbp_peps_store_sites": [_vp, _vp, C.POINTER(_vp)],
    "hbp_peps_create": [_vp, C.c_int, _i32p, _i32p, C.c_int, C.POINTER(_vp)],
    "hbp_peps_destroy": [_vp, _vp],
    "hbp_peps_set_mode": [_vp, C.c_int],
    "hbp_peps_set_residual_floor": [_vp, C.c_double],
    "hbp_peps_set_schedule": [_vp, C.c_int],
    "hbp_peps_set_accel": [_vp, C.c_int, C.c_double],
    "hbp_peps_set_pair_cache": [_vp, C.c_int],
    "hbp_peps_set_phase_gauge": [_vp, C.c_int],
    "hbp_peps_set_absorb_mode": [_vp, C.c_int],
    "hbp_peps_set_sweep_kernel": [_vp, C.c_int],
    "hbp_peps_set_bond_dims": [_vp, _vp, _i32p, _i32p],
    "hbp_peps_num_sites": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "hbp_peps_site_shape": [_vp, C.c_int, C.POINTER(C.c_int), _i32p],
    "hbp_peps_get_site": [_vp, _vp, C.c_int, _dp],
    "hbp_peps_set_site": [_vp, _vp, C.c_int, C.c_int, _i32p, _dp],
    "hbp_peps_get_message": [_vp, _vp, C.c_int, _i32p, _dp],
    "hbp_peps_set_message": [_vp, _vp, C.c_int, C.c_int32, _dp],
    "hbp_peps_absorb_tpo": [_vp, _vp, C.POINTER(_dp), _i32p],
    "hbp_peps_absorb_tpo_batched": [_vp, C.c_int, C.POINTER(_vp), C.POINTER(C.POINTER(_dp)),
                                    C.POINTER(_i32p)],
    "hbp_peps_bp_sweep": [_vp, _vp],
    "hbp_peps_bp_condition": [_vp, _vp, _dp],
    "hbp_peps_bp_run": [_vp, _vp, C.c_int, C.c_double, C.POINTER(C.c_int), _dp],
    "hbp_peps_bp_run_batched": [_vp, C.c_int, C.POINTER(_vp), C.c_int, C.c_double,
                                C.POINTER(C.c_int), _dp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "hbp_peps_bp_iterate": [_vp, _vp, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), _dp,
                            C.POINTER(C.c_int)],
    "hbp_peps_truncate": [_vp, _vp, C.c_int, C.c_double, _i32p, _dp],
    "hbp_peps_truncate_batched": [_vp, C.c_int, C.POINTER(_vp), _i32p, _dp, _i32p, _dp],
    "hbp_peps_measure_sites": [_vp, _vp, C.c_int, _i32p, _dp, _dp],
    "hbp_peps_measure_edges": [_vp, _vp, C.c_int, _i32p, _dp, _dp],
    "hbp_peps_measure_paths": [_vp, _vp, C.c_int, _i32p, _i32p, _dp, _dp],
    "hbp_last_error": [],
    "hbp_version": [],
}
_RES = {"hbp_last_error": C.c_char_p, "hbp_version": C.c_char_p}

_lib = None


def load(path: str = LIB_PATH):
    """Load libhbp.so and declare every entry point; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"libhbp.so not built at {path}: run `python -m tnbp_amd.build`")
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RES.get(name, C.c_int)
    _lib = lib
    return lib


def check(status: int):
    if status != 0:
        msg = load().hbp_last_error().decode(errors="replace")
        raise STATUS.get(status, RuntimeError)(f"hbp error {status}: {msg}")


def i32(seq):
    seq = list(seq)
    return (C.c_int32 * max(1, len(seq)))(*seq)
