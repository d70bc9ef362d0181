"""ctypes binding of libsparsebackends_amd.so (include/sparsebackends_amd.h).

The library is required: there is no fallback path.  If it is missing or no
HIP device is present, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsparsebackends_amd.so")

ERROR_NAMES = {1: "AssertionError", 2: "ErrorException", 3: "ArgumentError", 4: "KeyError",
               5: "MethodError", 6: "BoundsError", 7: "DeviceError", 8: "OutOfMemory",
               9: "InvalidHandle", 10: "LimitExceeded", 11: "LAPACKException"}

# every entry point the header declares: (name, argtypes)
P = C.c_void_p
PP = C.POINTER(C.c_void_p)
I64P = C.POINTER(C.c_int64)
U64P = C.POINTER(C.c_uint64)
F64P = C.POINTER(C.c_double)
I32P = C.POINTER(C.c_int)
SIGNATURES = {
    "sb_last_error": [],
    "sb_version": [],
    "sb_ctx_create": [C.c_int, PP],
    "sb_ctx_destroy": [P],
    "sb_ctx_set_stream": [P, P],
    "sb_ctx_synchronize": [P],
    "sb_ctx_memory": [P, I64P, I64P],
    "sb_ctx_trim": [P],
    "sb_ctx_profile": [P, C.c_int],
    "sb_ctx_set_svd_sweeps": [P, C.c_int],
    "sb_ctx_profile_reset": [P],
    "sb_ctx_profile_read": [P, C.c_char_p, C.c_int64, I64P],
    "sb_tensor_free": [P],
    "sb_tensor_kind": [P, I32P],
    "sb_tensor_types": [P, I32P, I32P],
    "sb_rekey": [P, P, C.c_int, C.c_int, PP],
    "sb_coo_retypeval": [P, P, C.c_int, PP],
    "sb_bs_new": [P, C.c_int, C.c_int, I64P, PP],
    "sb_bs_from_host": [P, C.c_int, C.c_int, I64P, C.c_int64, I64P, I64P, F64P, PP],
    "sb_bs_info": [P, I32P, I32P, I64P, I64P, I64P],
    "sb_bs_to_host": [P, P, I64P, I64P, F64P],
    "sb_bs_device_ptrs": [P, PP, PP, PP],
    "sb_bs_set_payload": [P, P, P, C.c_int],
    "sb_al_new": [P, C.c_int, C.c_int, I64P, PP],
    "sb_al_from_host": [P, C.c_int, C.c_int, I64P, C.c_int64, F64P, C.c_int64, I64P, I64P, F64P, PP],
    "sb_al_info": [P, I32P, I32P, I64P, I64P, I64P, I64P],
    "sb_al_to_host": [P, P, I64P, I64P, F64P, F64P],
    "sb_al_getindex": [P, P, I64P, F64P],
    "sb_al_compression_ratio": [P, F64P],
    "sb_al_to_blocksparse": [P, P, PP],
    "sb_al_to_dense": [P, P, C.c_double, PP],
    "sb_bs_getindex": [P, P, I64P, F64P],
    "sb_bs_setindex": [P, P, C.c_double, I64P],
    "sb_bs_sort_lex": [P, P],
    "sb_bs_is_dense": [P, I32P],
    "sb_bs_from_dense": [P, P, C.c_int, C.c_double, C.c_double, C.c_int, PP],
    "sb_bs_to_dense": [P, P, C.c_double, C.c_int, I64P, PP],
    "sb_bs_fuse_two_axes": [P, P, C.c_int64, C.c_int64, PP],
    "sb_coo_new": [P, C.c_int, I64P, PP],
    "sb_coo_from_host": [P, C.c_int, I64P, C.c_int64, I64P, F64P, C.c_int, PP],
    "sb_coo_info": [P, I32P, I64P, I64P, I32P],
    "sb_coo_to_host": [P, P, I64P, F64P],
    "sb_coo_device_ptrs": [P, PP, PP],
    "sb_coo_getindex": [P, P, I64P, F64P],
    "sb_coo_setindex": [P, P, C.c_double, I64P, C.c_int],
    "sb_coo_sort": [P, P],
    "sb_coo_from_dense": [P, P, C.c_double, C.c_double, PP],
    "sb_coo_to_dense": [P, P, PP],
    "sb_coo_fuse_axes": [P, P, C.c_int64, C.c_int64, PP],
    "sb_dense_from_host": [P, C.c_int, I64P, F64P, PP],
    "sb_dense_from_device": [P, C.c_int, I64P, P, C.c_int, PP],
    "sb_dense_info": [P, I32P, I64P],
    "sb_dense_to_host": [P, P, F64P],
    "sb_dense_device_ptr": [P, PP],
    "sb_permutedims": [P, P, I64P, PP],
    "sb_permutedims_inplace": [P, P, I64P],
    "sb_contract": [P, P, U64P, C.c_int, P, U64P, C.c_int, P, U64P, C.c_int, C.c_uint32],
    "sb_contract_bs_bs_to_dense": [P, P, U64P, C.c_int, P, U64P, C.c_int, P, U64P, C.c_int, C.c_uint32],
    "sb_plan_create": [P, P, U64P, C.c_int, P, U64P, C.c_int, P, U64P, C.c_int, C.c_uint32, PP],
    "sb_plan_execute": [P, P],
    "sb_heff_plan_create": [P, P, P, P, P, P, P, C.c_uint32, PP],
    "sb_plan_stats": [P, F64P, F64P, F64P, I64P, I64P],
    "sb_plan_box_config": [P, I32P, I32P, I32P, F64P],
    "sb_autotune_reset": [],
    "sb_plan_operand_bytes": [P, F64P, F64P, F64P],
    "sb_plan_mfma_flops": [P, F64P],
    "sb_plan_free": [P],
    "sb_graph_create": [P, PP, C.c_int, PP],
    "sb_graph_launch": [P, P],
    "sb_batch_graph_create": [P, P, PP],
    "sb_graph_stats": [P, I64P, I64P, I64P, I64P],
    "sb_graph_free": [P],
    "sb_ctx_graphs": [P, C.c_int],
    "sb_batch_create": [P, PP, C.c_int, C.c_int, PP],
    "sb_batch_execute": [P, P, I32P, C.c_int],
    "sb_batch_stats": [P, F64P, I32P],
    "sb_batch_free": [P],
    "sb_batch_autotune": [P, P],
    "sb_batch_stage_tile": [P, C.c_int, I32P, F64P],
    "sb_bs_norm": [P, P, F64P],
    "sb_bs_copy": [P, P, PP],
    "sb_bs_similar": [P, P, PP],
    "sb_bs_scale": [P, P, C.c_double, PP],
    "sb_bs_fill": [P, P, C.c_double],
    "sb_bs_axpby": [P, C.c_double, P, C.c_double, P],
    "sb_bs_add": [P, P, P, PP],
    "sb_bs_dot": [P, P, P, F64P],
    "sb_bs_norm_noalloc": [P, P, P],
    "sb_bs_restrict": [P, P, P, PP],
    "sb_bs_left_qr": [P, P, PP, PP],
    "sb_bs_left_factorize": [P, P, PP, PP],
    "sb_bs_qr_sectors": [P, P, I64P, I64P, I64P, PP, PP, I64P],
    "sb_bs_svd": [P, P, C.c_int, C.c_int64, C.c_int64, C.c_double, PP, PP, F64P, C.c_int64, I64P, F64P],
    "sb_bs_svd_sectors": [P, P, C.c_int64, I64P, I64P, I64P, I64P, C.c_int, C.c_int64, C.c_int64, C.c_double, PP,
                          PP, I64P, F64P, C.c_int64, I64P, F64P],
    "sb_eigsolve_lanczos_batch": [P, P, PP, PP, C.c_int, C.c_int, C.c_int, C.c_double, F64P, F64P, I64P, I32P,
                                  I32P],
    "sb_eigsolve_lanczos_multi": [P, PP, C.c_int, C.c_int, PP, PP, C.c_int, C.c_int, C.c_double, F64P, F64P, I64P,
                                  I32P, I32P],
    "sb_eigsolve_lanczos": [P, PP, C.c_int, P, P, C.c_int, C.c_int, C.c_double, F64P, F64P, I32P, F64P, I64P,
                            I32P, I32P],
}

_lib = None


class SBError(Exception):
    """A status returned by the library; .kind names the reference's exception class."""

    def __init__(self, code, msg):
        self.code = code
        self.kind = ERROR_NAMES.get(code, str(code))
        super().__init__(f"{self.kind}: {msg}")


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C sparsebackends_amd` "
                               "(or __graft_entry__.build())")
        lib = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = C.c_int
        lib.sb_last_error.restype = C.c_char_p
        _lib = lib
    return _lib


def check(rc):
    if rc != 0:
        raise SBError(rc, load().sb_last_error().decode())
