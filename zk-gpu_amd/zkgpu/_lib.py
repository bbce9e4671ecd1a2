"""ctypes binding of libzkgpu.so (include/zkgpu.h).

The library is the product; this module only declares its symbols. It fails loudly when the
library is missing: there is no CPU fallback behind any of the calls.
"""
from __future__ import annotations

import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("ZKGPU_LIB", os.path.join(PKG_ROOT, "_build", "libzkgpu.so"))

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_void_pp = ctypes.POINTER(ctypes.c_void_p)
# zkg_error_fn: void (*)(const char* message, void* user)
ERROR_FN = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_void_p)

# name: (restype, argtypes)
SIGNATURES = {
    "zkg_last_error": (ctypes.c_char_p, []),
    "zkg_version": (ctypes.c_int, []),
    "zkg_device_open": (ctypes.c_int, [ctypes.c_int, c_void_pp]),
    "zkg_device_close": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_device_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_size_t), ctypes.c_char_p, ctypes.c_size_t]),
    "zkg_device_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "zkg_device_report_errors": (ctypes.c_int, [ctypes.c_void_p, ERROR_FN, ctypes.c_void_p]),
    "zkg_buffer_create": (ctypes.c_int, [ctypes.c_void_p, c_u32p, ctypes.c_size_t, c_void_pp]),
    "zkg_buffer_empty": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, c_void_pp]),
    "zkg_buffer_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, c_u32p]),
    "zkg_buffer_read_async": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, c_u32p, c_void_pp]),
    "zkg_read_done": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "zkg_read_wait": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_buffer_write": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, c_u32p, ctypes.c_size_t]),
    "zkg_buffer_size": (ctypes.c_size_t, [ctypes.c_void_p]),
    "zkg_buffer_wrap": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, c_void_pp]),
    "zkg_buffer_device_ptr": (ctypes.c_void_p, [ctypes.c_void_p]),
    "zkg_buffer_free": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_wait": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_msm_default_window": (ctypes.c_uint32, [ctypes.c_size_t]),
    "zkg_msm_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                      c_void_pp]),
    "zkg_msm_create_tuned": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, c_void_pp]),
    "zkg_msm_plan_check": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
    "zkg_msm_item_points": (ctypes.c_uint32, [ctypes.c_void_p]),
    "zkg_msm_create_precomputed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_uint32,
                                                  ctypes.c_uint32, c_void_pp]),
    "zkg_msm_precompute_bases": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                                c_void_pp]),
    "zkg_msm_base_points": (ctypes.c_size_t, [ctypes.c_void_p]),
    "zkg_msm_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_msm_group": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_msm_window": (ctypes.c_uint32, [ctypes.c_void_p]),
    "zkg_msm_points": (ctypes.c_size_t, [ctypes.c_void_p]),
    "zkg_msm_upload_bases": (ctypes.c_int, [ctypes.c_void_p, c_u32p, ctypes.c_size_t, c_void_pp]),
    "zkg_msm_upload_bases_montgomery_256": (ctypes.c_int, [ctypes.c_void_p, c_u32p, ctypes.c_size_t, c_void_pp]),
    "zkg_msm_convert_bases": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t]),
    "zkg_msm_check_bases": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "zkg_msm_run_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_uint32,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "zkg_msm_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_void_pp]),
    "zkg_msm_run_yielding": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_void_pp]),
    "zkg_msm_result": (ctypes.c_int, [ctypes.c_void_p, c_void_pp]),
    "zkg_msm_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                       ctypes.c_size_t, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]),
    "zkg_xyzz_to_affine": (ctypes.c_int, [c_u32p, c_u32p, ctypes.POINTER(ctypes.c_int)]),
    "zkg_xyzz_to_affine_g2": (ctypes.c_int, [c_u32p, c_u32p, ctypes.POINTER(ctypes.c_int)]),
    "zkg_ntt_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_void_pp]),
    "zkg_ntt_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_ntt_log_n": (ctypes.c_uint32, [ctypes.c_void_p]),
    "zkg_ntt_passes": (ctypes.c_uint32, [ctypes.c_void_p]),
    "zkg_ntt_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "zkg_ntt_inverse": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "zkg_batch_create": (ctypes.c_int, [ctypes.c_void_p, c_void_pp]),
    "zkg_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_batch_add_ntt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int]),
    "zkg_batch_add_msm": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "zkg_batch_run": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_batch_clear": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_batch_capture": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_batch_captured": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_groth16_h_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_void_pp]),
    "zkg_groth16_h_create_form": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, c_void_pp]),
    "zkg_groth16_h_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_groth16_h_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "zkg_tune_msm": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, c_u32p, c_u32p, c_u32p,
                                    ctypes.POINTER(ctypes.c_double)]),
    "zkg_tune_ntt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u32p, ctypes.POINTER(ctypes.c_double)]),
    "zkg_tune_lookup_msm": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int),
                                           c_u32p, c_u32p, c_u32p]),
    "zkg_tune_lookup_ntt": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int), c_u32p]),
    "zkg_tune_load": (ctypes.c_int, [ctypes.c_char_p]),
    "zkg_tune_save": (ctypes.c_int, [ctypes.c_char_p]),
    "zkg_tune_clear": (ctypes.c_int, []),
    "zkg_server_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, c_void_pp]),
    "zkg_server_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_server_add_key": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p * 4, ctypes.c_void_p,
                                          ctypes.POINTER(ctypes.c_int)]),
    "zkg_server_slot_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_size_t)]),
    "zkg_server_add_key_shaped": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p * 4,
                                                 ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "zkg_server_submit_groth16": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "zkg_server_add_zkey": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "zkg_server_key_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "zkg_server_key_slot_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
    "zkg_zkey_open": (ctypes.c_int, [ctypes.c_char_p, c_void_pp]),
    "zkg_zkey_close": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_zkey_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "zkg_zkey_points": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
    "zkg_zkey_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, c_void_pp]),
    "zkg_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p * 7, ctypes.c_uint32,
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "zkg_server_done": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "zkg_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, c_u32p, ctypes.POINTER(ctypes.c_double)]),
    "zkg_server_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "zkg_server_reset_stats": (ctypes.c_int, [ctypes.c_void_p]),
    "zkg_test_field_op": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_u32p, c_u32p, c_u32p,
                                         ctypes.c_size_t]),
    "zkg_test_point_op": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_u32p, c_u32p, c_u32p,
                                         ctypes.c_size_t]),
    "zkg_bench_field_mul": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "zkg_bench_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "zkg_test_glv_split": (ctypes.c_int, [ctypes.c_void_p, c_u32p, ctypes.c_size_t, c_u32p]),
    "zkg_test_gls_split": (ctypes.c_int, [ctypes.c_void_p, c_u32p, ctypes.c_size_t, c_u32p]),
    "zkg_util_generator_multiples": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                                    ctypes.c_void_p]),
}

_lib = None


class ZkgError(RuntimeError):
    """A failed library call; the message is the library's (the reference's panic text)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def load(required_symbols=None):
    """Load libzkgpu.so and declare every symbol; raises if the library or a symbol is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libzkgpu.so not built at {LIB_PATH}: run `make -C zk-gpu_amd` or __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        if required_symbols is not None and name not in required_symbols:
            continue
        fn = getattr(lib, name)  # AttributeError = missing export: fail loudly
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        msg = load().zkg_last_error().decode(errors="replace")
        raise ZkgError(status, msg)


def u32_ptr(arr):
    return arr.ctypes.data_as(c_u32p)
