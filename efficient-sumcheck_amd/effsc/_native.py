"""Loads the gfx950 library (libeffsc_hip.so) and declares its C ABI.

There is no CPU fallback: if the library is missing this raises, so every
caller runs the HIP path or fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("EFFSC_HIP_LIB", os.path.join(PKG_ROOT, "lib", "libeffsc_hip.so"))

u64 = C.c_uint64
u64p = C.POINTER(C.c_uint64)
sz = C.c_size_t
szp = C.POINTER(C.c_size_t)
vp = C.c_void_p


class Config(C.Structure):
    """effsc_hip_config (include/effsc_hip.h)."""

    _fields_ = [
        ("ext", C.c_int),
        ("input_ext", C.c_int),
        ("repr", C.c_int),
        ("ext_w", u64),
        ("device", C.c_int),
        ("flags", C.c_uint),
    ]


class Request(C.Structure):
    """effsc_hip_request (include/effsc_hip.h)."""

    _fields_ = [
        ("kind", C.c_int),
        ("v", C.c_uint),
        ("seeds", u64 * 3),
        ("tables", u64p * 3),
        ("transcript", vp),
        ("priority", C.c_double),
    ]


HOOK = C.CFUNCTYPE(C.c_int, vp, sz, vp)

_SIGS = {
    "effsc_hip_config_default": (None, [C.POINTER(Config)]),
    "effsc_hip_last_error": (C.c_char_p, []),
    "effsc_hip_multilinear_new": (C.c_int, [C.POINTER(Config), u64p, sz, C.POINTER(vp)]),
    "effsc_hip_inner_product_new": (C.c_int, [C.POINTER(Config), u64p, u64p, sz, C.POINTER(vp)]),
    "effsc_hip_product_new": (C.c_int, [C.POINTER(Config), C.POINTER(u64p), C.c_int, sz, C.POINTER(vp)]),
    "effsc_hip_eq_factored_new": (C.c_int, [C.POINTER(Config), u64p, sz, u64p, sz, C.POINTER(vp)]),
    "effsc_hip_gkr_new": (C.c_int, [C.POINTER(Config), u64p, u64p, u64p, sz, C.POINTER(vp)]),
    "effsc_hip_multilinear_lsb_new": (C.c_int, [C.POINTER(Config), u64p, sz, C.POINTER(vp)]),
    "effsc_hip_inner_product_lsb_new": (C.c_int, [C.POINTER(Config), u64p, u64p, sz, C.POINTER(vp)]),
    "effsc_hip_product_lsb_new": (C.c_int, [C.POINTER(Config), C.POINTER(u64p), C.c_int, sz, C.POINTER(vp)]),
    "effsc_hip_bit_reverse": (C.c_int, [C.c_int, u64p, u64p, sz, C.c_int]),
    "effsc_hip_gkr_claimed_w": (C.c_int, [vp, u64p]),
    "effsc_hip_degree": (C.c_int, [vp, szp]),
    "effsc_hip_ext": (C.c_int, [vp, szp]),
    "effsc_hip_round": (C.c_int, [vp, u64p, u64p]),
    "effsc_hip_finalize": (C.c_int, [vp, u64p]),
    "effsc_hip_final_value": (C.c_int, [vp, u64p]),
    "effsc_hip_table_count": (C.c_int, [vp, szp]),
    "effsc_hip_table_len": (C.c_int, [vp, sz, szp]),
    "effsc_hip_read_table": (C.c_int, [vp, sz, u64p, sz, szp]),
    "effsc_hip_reset": (C.c_int, [vp]),
    "effsc_hip_free": (None, [vp]),
    "effsc_hip_profile": (C.c_int, [vp, C.c_int]),
    "effsc_hip_profile_read": (C.c_int, [vp, u64p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "effsc_hip_profile_read_muls": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "effsc_hip_synchronize": (C.c_int, [C.c_int]),
    "effsc_hip_fold": (C.c_int, [C.c_int, u64p, sz, u64, szp]),
    "effsc_hip_fold_ext": (C.c_int, [C.c_int, C.c_int, u64, u64p, sz, u64p, szp]),
    "effsc_hip_transcript_new": (vp, [C.c_char_p, sz]),
    "effsc_hip_transcript_new_test": (vp, [u64]),
    "effsc_hip_transcript_send": (C.c_int, [vp, C.c_int, u64p]),
    "effsc_hip_transcript_challenge": (C.c_int, [vp, C.c_int, u64p]),
    "effsc_hip_transcript_free": (None, [vp]),
    "effsc_hip_sumcheck": (C.c_int, [vp, sz, vp, HOOK, vp, u64p, u64p, u64p]),
    "effsc_hip_sumcheck_noalloc": (C.c_int, [C.c_int, C.c_int, u64p, u64p, sz, vp, u64p, u64p, u64p]),
    "effsc_hip_sumcheck_verify": (
        C.c_int,
        [C.c_int, u64, u64p, sz, sz, vp, u64p, sz, HOOK, vp, u64p, u64p, szp, C.POINTER(C.c_int)],
    ),
    "effsc_hip_batch_new": (C.c_int, [C.c_int, C.c_int, sz, szp, C.POINTER(u64p), C.POINTER(u64p), C.c_uint,
                                      C.POINTER(vp)]),
    "effsc_hip_batch_rounds": (C.c_int, [vp, sz, szp, szp]),
    "effsc_hip_batch_prove": (C.c_int, [vp, C.POINTER(vp), u64p, u64p, u64p]),
    "effsc_hip_batch_autotune": (C.c_int, [vp, szp, sz, C.c_int, szp, C.POINTER(C.c_double)]),
    "effsc_hip_batch_free": (None, [vp]),
    "effsc_hip_device_alloc": (C.c_int, [C.c_int, sz, C.POINTER(u64p)]),
    "effsc_hip_device_free": (C.c_int, [C.c_int, u64p]),
    "effsc_hip_copy": (C.c_int, [C.c_int, u64p, u64p, sz]),
    "effsc_hip_fill_uniform": (C.c_int, [C.c_int, u64p, sz, u64, u64]),
    "effsc_hip_mul_rate": (C.c_int, [C.c_int, C.POINTER(C.c_double)]),
    "effsc_hip_hbm_rate": (C.c_int, [C.c_int, C.c_int, sz, C.POINTER(C.c_double)]),
    "effsc_hip_stream_open": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(vp)]),
    "effsc_hip_stream_num_variables": (C.c_int, [vp, szp]),
    "effsc_hip_stream_read": (C.c_int, [vp, sz, sz, u64p]),
    "effsc_hip_stream_close": (None, [vp]),
    "effsc_hip_stream_claim": (C.c_int, [C.c_int, C.POINTER(vp), C.c_int, u64, u64p]),
    "effsc_hip_stream_prover_new": (C.c_int, [C.POINTER(Config), C.c_int, C.POINTER(vp), C.POINTER(vp)]),
    "effsc_hip_stream_prover_new_ex": (C.c_int, [C.POINTER(Config), C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp)]),
    "effsc_hip_server_new": (C.c_int, [C.c_int, sz, sz, C.POINTER(vp)]),
    "effsc_hip_server_footprint": (C.c_int, [C.POINTER(Request), szp]),
    "effsc_hip_server_submit": (C.c_int, [vp, C.POINTER(Request), u64p]),
    "effsc_hip_server_wait": (C.c_int, [vp, u64, u64p, u64p, u64p, C.POINTER(C.c_double)]),
    "effsc_hip_server_stats": (C.c_int, [vp, u64p, u64p, u64p, u64p]),
    "effsc_hip_server_now": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "effsc_hip_server_free": (None, [vp]),
    "effsc_hip_mlp_new": (C.c_int, [C.c_int, C.c_int, C.c_int, u64, C.POINTER(vp)]),
    "effsc_hip_mlp_param_count": (C.c_int, [vp, szp]),
    "effsc_hip_mlp_get_params": (C.c_int, [vp, C.POINTER(C.c_float)]),
    "effsc_hip_mlp_set_params": (C.c_int, [vp, C.POINTER(C.c_float)]),
    "effsc_hip_mlp_forward": (C.c_int, [vp, C.POINTER(C.c_float), sz, C.POINTER(C.c_float)]),
    "effsc_hip_mlp_loss_grad": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), sz, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "effsc_hip_mlp_train": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), sz, C.c_int, C.c_float, C.POINTER(C.c_float)]),
    "effsc_hip_mlp_free": (None, [vp]),
}

# Entry points declared in include/effsc_hip.h (checked by the CPU tests).
EXPORTS = tuple(_SIGS)

_lib = None


class NativeError(RuntimeError):
    """A non-OK effsc_hip_status."""

    def __init__(self, code, msg):
        super().__init__(f"effsc_hip status {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libeffsc_hip.so not found at {LIB_PATH}; build it with `make -C efficient-sumcheck_amd` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise NativeError(rc, lib().effsc_hip_last_error().decode())
    return rc
