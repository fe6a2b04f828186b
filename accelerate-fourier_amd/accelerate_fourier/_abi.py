"""ctypes binding of the C ABI (include/accelerate_fourier.h).

The shared library is built in tree by ``accelerate-fourier_amd/build.py``.
Loading fails loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "libaccelerate_fourier_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_PKG_ROOT), "include", "accelerate_fourier.h")

# af_status
SUCCESS, INVALID_ARG, UNSUPPORTED, NO_DEVICE, HIP_ERROR, OUT_OF_MEMORY, INTERNAL = range(7)
# af_precision
F64, F32 = 0, 1
# af_kind
C2C, R2C_FULL, C2R_FULL, CONV_CYCLIC, REAL_CONV = range(5)
# af_algorithm
ALG_AUTO, ALG_DECOMPOSE, ALG_CHIRP2, ALG_CHIRP235 = range(4)
FLAG_PLAIN_TWIDDLES = 0x1
FLAG_COOLEY_TUKEY = 0x2


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("length", ctypes.c_int64),
        ("sign", ctypes.c_int32),
        ("kind", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("algorithm", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
    ]


class AfError(RuntimeError):
    def __init__(self, status: int, what: str, detail: str = ""):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} (status {status})" + (f" [{detail}]" if detail else ""))


_lib = None
_v = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_p64 = ctypes.POINTER(ctypes.c_int64)

_SIGNATURES = {
    "af_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "af_version": (ctypes.c_char_p, []),
    "af_last_error": (ctypes.c_char_p, []),
    "af_plan_create": (ctypes.c_int, [ctypes.POINTER(PlanDesc), ctypes.POINTER(_v)]),
    "af_plan_destroy": (ctypes.c_int, [_v]),
    "af_plan_describe": (ctypes.c_int, [ctypes.POINTER(PlanDesc), ctypes.c_char_p, ctypes.c_size_t]),
    "af_plan_workspace_size": (ctypes.c_int, [_v, _i64, ctypes.POINTER(ctypes.c_size_t)]),
    "af_execute": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v, _v]),
    "af_execute_noalloc": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v]),
    "af_plan_num_passes": (ctypes.c_int, [_v, _i64, ctypes.POINTER(_i32)]),
    "af_execute_pass": (ctypes.c_int, [_v, _i32, _i64, _v, _v, _v, _v, _v, _v]),
    "af_two_to_spectrum": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v, _v, _v]),
    "af_two_from_spectrum": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v, _v, _v]),
    "af_untangle_spectra": (ctypes.c_int, [_i32, _i64, _i64, _v, _v, _v, _v, _v, _v, _v]),
    "af_execute_2d": (ctypes.c_int, [_v, _v, _i64, _v, _v, _v, _v, _v]),
    "af_execute_3d": (ctypes.c_int, [_v, _v, _v, _i64, _v, _v, _v, _v, _v]),
    "af_transpose": (ctypes.c_int, [_i32, _i64, _i64, _i64, _v, _v, _v, _v, _v]),
    "af_two_to_spectrum_2d": (ctypes.c_int, [_v, _v, _i64, _v, _v, _v, _v, _v, _v, _v]),
    "af_two_from_spectrum_2d": (ctypes.c_int, [_v, _v, _i64, _v, _v, _v, _v, _v, _v, _v]),
    "af_untangle_spectra_2d": (ctypes.c_int, [_i32, _i64, _i64, _i64, _v, _v, _v, _v, _v, _v, _v]),
    "af_entangle_spectra": (ctypes.c_int, [_i32, _i64, _i64, _v, _v, _v, _v, _v, _v, _v]),
    "af_plan_tune": (ctypes.c_int, [_v, _i64, _i32, _v, ctypes.c_char_p, ctypes.c_size_t]),
    "af_wisdom_export": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "af_wisdom_import": (ctypes.c_int, [ctypes.c_char_p]),
    "af_to_spectrum": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v]),
    "af_from_spectrum": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v]),
    "af_convolve_cyclic": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v, _v, _v, _v]),
    "af_real_convolve": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v]),
    "af_vjp_execute": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v, _v]),
    "af_vjp_to_spectrum": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v]),
    "af_vjp_from_spectrum": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v]),
    "af_vjp_real_convolve": (ctypes.c_int, [_v, _i64, _v, _v, _v, _v, _v, _v]),
    "af_karatsuba_i64": (ctypes.c_int, [_i64, _i64, _v, _v, _v, _v]),
    "af_karatsuba_ragged_i64": (ctypes.c_int, [_i64, _p64, _v, _v, _v, _v]),
    "af_karatsuba_cyclic_i64": (ctypes.c_int, [_i64, _i64, _v, _v, _v, _v]),
    "af_karatsuba_complex_i64": (ctypes.c_int, [_i64, _i64, ctypes.c_int32, _v, _v, _v, _v, _v, _v, _v]),
    "af_batch_create": (ctypes.c_int, [_i64, _p64, _i32, _i32, ctypes.POINTER(_v)]),
    "af_batch_destroy": (ctypes.c_int, [_v]),
    "af_batch_execute": (ctypes.c_int, [_v, _v, _v, _v, _v, _v]),
    "af_batch_describe": (ctypes.c_int, [_v, ctypes.c_char_p, ctypes.c_size_t]),
    "af_batch_tune": (ctypes.c_int, [_v, _v, ctypes.c_char_p, ctypes.c_size_t]),
    "af_batch_set_schedule": (ctypes.c_int, [_v, _i64, _i32]),
    "af_rader_permutations": (ctypes.c_int, [_i64, _p64, _p64, _p64]),
    "af_good_thomas_maps": (ctypes.c_int, [_i64, _i64, _p64, _p64]),
    "af_multiplicative_generator": (_i64, [_i64]),
    "af_ceiling_5smooth": (_i64, [_i64]),
    "af_ceiling_power_of_two": (_i64, [_i64]),
}


def lib():
    """The loaded native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python accelerate-fourier_amd/build.py` "
                "(the transform path has no non-native fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def status_string(status: int) -> str:
    return lib().af_status_string(status).decode()


def check(status: int, what: str) -> None:
    if status != SUCCESS:
        detail = lib().af_last_error().decode() if status in (HIP_ERROR, OUT_OF_MEMORY) else ""
        raise AfError(status, what, detail)


def int64_array(values):
    arr = (ctypes.c_int64 * len(values))(*[int(v) for v in values])
    return arr
