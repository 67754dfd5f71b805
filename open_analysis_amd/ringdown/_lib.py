"""ctypes binding of ``liboaring_hip.so`` (include/oaring.h). Loading fails loudly when the
library is not built; there is no CPU path."""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parents[1] / "csrc" / "liboaring_hip.so"
OK, E_INVALID, E_OOM, E_DEVICE, E_NO_DEVICE = 0, 1, 2, 3, 5
P = C.c_void_p
_LIB: C.CDLL | None = None

EXPORTS = (
    "oaring_param_count", "oaring_create", "oaring_set_params", "oaring_get_params",
    "oaring_get_grads", "oaring_reset_adam", "oaring_set_batch", "oaring_forward",
    "oaring_backward", "oaring_adam", "oaring_gemm_check", "oaring_last_error", "oaring_destroy",
)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            from ..accel import AccelError, ErrorInfo

            raise AccelError(ErrorInfo("ACCEL_NOT_BUILT", f"{LIB_PATH.name} is not built",
                                       str(LIB_PATH),
                                       "Run python -m open_analysis_amd.csrc.build (hipcc)."))
        h = C.CDLL(str(LIB_PATH))
        sig = {
            "oaring_param_count": (C.c_int64, []),
            "oaring_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(P)]),
            "oaring_set_params": (C.c_int, [P, P, C.c_int64]),
            "oaring_get_params": (C.c_int, [P, P, C.c_int64]),
            "oaring_get_grads": (C.c_int, [P, P, C.c_int64]),
            "oaring_reset_adam": (C.c_int, [P]),
            "oaring_set_batch": (C.c_int, [P, P, P, P, P, P, P, C.c_float, C.c_float, C.c_float]),
            "oaring_forward": (C.c_int, [P, P, P, P]),
            "oaring_backward": (C.c_int, [P]),
            "oaring_adam": (C.c_int, [P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]),
            "oaring_gemm_check": (C.c_int, [C.c_int, C.c_int, C.c_int, P, P, P]),
            "oaring_last_error": (C.c_char_p, []),
            "oaring_destroy": (None, [P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


def check(code: int, what: str) -> None:
    if code == OK:
        return
    from ..accel import AccelError, ErrorInfo

    msg = (lib().oaring_last_error() or b"").decode(errors="replace")
    if code == E_INVALID:
        raise ValueError(f"{what}: {msg}")
    name = {E_OOM: "ACCEL_OUT_OF_MEMORY", E_NO_DEVICE: "ACCEL_NO_DEVICE"}.get(code,
                                                                           "ACCEL_DEVICE_ERROR")
    raise AccelError(ErrorInfo(name, f"{what}: {msg}", what))
