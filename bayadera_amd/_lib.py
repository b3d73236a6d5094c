"""ctypes binding of ``libbayadera_hip.so`` (declarations: include/bayadera_hip.h).

The library is the only compute path: if it is missing this module raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbayadera_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "bayadera_hip.h")

BH_OK = 0
BH_ERR_ARG = -1
BH_ERR_HIP = -2
BH_ERR_RTC = -3
BH_ERR_RANGE = -4
BH_ERR_STATE = -5

BH_PARTNER_FLAT = 0
BH_PARTNER_WALKER = 1
BH_BATCH_STATE_LDS = 0
BH_BATCH_STATE_GLOBAL = 1
BH_EXEC_AUTO = 0
BH_EXEC_STEPWISE = 1
BH_EXEC_RESIDENT = 2
BH_GLM_LOGISTIC = 1
BH_GLM_GAUSSIAN = 2
BH_GLM_FP64 = 0
BH_GLM_FP32 = 1
BH_GLM_BF16 = 2
BH_GLM_BF16X3 = 3
BH_GLM_FP16X2 = 4
BH_GLM_FP16 = 5

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_S = ctypes.c_char_p
_SZ = ctypes.c_size_t
_PP = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "bh_last_error": [],
    "bh_version": [],
    "bh_rtc_compile_check": [_S, _S, _S, _S, _U32, _U32, _S, _P, _SZ],
    "bh_ctx_create": [_I, _U32, _PP],
    "bh_ctx_destroy": [_P],
    "bh_ctx_info": [_P, _P, _P],
    "bh_uniform_sample": [_P, _U32, _F, _F, _P, _U64],
    "bh_direct_sample": [_P, _I, _U32, _P, _U32, _P, _U64],
    "bh_dist_engine_create": [_P, _S, _S, _U32, _PP, _P, _SZ],
    "bh_dist_engine_destroy": [_P],
    "bh_log_density": [_P, _P, _U32, _P, _U32, _U64, _P],
    "bh_density": [_P, _P, _U32, _P, _U32, _U64, _P],
    "bh_lik_engine_create": [_P, _S, _S, _PP, _P, _SZ],
    "bh_lik_engine_destroy": [_P],
    "bh_loglik": [_P, _P, _U32, _P, _U32, _U64, _P],
    "bh_lik": [_P, _P, _U32, _P, _U32, _U64, _P],
    "bh_evidence": [_P, _P, _U32, _P, _U32, _U64, _P],
    "bh_dataset_mean": [_P, _P, _U32, _U64, _U32, _P],
    "bh_dataset_variance": [_P, _P, _U32, _U64, _U32, _P],
    "bh_dataset_variance_noalloc": [_P, _P, _U32, _U64, _U32, _P],
    "bh_dataset_histogram": [_P, _P, _U32, _U64, _U32, _P, _P, _P],
    "bh_acor": [_P, _P, _U32, _U32, _P, _P, _P, _P],
    "bh_acor_batch": [_P, _P, _U32, _P, _U32, _P, _P, _P, _P],
    "bh_stretch_factory_create": [_P, _S, _S, _U32, _U32, _PP, _P, _SZ],
    "bh_stretch_factory_destroy": [_P],
    "bh_glm_stretch_factory_create": [_P, _I, _P, _P, _U64, _U32, _F, _PP, _P, _SZ],
    "bh_glm_gaussian_factory_create": [_P, _P, _P, _U64, _U32, _F, _F, _PP, _P, _SZ],
    "bh_glm_log_density": [_P, _P, _U64, _I, _P],
    "bh_stretch_set_glm_precision": [_P, _I],
    "bh_stretch_set_glm_guard": [_P, ctypes.c_double],
    "bh_stretch_set_kernel_timing": [_P, _I],
    "bh_stretch_kernel_timing": [_P, _P, _P],
    "bh_stretch_glm_guard_stats": [_P, _P, _P],
    "bh_stretch_create": [_P, _U32, _U32, _P, _U32, _PP],
    "bh_stretch_destroy": [_P],
    "bh_stretch_set_partner": [_P, _I],
    "bh_stretch_set_exec": [_P, _I],
    "bh_stretch_info": [_P, _P, _P],
    "bh_stretch_init": [_P, _U32],
    "bh_stretch_init_position": [_P, _U32, _P],
    "bh_stretch_init_position_from": [_P, _P],
    "bh_stretch_burn_in": [_P, _U64, _F],
    "bh_stretch_anneal": [_P, _P, _U64, _F],
    "bh_stretch_run": [_P, _U64, _F, _P, _P, _P, _P, _P, _P],
    "bh_stretch_acc_rate": [_P, _F, _P],
    "bh_stretch_init_move": [_P, _F, _U64],
    "bh_stretch_move": [_P],
    "bh_stretch_move_bare": [_P],
    "bh_stretch_set_temperature": [_P, _D],
    "bh_stretch_read_accounting": [_P, _P, _P, _U64],
    "bh_stretch_sample": [_P, _P, _U64],
    "bh_stretch_histogram": [_P, _U32, _P, _P, _P],
    "bh_stretch_mean": [_P, _P],
    "bh_stretch_variance": [_P, _P],
    "bh_batch_create": [_P, _S, _S, _U32, _U32, _P, _P, _P, _P, _U32, _PP, _P, _SZ],
    "bh_batch_destroy": [_P],
    "bh_batch_set_partner": [_P, _I],
    "bh_batch_init_position": [_P, _U32, _P],
    "bh_batch_burn_in": [_P, _U32, _U64, _F],
    "bh_batch_run": [_P, _U32, _U64, _F, _P, _P, _P],
    "bh_batch_autotune": [_P, _U32, _P, _P, _P],
    "bh_batch_set_state": [_P, _I],
    "bh_batch_class_states": [_P, _P, _P],
    "bh_batch_run_series": [_P, _U32, _U64, _F, _P, _P, _P, _P],
    "bh_glm_batch_create": [_P, _I, _U32, _P, _P, _P, _P, _P, _F, _U32, _PP],
    "bh_glm_batch_destroy": [_P],
    "bh_glm_batch_set_precision": [_P, _I],
    "bh_glm_batch_set_tiles": [_P, _U32, _U32, _U32],
    "bh_glm_batch_autotune": [_P, _U32, _P, _P, _P],
    "bh_glm_batch_init_position": [_P, _U32, _P],
    "bh_glm_batch_burn_in": [_P, _U32, _U64, _F],
    "bh_glm_batch_run": [_P, _U32, _U64, _F, _P, _P, _P, _P],
    "bh_glm_batch_state_get": [_P, _P, _P],
    "bh_glm_batch_last_gpu_ms": [_P, _P],
    "bh_glm_batch_log_density": [_P, _I, _P],
    "bh_fp8_decode": [_P, _P, _P, _U64],
    "bh_device_rates": [_P, _P, _P, _P, _P],
    "bh_device_copy_rate": [_P, _U64, _P, _P],
    "bh_device_gather_rate": [_P, _U64, _I, _P],
    "bh_fp8_decode_f16": [_P, _P, _P, _U64],
    "bh_batch_state_get": [_P, _P, _P],
    "bh_batch_last_gpu_ms": [_P, _P],
    "bh_stretch_set_timing": [_P, _I],
    "bh_stretch_last_gpu_ms": [_P, _P],
    "bh_stretch_state_get": [_P, _P, _P, _P],
    "bh_stretch_state_set": [_P, _P, _P, _P],
}

_RESTYPE = {"bh_last_error": ctypes.c_char_p}

_lib = None


class BayaderaError(RuntimeError):
    """A failed engine call; ``code`` is the BH_ERR_* status."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class ArgumentError(BayaderaError, ValueError):
    pass


def header_functions() -> list[str]:
    """Names of the functions declared in include/bayadera_hip.h."""
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(bh_\w+)\s*\(", text, flags=re.M)


def load(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 and
    # only works with that one, while this library needs any libamdhip64.so.7;
    # whichever loads first serves both, so let torch's load first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build the HIP engine first (python -c "
            "'import __graft_entry__ as g; g.build()')")
    L = ctypes.CDLL(path)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = argtypes
        fn.restype = _RESTYPE.get(name, ctypes.c_int)
    _lib = L
    return L


def last_error() -> str:
    return load().bh_last_error().decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    if rc == BH_OK:
        return
    msg = last_error()
    if what:
        msg = f"{what}: {msg}"
    if rc in (BH_ERR_ARG, BH_ERR_RANGE):
        raise ArgumentError(rc, msg)
    raise BayaderaError(rc, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
