"""ctypes binding of libpa_hip.so (the C ABI declared in include/pa_hip.h).

The library is the only compute path of this package: if it cannot be loaded
every device operation raises; there is no host fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PA_HIP_LIB") or os.path.join(_HERE, "libpa_hip.so")

PA_MAX_PUPILS = 8
PA_NUM_MOMENTS = 7 + PA_MAX_PUPILS
PA_PART_SUMS = 6 + PA_MAX_PUPILS
PA_CH_PARAMS = 4
PA_ENS_QUANTITIES = 11
PA_PDT_BINS = 64
PA_LOWRANK_TERMS = 32
PA_SCREEN_PARAMS = 8 + PA_LOWRANK_TERMS * PA_LOWRANK_TERMS + 2 * PA_LOWRANK_TERMS

GEMM_MODES = {"f16x3": 0, "bf16x3": 1, "f32": 2}

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_longlong
f64 = C.c_double
f32 = C.c_float
u64 = C.c_uint64

# name -> (restype, argtypes), mirroring include/pa_hip.h
SIGNATURES = {
    "pa_last_error": (C.c_char_p, []),
    "pa_version": (i32, []),
    "pa_tile_rows": (i32, [i32]),
    "pa_operand_bytes": (i64, [i32, i32, i32]),
    "pa_fill_twiddles": (i32, [i32, vp, vp]),
    "pa_row_first": (i32, [i32, vp, vp, vp, vp, vp, i32, vp]),
    "pa_col_step": (i32, [i32, vp, vp, vp, vp, i32, i32, vp, i32, i32, vp]),
    "pa_col_groups": (i32, [i32, vp]),
    "pa_row_mid": (i32, [i32, vp, vp, vp, i32, vp, i32, vp]),
    "pa_source_fresnel": (i32, [i32, vp, vp, vp, vp, i32, vp, i32, vp]),
    "pa_row_mid_source": (i32, [i32, vp, vp, vp, vp, i32, vp]),
    "pa_expmi": (i32, [i32, vp, vp, vp]),
    "pa_row_last": (i32, [i32, vp, vp, vp, vp, vp, i32, vp, vp, vp, i32, vp]),
    "pa_field_partials": (i32, [i32, vp, vp, vp, vp, i32, vp, vp, i32, vp]),
    "pa_moments_finalize": (i32, [i32, vp, vp, i32, vp, vp, i32, vp]),
    "pa_ensemble_accumulate": (i32, [i32, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]),
    "pa_fft_screen": (i32, [i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, i32, vp]),
    "pa_channel_create": (i32, [C.POINTER(vp), i32, f64, f64, f64, f64, f64, i32, vp, vp, i32, vp, vp, i32, vp,
                                f64, f64, i32, i32, vp]),
    "pa_channel_run": (i32, [vp, i32, vp, vp, vp, vp, u64, vp, vp, vp, vp]),
    "pa_channel_destroy": (i32, [vp]),
    "pa_hbm_stream": (i32, [i32, vp, vp, i64, i32, i32, vp]),
    "pa_field_moments_static": (i32, [i32, vp, vp, vp, i32, vp, vp, vp, i32, vp]),
    "pa_field_moments_static_max": (i32, [i32]),
    "pa_spectral_draw": (i32, [i32, vp, vp, vp, vp, u64, i32, vp, vp, vp, i32, vp]),
    "pa_screen_prep": (i32, [i32, vp, vp, vp, vp, vp, i32, vp, vp, vp, i32, i32, vp]),
    "pa_screen_draw_prep": (i32, [i32, vp, vp, vp, vp, u64, i32, vp, vp, i32, vp, vp, i32, i32, vp]),
    "pa_screen_tables": (i32, [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp, i32, i32, vp]),
    "pa_screen_gemm": (i32, [i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp, i32, i32, vp]),
    "pa_screen_compose": (i32, [i32, vp, vp, vp, vp, i32, vp]),
    "pa_apply_pupil": (i32, [i32, vp, f64, f64, f64, f64, i32, vp]),
    "pa_structure_function_scratch": (i64, [i32, i32]),
    "pa_structure_function": (i32, [i32, vp, vp, vp, i32, vp]),
    "pa_outer": (i32, [i32, vp, vp, vp, i32, vp]),
    "pa_intensity": (i32, [i64, vp, vp, vp]),
    "pa_mlp_gemm_f32": (i32, [i32, i32, i32, vp, i32, i32, vp, i32, i32, vp, i32, f32, f32, vp]),
    "pa_mlp_bias_act": (i32, [i32, i32, vp, vp, vp, i32, vp]),
    "pa_mlp_act_backward": (i32, [i32, i32, vp, vp, vp, vp]),
    "pa_mlp_bias_grad": (i32, [i32, i32, vp, vp, vp]),
    "pa_mlp_head_loss": (i32, [i32, i32, vp, vp, vp, f32, f32, vp, vp, vp]),
    "pa_mlp_adam": (i32, [i32, vp, vp, vp, vp, f32, f32, f32, f32, i32, vp]),
    "pa_mlp_to_bf16": (i32, [i64, vp, vp, vp]),
    "pa_mlp_layer_bf16": (i32, [i32, i32, i32, vp, vp, vp, vp, i32, vp]),
}

_lib: Optional[C.CDLL] = None


class NativeLibraryError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load libpa_hip.so from the package directory (built by build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryError(
            f"{LIB_PATH} is missing: build it with pyatmosphere_amd/csrc/build.sh "
            "(or __graft_entry__.build()); this package has no host fallback")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> int:
    if rc < 0:
        msg = load().pa_last_error().decode(errors="replace")
        raise RuntimeError(f"libpa_hip {what}: {msg} (status {rc})")
    return rc


def call(name: str, *args) -> int:
    return check(getattr(load(), name)(*args), name)


def ptr(t) -> Optional[int]:
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle() -> Optional[int]:
    import torch
    return torch.cuda.current_stream().cuda_stream
