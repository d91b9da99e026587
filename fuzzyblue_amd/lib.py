"""ctypes binding of libfuzzyblue_hip.so (the C ABI in include/fuzzyblue_hip.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import re

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfuzzyblue_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "fuzzyblue_hip.h")

_lib = None


class FuzzyblueError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def declared_functions() -> list:
    """Names of every function include/fuzzyblue_hip.h declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(fb_[a-z0-9_]+)\s*\(", text)
    return sorted(set(n for n in names if not n.endswith("_t")))


def load():
    """Load the in-tree library (built by fuzzyblue_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FuzzyblueError(
            abi.FB_ENODEV,
            f"{LIB_PATH} is missing: build it with `python -m fuzzyblue_amd.build`",
        )
    # HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues per process
    # (default 4); streams beyond that share a queue and run in FIFO order
    # with each other.  The serving loop's render stream must not queue
    # behind a precompute stream, so ask for one queue per stream unless the
    # caller chose a value (read when the HIP runtime initialises).
    os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER(abi.fb_parameters)
    V = C.c_void_p
    S = C.c_int
    sigs = {
        "fb_status_string": (C.c_char_p, [S]),
        "fb_last_error": (C.c_char_p, []),
        "fb_abi_version": (C.c_uint32, []),
        "fb_parameters_default": (None, [P]),
        "fb_parameters_validate": (S, [P]),
        "fb_parameters_extents": (None, [P, V, V, V]),
        "fb_parameters_footprint": (C.c_size_t, [P, V, V]),
        "fb_builder_create": (S, [C.c_int, C.POINTER(V)]),
        "fb_builder_destroy": (None, [V]),
        "fb_atmosphere_build": (S, [V, V, P, C.POINTER(V)]),
        "fb_atmosphere_build_batch": (S, [V, V, P, C.c_size_t, V]),
        "fb_atmosphere_build_into": (
            S, [V, V, P, V, C.c_size_t, V, C.c_size_t, V, C.c_size_t]),
        "fb_build_into_scratch_limit": (C.c_size_t, []),
        "fb_pending_relaunch": (S, [V, V]),
        "fb_builder_autotune": (S, [V, V, P, C.POINTER(abi.fb_tuning)]),
        "fb_builder_get_tuning": (S, [V, P, C.POINTER(abi.fb_tuning)]),
        "fb_builder_set_tuning": (S, [V, P, C.POINTER(abi.fb_tuning)]),
        "fb_pending_profile": (
            S, [V, V, C.POINTER(abi.fb_pass_timing), C.c_size_t, C.POINTER(C.c_size_t)]),
        "fb_pending_peek": (V, [V]),
        "fb_pending_query": (S, [V]),
        "fb_pending_wait": (S, [V]),
        "fb_pending_stream_wait": (S, [V, V]),
        "fb_pending_record_ready_event": (S, [V, V]),
        "fb_pending_assert_ready": (S, [V, C.POINTER(V)]),
        "fb_pending_destroy": (None, [V]),
        "fb_pending_lut": (S, [V, S, C.POINTER(abi.fb_lut_view)]),
        "fb_atmosphere_lut": (S, [V, S, C.POINTER(abi.fb_lut_view)]),
        "fb_atmosphere_parameters": (S, [V, P]),
        "fb_atmosphere_download": (S, [V, S, C.c_uint32, V, C.c_size_t]),
        "fb_atmosphere_destroy": (None, [V]),
        "fb_atmosphere_save": (S, [V, S, C.c_uint32, C.c_char_p]),
        "fb_atmosphere_load": (S, [V, C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(V)]),
        "fb_atmosphere_export": (S, [V, S, V, C.POINTER(abi.fb_lut_export)]),
        "fb_lut_export_release": (None, [C.POINTER(abi.fb_lut_export)]),
        "fb_render_sky": (
            S,
            [V, V, V, C.POINTER(abi.fb_draw_parameters), V, C.c_uint32, C.c_uint32, V, V,
             C.c_uint32],
        ),
        "fb_render_sky_rays": (
            S,
            [V, V, V, V, C.c_uint32, V, V, V, C.c_uint32, C.c_size_t, V, V],
        ),
        "fb_render_sky_rays_shadowed": (
            S,
            [V, V, V, V, C.c_uint32, V, V, V, C.c_uint32, V, C.c_size_t, V, V],
        ),
        "fb_render_sky_frames": (
            S,
            [V, V, V, V, V, C.c_uint32, C.c_uint32, C.c_uint64, V, V],
        ),
        "fb_ray_direction": (None, [C.c_uint64, C.c_uint32, C.c_uint32, V]),
        "fb_sun_and_sky_irradiance": (
            S,
            [V, V, V, V, V, V, C.c_uint32, C.c_size_t, V, V],
        ),
        "fb_server_create": (S, [V, C.POINTER(abi.fb_server_config), C.POINTER(V)]),
        "fb_server_destroy": (None, [V]),
        "fb_server_submit_precompute": (
            S, [V, P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "fb_server_submit_precomputes": (S, [V, P, C.c_size_t, C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint64)]),
        "fb_server_submit_render": (
            S,
            [V, C.c_uint64, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_uint64,
             C.POINTER(C.c_uint64)],
        ),
        "fb_server_poll": (S, [V, V, C.c_size_t, C.POINTER(C.c_size_t)]),
        "fb_server_drain": (S, [V]),
        "fb_server_download": (S, [V, C.c_uint64, S, V, C.c_size_t]),
        "fb_server_render_result": (S, [V, C.c_uint64, V, C.c_size_t]),
        "fb_server_get_stats": (S, [V, C.POINTER(abi.fb_server_stats)]),
        "fb_server_resident": (S, [V, V, C.c_size_t, C.POINTER(C.c_size_t)]),
        "fb_device_count": (S, [C.POINTER(C.c_int)]),
        "fb_device_alloc": (S, [C.c_size_t, C.POINTER(V)]),
        "fb_device_free": (S, [V]),
        "fb_memcpy_h2d": (S, [V, V, C.c_size_t]),
        "fb_memcpy_d2h": (S, [V, V, C.c_size_t]),
        "fb_memset_d": (S, [V, S, C.c_size_t]),
        "fb_device_copy": (S, [V, V, C.c_size_t, V]),
        "fb_stream_create": (S, [C.POINTER(V)]),
        "fb_stream_destroy": (S, [V]),
        "fb_stream_synchronize": (S, [V]),
        "fb_device_synchronize": (S, []),
        "fb_event_create": (S, [C.POINTER(V)]),
        "fb_event_destroy": (S, [V]),
        "fb_event_record": (S, [V, V]),
        "fb_event_elapsed_ms": (S, [V, V, C.POINTER(C.c_float)]),
        "fb_event_synchronize": (S, [V]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != abi.FB_OK:
        msg = load().fb_last_error()
        raise FuzzyblueError(status, msg.decode() if msg else "")
