"""Python mirror of the reference's public API over the C ABI.

``Builder``, ``Atmosphere``, ``PendingAtmosphere``, ``Renderer`` and
``DrawParameters`` follow src/lib.rs:8-12 (Parameters lives in
:mod:`fuzzyblue_amd.params`).  Work is enqueued on HIP streams instead of
recorded into Vulkan command buffers; readiness is tracked by HIP events.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import abi
from .lib import check, load
from .params import Parameters

LUT_NAMES = {
    "transmittance": abi.FB_LUT_TRANSMITTANCE,
    "irradiance": abi.FB_LUT_IRRADIANCE,
    "scattering": abi.FB_LUT_SCATTERING,
    "delta_irradiance": abi.FB_LUT_DELTA_IRRADIANCE,
    "delta_rayleigh": abi.FB_LUT_DELTA_RAYLEIGH,
    "delta_mie": abi.FB_LUT_DELTA_MIE,
    "delta_multiple_scattering": abi.FB_LUT_DELTA_MULTIPLE_SCATTERING,
    "scattering_density": abi.FB_LUT_SCATTERING_DENSITY,
}


def _lut_id(which) -> int:
    return LUT_NAMES[which] if isinstance(which, str) else int(which)


class Stream:
    """A non-blocking HIP stream (None = the default stream)."""

    def __init__(self):
        self.handle = C.c_void_p()
        check(load().fb_stream_create(C.byref(self.handle)))

    def synchronize(self):
        check(load().fb_stream_synchronize(self.handle))

    def close(self):
        if self.handle:
            check(load().fb_stream_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_handle(stream) -> Optional[int]:
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class DeviceBuffer:
    """Plain device allocation (for render targets and ray inputs)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(load().fb_device_alloc(self.nbytes, C.byref(self.ptr)))

    @staticmethod
    def from_array(a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = DeviceBuffer(a.nbytes)
        check(load().fb_memcpy_h2d(b.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return b

    def to_array(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(load().fb_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def zero(self):
        check(load().fb_memset_d(self.ptr, 0, self.nbytes))

    def close(self):
        if self.ptr:
            load().fb_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Builder:
    """Builder::new (src/precompute.rs:61-68): binds a HIP device."""

    def __init__(self, device: int = 0):
        self.handle = C.c_void_p()
        check(load().fb_builder_create(int(device), C.byref(self.handle)))

    def autotune(self, params: Parameters, stream=None) -> dict:
        """Time the candidate block sizes for params' LUT-size class and keep
        the fastest (fb_builder_autotune)."""
        t = abi.fb_tuning()
        c = params.to_c()
        check(load().fb_builder_autotune(self.handle, _stream_handle(stream), C.byref(c),
                                         C.byref(t)))
        return t.as_dict()

    def tuning(self, params: Parameters):
        """(tuned, block sizes) of params' size class (fb_builder_get_tuning)."""
        t = abi.fb_tuning()
        c = params.to_c()
        st = load().fb_builder_get_tuning(self.handle, C.byref(c), C.byref(t))
        if st not in (abi.FB_OK, abi.FB_ENOENT):
            check(st)
        return st == abi.FB_OK, t.as_dict()

    def set_tuning(self, params: Parameters, **sizes):
        """Install block sizes (fb_tuning field names; omitted = default)."""
        t = abi.fb_tuning()
        for k, v in sizes.items():
            if k == "best_ms" or not hasattr(t, k):
                raise AttributeError(k)
            setattr(t, k, int(v))
        c = params.to_c()
        check(load().fb_builder_set_tuning(self.handle, C.byref(c), C.byref(t)))

    def close(self):
        if self.handle:
            load().fb_builder_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Atmosphere:
    """A finished LUT set (src/precompute.rs:1045-1057, accessors :2075-2101)."""

    def __init__(self, handle, builder: Builder, owned: bool):
        self.handle = handle
        self.builder = builder
        self._owned = owned

    @staticmethod
    def build(builder: Builder, stream, params: Parameters) -> "PendingAtmosphere":
        """Atmosphere::build (src/precompute.rs:1077-1081): enqueue only."""
        c = params.to_c() if isinstance(params, Parameters) else params
        out = C.c_void_p()
        check(load().fb_atmosphere_build(builder.handle, _stream_handle(stream), C.byref(c),
                                         C.byref(out)))
        return PendingAtmosphere(out, builder)

    @staticmethod
    def build_batch(builder: Builder, stream, params: Sequence[Parameters]) -> List["PendingAtmosphere"]:
        n = len(params)
        arr = (abi.fb_parameters * n)(*[p.to_c() for p in params])
        outs = (C.c_void_p * n)()
        check(load().fb_atmosphere_build_batch(builder.handle, _stream_handle(stream), arr, n,
                                               outs))
        return [PendingAtmosphere(C.c_void_p(outs[i]), builder) for i in range(n)]

    def lut(self, which) -> abi.fb_lut_view:
        v = abi.fb_lut_view()
        check(load().fb_atmosphere_lut(self.handle, _lut_id(which), C.byref(v)))
        return v

    def transmittance(self):
        return self.lut("transmittance")

    def irradiance(self):
        return self.lut("irradiance")

    def scattering(self):
        return self.lut("scattering")

    def transmittance_extent(self):
        v = self.lut("transmittance")
        return (v.width, v.height)

    def irradiance_extent(self):
        v = self.lut("irradiance")
        return (v.width, v.height)

    def scattering_extent(self):
        v = self.lut("scattering")
        return (v.width, v.height, v.depth)

    def parameters(self) -> Parameters:
        c = abi.fb_parameters()
        check(load().fb_atmosphere_parameters(self.handle, C.byref(c)))
        return Parameters.from_c(c)

    def download(self, which, fp16: bool = False) -> np.ndarray:
        """Linear readback as in examples/dump.rs:150-193: (H, W, 4) for 2D
        LUTs, (D, H, W, 4) for the scattering LUT."""
        v = self.lut(which)
        shape = (v.height, v.width, 4) if v.depth == 1 else (v.depth, v.height, v.width, 4)
        out = np.empty(shape, dtype=np.float16 if fp16 else np.float32)
        check(load().fb_atmosphere_download(
            self.handle, _lut_id(which), abi.FB_FORMAT_F16 if fp16 else abi.FB_FORMAT_F32,
            out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def export(self, which, stream=None) -> abi.fb_lut_export:
        """A copy of the LUT in memory exportable as a dma-buf descriptor
        (fb_atmosphere_export); release it with release_export()."""
        e = abi.fb_lut_export()
        check(load().fb_atmosphere_export(self.handle, _lut_id(which), _stream_handle(stream),
                                          C.byref(e)))
        return e

    @staticmethod
    def release_export(e: abi.fb_lut_export):
        load().fb_lut_export_release(C.byref(e))

    def save(self, which, path: str, fp16: bool = False):
        """Write one LUT as an OpenEXR image laid out like the reference's
        dump (examples/dump.rs:331-367); fp16 writes the scattering LUT as
        half (the dump's format)."""
        check(load().fb_atmosphere_save(
            self.handle, _lut_id(which), abi.FB_FORMAT_F16 if fp16 else abi.FB_FORMAT_F32,
            str(path).encode()))

    def save_all(self, prefix: str, fp16: bool = False) -> List[str]:
        """prefix + "transmittance.exr", "irradiance.exr", "scattering.exr"."""
        paths = [f"{prefix}{n}.exr" for n in ("transmittance", "irradiance", "scattering")]
        for name, path in zip(("transmittance", "irradiance", "scattering"), paths):
            self.save(name, path, fp16)
        return paths

    @staticmethod
    def load(builder: Builder, transmittance: str, irradiance: str,
             scattering: str) -> "Atmosphere":
        """A finished atmosphere from files written by save()."""
        out = C.c_void_p()
        check(load().fb_atmosphere_load(builder.handle, str(transmittance).encode(),
                                        str(irradiance).encode(), str(scattering).encode(),
                                        C.byref(out)))
        return Atmosphere(out, builder, owned=True)

    def close(self):
        if self.handle and self._owned:
            load().fb_atmosphere_destroy(self.handle)
        self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PendingAtmosphere:
    """PendingAtmosphere (src/precompute.rs:2108-2212)."""

    def __init__(self, handle, builder: Builder):
        self.handle = handle
        self.builder = builder

    def atmosphere(self) -> Atmosphere:
        """Access the atmosphere while it may not be ready (:2204)."""
        h = load().fb_pending_peek(self.handle)
        return Atmosphere(C.c_void_p(h), self.builder, owned=False)

    def is_ready(self) -> bool:
        st = load().fb_pending_query(self.handle)
        if st == abi.FB_ENOTREADY:
            return False
        check(st)
        return True

    def wait(self):
        check(load().fb_pending_wait(self.handle))

    def acquire_ownership(self, stream):
        """Make `stream` wait for the precompute (:2142-2201)."""
        check(load().fb_pending_stream_wait(self.handle, _stream_handle(stream)))

    def record_ready_event(self, event):
        """Make an fb_event (a hipEvent_t handle) complete once the
        precompute has completed (acquire_ownership, :2147-2152)."""
        check(load().fb_pending_record_ready_event(self.handle, event))

    def relaunch(self, stream=None):
        check(load().fb_pending_relaunch(self.handle, _stream_handle(stream)))

    def profile(self, stream=None):
        """Per-launch GPU times of one re-run: list of (pass, order, ms)."""
        cap = 8 + 3 * 64
        arr = (abi.fb_pass_timing * cap)()
        n = C.c_size_t()
        check(load().fb_pending_profile(self.handle, _stream_handle(stream), arr, cap,
                                        C.byref(n)))
        return [(abi.FB_PASS_NAMES[arr[i].pass_], arr[i].order, arr[i].ms)
                for i in range(min(n.value, cap))]

    def scratch(self, which) -> abi.fb_lut_view:
        v = abi.fb_lut_view()
        check(load().fb_pending_lut(self.handle, _lut_id(which), C.byref(v)))
        return v

    def download_scratch(self, which) -> np.ndarray:
        self.wait()
        v = self.scratch(which)
        shape = (v.height, v.width, 4) if v.depth == 1 else (v.depth, v.height, v.width, 4)
        # Binary16 texels (FB_FLAG_F16_STORAGE) are widened to float32.
        out = np.empty(shape, dtype=np.float16 if v.texel_bytes == 8 else np.float32)
        check(load().fb_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(v.device_ptr),
                                   out.nbytes))
        return out.astype(np.float32, copy=False)

    def assert_ready(self) -> Atmosphere:
        """Consume; the scratch LUTs are released (:2209)."""
        out = C.c_void_p()
        check(load().fb_pending_assert_ready(self.handle, C.byref(out)))
        self.handle = C.c_void_p()
        return Atmosphere(out, self.builder, owned=True)

    def close(self):
        if self.handle:
            load().fb_pending_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class DrawParameters:
    """DrawParameters (src/render.rs:243-252); inverse_viewproj[col][row]."""

    inverse_viewproj: List[List[float]] = field(
        default_factory=lambda: [[1.0 if i == j else 0.0 for j in range(4)] for i in range(4)]
    )
    camera_position: List[float] = field(default_factory=lambda: [0.0, 0.0, 6361.0])
    sun_direction: List[float] = field(default_factory=lambda: [0.0, 0.0, 1.0])

    def to_c(self) -> abi.fb_draw_parameters:
        c = abi.fb_draw_parameters()
        for i in range(4):
            for j in range(4):
                c.inverse_viewproj[i][j] = float(self.inverse_viewproj[i][j])
        for i in range(3):
            c.camera_position[i] = float(self.camera_position[i])
            c.sun_direction[i] = float(self.sun_direction[i])
        return c


class Renderer:
    """Renderer (src/render.rs:12-236): per-pixel sky over a render target."""

    def __init__(self, builder: Builder):
        self.builder = builder

    def draw(self, stream, atmosphere: Atmosphere, params: DrawParameters, width: int,
             height: int, color: DeviceBuffer, transmittance: Optional[DeviceBuffer] = None,
             depth: Optional[DeviceBuffer] = None, composite: bool = False):
        c = params.to_c()
        check(load().fb_render_sky(
            self.builder.handle, _stream_handle(stream), atmosphere.handle, C.byref(c),
            depth.ptr if depth is not None else None, int(width), int(height), color.ptr,
            transmittance.ptr if transmittance is not None else None,
            abi.FB_RENDER_COMPOSITE if composite else 0))

    def rays(self, stream, atmosphere: Atmosphere, cameras: DeviceBuffer, camera_stride: int,
             views: DeviceBuffer, suns: DeviceBuffer, sun_stride: int, n: int,
             radiance: DeviceBuffer, transmittance: Optional[DeviceBuffer] = None,
             points: Optional[DeviceBuffer] = None,
             shadow_lengths: Optional[DeviceBuffer] = None):
        """Per-ray lookups; shadow_lengths (float per ray, km) adds light
        shafts (fb_render_sky_rays_shadowed)."""
        pts = points.ptr if points is not None else None
        tr = transmittance.ptr if transmittance is not None else None
        if shadow_lengths is None:
            check(load().fb_render_sky_rays(
                self.builder.handle, _stream_handle(stream), atmosphere.handle, cameras.ptr,
                camera_stride, views.ptr, pts, suns.ptr, sun_stride, n, radiance.ptr, tr))
        else:
            check(load().fb_render_sky_rays_shadowed(
                self.builder.handle, _stream_handle(stream), atmosphere.handle, cameras.ptr,
                camera_stride, views.ptr, pts, suns.ptr, sun_stride, shadow_lengths.ptr, n,
                radiance.ptr, tr))

    def frames(self, stream, atmosphere: Atmosphere, altitudes: DeviceBuffer,
               zeniths: DeviceBuffer, n_frames: int, pixels: int, seed: int,
               radiance: DeviceBuffer, transmittance: Optional[DeviceBuffer] = None):
        check(load().fb_render_sky_frames(
            self.builder.handle, _stream_handle(stream), atmosphere.handle, altitudes.ptr,
            zeniths.ptr, n_frames, pixels, seed, radiance.ptr,
            transmittance.ptr if transmittance is not None else None))

    def sun_and_sky_irradiance(self, stream, atmosphere: Atmosphere, points: DeviceBuffer,
                               normals: DeviceBuffer, suns: DeviceBuffer, sun_stride: int,
                               n: int, direct: DeviceBuffer, sky: DeviceBuffer):
        check(load().fb_sun_and_sky_irradiance(
            self.builder.handle, _stream_handle(stream), atmosphere.handle, points.ptr,
            normals.ptr, suns.ptr, sun_stride, n, direct.ptr, sky.ptr))


def ray_direction(seed: int, frame: int, pixel: int) -> np.ndarray:
    out = (C.c_float * 3)()
    load().fb_ray_direction(seed, frame, pixel, out)
    return np.array(out[:], dtype=np.float32)


def device_count() -> int:
    n = C.c_int(0)
    check(load().fb_device_count(C.byref(n)))
    return n.value


def synchronize():
    check(load().fb_device_synchronize())
