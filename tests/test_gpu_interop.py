"""LUT export (fb_atmosphere_export): the dma-buf descriptor of a finished
LUT, imported again through HIP's external-memory API (the import path a
Vulkan renderer takes with VK_EXT_external_memory_dma_buf), holds the same
bytes as the LUT read back with fb_atmosphere_download, in the layout and
format of the reference's images (src/precompute.rs:1167-1228); the shared
memory outlives both the atmosphere and the library's release of the
export, and later LUT sets do not overwrite it."""
import ctypes as C
import os

import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, Parameters, abi
from fuzzyblue_amd.lib import load
from fuzzyblue_amd.params import DUMP_SIZES

pytestmark = pytest.mark.gpu

HIP_EXT_MEM_OPAQUE_FD = 1  # hipExternalMemoryHandleTypeOpaqueFd
HIP_MEMCPY_D2H = 2


class _Win32(C.Structure):
    _fields_ = [("handle", C.c_void_p), ("name", C.c_void_p)]


class _Handle(C.Union):
    _fields_ = [("fd", C.c_int), ("win32", _Win32), ("nvSciBufObject", C.c_void_p)]


class hipExternalMemoryHandleDesc(C.Structure):
    _fields_ = [("type", C.c_int), ("handle", _Handle), ("size", C.c_ulonglong),
                ("flags", C.c_uint), ("reserved", C.c_uint * 16)]


class hipExternalMemoryBufferDesc(C.Structure):
    _fields_ = [("offset", C.c_ulonglong), ("size", C.c_ulonglong), ("flags", C.c_uint),
                ("reserved", C.c_uint * 16)]


def hip_runtime():
    """The HIP runtime the library itself is linked against (one runtime per
    process)."""
    load()
    paths = {line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64" in line}
    assert len(paths) == 1, paths
    hip = C.CDLL(paths.pop())
    hip.hipImportExternalMemory.argtypes = [C.POINTER(C.c_void_p),
                                            C.POINTER(hipExternalMemoryHandleDesc)]
    hip.hipExternalMemoryGetMappedBuffer.argtypes = [C.POINTER(C.c_void_p), C.c_void_p,
                                                     C.POINTER(hipExternalMemoryBufferDesc)]
    hip.hipDestroyExternalMemory.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


class Imported:
    """A dma-buf descriptor imported and mapped by HIP."""

    def __init__(self, hip, e):
        self.hip = hip
        d = hipExternalMemoryHandleDesc()
        d.type = HIP_EXT_MEM_OPAQUE_FD
        # The import may take ownership of the descriptor: give it a copy.
        d.handle.fd = os.dup(e.fd)
        d.size = e.alloc_bytes
        self.ext = C.c_void_p()
        assert hip.hipImportExternalMemory(C.byref(self.ext), C.byref(d)) == 0
        b = hipExternalMemoryBufferDesc()
        b.offset = e.offset
        b.size = e.bytes
        self.ptr = C.c_void_p()
        assert hip.hipExternalMemoryGetMappedBuffer(C.byref(self.ptr), self.ext, C.byref(b)) == 0
        assert self.ptr.value
        self.bytes = e.bytes

    def read(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.bytes
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.bytes,
                                  HIP_MEMCPY_D2H) == 0
        return out

    def close(self):
        assert self.hip.hipDestroyExternalMemory(self.ext) == 0


def build(builder, p):
    pa = Atmosphere.build(builder, None, p)
    pa.wait()
    return pa.assert_ready()


def shape_of(e):
    return (e.height, e.width, 4) if e.depth == 1 else (e.depth, e.height, e.width, 4)


# Extents that are no powers of two (odd row and slice pitches).
ODD_SIZES = dict(transmittance_mu_size=250, transmittance_r_size=60, scattering_r_size=4,
                 scattering_mu_size=36, scattering_mu_s_size=26, scattering_nu_size=18,
                 irradiance_mu_s_size=60, irradiance_r_size=14)


@pytest.mark.parametrize("sizes", ["dump", "odd"])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_export_reimport_matches_download(builder, storage, sizes):
    hip = hip_runtime()
    flags = abi.FB_FLAG_PRECISE | (abi.FB_FLAG_F16_STORAGE if storage == "f16" else 0)
    p = Parameters().replace(**(DUMP_SIZES if sizes == "dump" else ODD_SIZES), order=3,
                             flags=flags)
    a = build(builder, p)
    f16 = storage == "f16"
    want_ext = {"transmittance": p.transmittance_extent() + (1,),
                "irradiance": p.irradiance_extent() + (1,),
                "scattering": p.scattering_extent()}
    for name in ("transmittance", "irradiance", "scattering"):
        print(f"export {name} {storage} {sizes}", flush=True)
        e = a.export(name)
        try:
            stored16 = name == "scattering" and f16
            tb = 8 if stored16 else 16
            assert e.fd >= 0 and e.offset == 0 and e.alloc_bytes >= e.bytes
            assert (e.width, e.height, e.depth) == want_ext[name]
            assert e.texel_bytes == tb and e.bytes == e.width * e.height * e.depth * tb
            assert e.row_pitch == e.width * tb and e.slice_pitch == e.row_pitch * e.height
            # VkFormat of the reference's images
            assert e.vk_format == (97 if stored16 else 109)
            imp = Imported(hip, e)
            got = imp.read(shape_of(e), np.float16 if stored16 else np.float32)
            imp.close()
            want = a.download(name, fp16=stored16)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), name
            # the library's own mapping of the shared memory holds the same
            mine = np.empty_like(want)
            assert load().fb_memcpy_d2h(mine.ctypes.data_as(C.c_void_p),
                                        C.c_void_p(e.device_ptr), mine.nbytes) == 0
            assert np.array_equal(mine.view(np.uint8), want.view(np.uint8)), name
        finally:
            Atmosphere.release_export(e)
            assert e.fd == -1 and not e.opaque
    a.close()


def test_exported_memory_outlives_atmosphere_and_release(builder):
    """After the atmosphere is destroyed and the library has released the
    export, the importer still reads the texels, and new LUT sets of the same
    sizes do not overwrite them."""
    hip = hip_runtime()
    p = Parameters().replace(**DUMP_SIZES, order=2, flags=abi.FB_FLAG_PRECISE)
    a = build(builder, p)
    print("export scattering, import, release, destroy", flush=True)
    e = a.export("scattering")
    imp = Imported(hip, e)
    first = imp.read(shape_of(e), np.float32)
    assert np.array_equal(first, a.download("scattering"))
    Atmosphere.release_export(e)
    a.close()
    kept = imp.read(shape_of(e), np.float32)
    assert np.array_equal(kept.view(np.uint8), first.view(np.uint8))
    q = p.replace(ground_albedo=[0.4, 0.4, 0.4], order=3)
    others = [build(builder, q) for _ in range(2)]
    assert not np.array_equal(others[0].download("scattering"), first)
    again = imp.read(shape_of(e), np.float32)
    assert np.array_equal(again.view(np.uint8), first.view(np.uint8))
    imp.close()
    for o in others:
        o.close()


def test_export_rejects_scratch_luts(builder):
    p = Parameters().replace(**DUMP_SIZES, order=1, flags=abi.FB_FLAG_PRECISE)
    a = build(builder, p)
    e = abi.fb_lut_export()
    assert load().fb_atmosphere_export(a.handle, abi.FB_LUT_DELTA_MIE, None,
                                       C.byref(e)) == abi.FB_EINVAL
    assert load().fb_atmosphere_export(a.handle, abi.FB_LUT_SCATTERING, None, None) == abi.FB_EINVAL
    a.close()
