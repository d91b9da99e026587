// interop.hip — fb_atmosphere_export: a finished LUT handed to another API
// (a Vulkan renderer, another process) as a Linux dma-buf file descriptor.
//
// The reference keeps its LUTs in Vulkan images that its own renderer
// samples (src/precompute.rs:1167-1228: transmittance and irradiance
// R32G32B32A32_SFLOAT 2D, scattering R16G16B16A16_SFLOAT 3D).  Here the LUTs
// live in pooled HIP device memory, and hipMalloc places small blocks inside
// larger driver allocations: a dma-buf of a pooled block would name the
// whole driver allocation, neighbours included, at an offset nobody reports.
// An export therefore gets an allocation of its own, made through HIP's
// virtual-memory API with a POSIX-descriptor handle type (one physical
// allocation, nothing else in it), filled by one device-to-device copy of the
// LUT (texels in the linear x, y, z order), and exported as a dma-buf.  The
// importer maps it (Vulkan: VK_EXT_external_memory_dma_buf, a VkBuffer bound
// at offset 0; HIP: hipImportExternalMemory).  The export has its own
// lifetime: destroying the atmosphere does not touch it, and releasing it
// drops only the library's mapping — the pages live until every importer
// and descriptor has let go.
#include <unistd.h>

#include "runtime.hpp"

using namespace fbrt;

namespace {

// VkFormat values of the reference's LUT images.
constexpr uint32_t kVkR16G16B16A16Sfloat = 97;
constexpr uint32_t kVkR32G32B32A32Sfloat = 109;

struct ExportAlloc {
    int device = 0;
    hipMemGenericAllocationHandle_t handle{};
    bool created = false;
    void* va = nullptr;
    bool mapped = false;
    size_t size = 0;
};

void free_alloc(ExportAlloc* x) {
    if (!x) return;
    (void)hipSetDevice(x->device);
    if (x->mapped) (void)hipMemUnmap(x->va, x->size);
    if (x->va) (void)hipMemAddressFree(x->va, x->size);
    if (x->created) (void)hipMemRelease(x->handle);
    delete x;
}

}  // namespace

extern "C" {

fb_status fb_atmosphere_export(const fb_atmosphere* a, fb_lut which, fb_stream stream,
                               fb_lut_export* out) {
    if (!a || !out) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    fb_lut_view v;
    fb_status st = fb_atmosphere_lut(a, which, &v);
    if (st != FB_OK) return st;
    const int dev = a->builder->device;
    FB_HIP_TRY(hipSetDevice(dev));
    int vmm = 0;
    FB_HIP_TRY(hipDeviceGetAttribute(&vmm, hipDeviceAttributeVirtualMemoryManagementSupported, dev));
    if (!vmm) {
        set_error("device %d has no virtual memory management (needed for exportable memory)",
                  dev);
        return FB_EHIP;
    }
    hipMemAllocationProp prop{};
    prop.type = hipMemAllocationTypePinned;
    prop.requestedHandleType = hipMemHandleTypePosixFileDescriptor;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = dev;
    size_t gran = 0;
    FB_HIP_TRY(hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum));
    if (gran == 0) gran = 4096;
    auto* x = new ExportAlloc();
    x->device = dev;
    x->size = (v.bytes + gran - 1) / gran * gran;
    auto fail = [&](hipError_t e, const char* what) {
        free_alloc(x);
        return hip_fail(e, what);
    };
    hipError_t e = hipMemCreate(&x->handle, x->size, &prop, 0);
    if (e != hipSuccess) return fail(e, "hipMemCreate");
    x->created = true;
    e = hipMemAddressReserve(&x->va, x->size, 0, nullptr, 0);
    if (e != hipSuccess) {
        x->va = nullptr;
        return fail(e, "hipMemAddressReserve");
    }
    e = hipMemMap(x->va, x->size, 0, x->handle, 0);
    if (e != hipSuccess) return fail(e, "hipMemMap");
    x->mapped = true;
    hipMemAccessDesc acc{};
    acc.location = prop.location;
    acc.flags = hipMemAccessFlagsProtReadWrite;
    e = hipMemSetAccess(x->va, x->size, &acc, 1);
    if (e != hipSuccess) return fail(e, "hipMemSetAccess");
    // The copy waits for the precompute (or upload) on the caller's stream,
    // then the host waits for the copy: the importer sees final texels.
    hipStream_t s = static_cast<hipStream_t>(stream);
    e = hipStreamWaitEvent(s, a->ready, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(x->va, v.device_ptr, v.bytes, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && x->size > v.bytes)
        e = hipMemsetAsync(static_cast<char*>(x->va) + v.bytes, 0, x->size - v.bytes, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(e, "copy into the exported allocation");
    // The descriptor is written into a zeroed 64-bit slot, so the call may
    // store an int or a pointer-sized value.
    uint64_t slot = 0;
    e = hipMemExportToShareableHandle(&slot, x->handle, hipMemHandleTypePosixFileDescriptor, 0);
    if (e != hipSuccess) return fail(e, "hipMemExportToShareableHandle");
    const int fd = static_cast<int>(slot & 0xffffffffu);
    if (fd < 0) {
        free_alloc(x);
        set_error("hipMemExportToShareableHandle returned no descriptor");
        return FB_EHIP;
    }
    out->fd = fd;
    out->width = v.width;
    out->height = v.height;
    out->depth = v.depth;
    out->texel_bytes = v.texel_bytes;
    out->vk_format = v.texel_bytes == 8 ? kVkR16G16B16A16Sfloat : kVkR32G32B32A32Sfloat;
    out->bytes = v.bytes;
    out->alloc_bytes = x->size;
    out->offset = 0;
    out->row_pitch = size_t(v.width) * v.texel_bytes;
    out->slice_pitch = out->row_pitch * v.height;
    out->device_ptr = x->va;
    out->opaque = x;
    return FB_OK;
}

void fb_lut_export_release(fb_lut_export* e) {
    if (!e) return;
    if (e->fd >= 0) (void)close(e->fd);
    e->fd = -1;
    free_alloc(static_cast<ExportAlloc*>(e->opaque));
    e->opaque = nullptr;
    e->device_ptr = nullptr;
}

}  // extern "C"
