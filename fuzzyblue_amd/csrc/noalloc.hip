// noalloc.hip — fb_atmosphere_build_into: the precompute of one atmosphere
// into caller-owned output LUTs with no device allocation.
//
// No reference counterpart.  It serves callers that cannot allocate device
// memory (code under graph capture, serving loops that forbid allocation):
// every transient LUT and per-set table of the pass schedule lives in one
// file-scope device array, `scratch`, of kScratchBytes.  Each GPU holds its
// own copy of the array from the first builder created on it until the
// process ends.  Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU race (enqueue them on one stream, or order them with
// events).  The schedule writes every scratch texel before reading it, so
// nothing needs clearing between calls.
#include <mutex>

#include "runtime.hpp"

namespace {

// 96 MiB: the default LUT sizes (256x64 transmittance, 64x16 irradiance,
// 32x128x32x8 scattering) need 93 MiB of scratch with FP32 storage and
// 61 MiB with binary16 storage.
constexpr size_t kScratchBytes = size_t(96) << 20;
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of each GPU's copy of `scratch` (host code must never
// pass the array's name itself to a kernel or a copy).
std::mutex g_mu;
bool g_done[kMaxDevices];
hipError_t g_err[kMaxDevices];
unsigned char* g_ptr[kMaxDevices];

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

}  // namespace

namespace fbrt {

// Called by fb_builder_create with `device` current.
void noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> g(g_mu);
    if (g_done[device]) return;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) (void)hipGetLastError();
    g_err[device] = e;
    g_ptr[device] = e == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
    g_done[device] = true;
}

}  // namespace fbrt

using namespace fbrt;

extern "C" {

size_t fb_build_into_scratch_limit(void) { return kScratchBytes; }

fb_status fb_atmosphere_build_into(fb_builder* b, fb_stream stream, const fb_parameters* params,
                                   void* transmittance, size_t transmittance_bytes,
                                   void* irradiance, size_t irradiance_bytes, void* scattering,
                                   size_t scattering_bytes) {
    if (!b || !params || !transmittance || !irradiance || !scattering) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    fb_status st = fb_parameters_validate(params);
    if (st != FB_OK) return st;
    const Sizes sz = sizes_of(*params);
    const size_t tb = (params->flags & FB_FLAG_F16_STORAGE) ? 8 : sizeof(float4);
    if (transmittance_bytes < sz.nt * sizeof(float4) || irradiance_bytes < sz.ni * sizeof(float4) ||
        scattering_bytes < sz.ns * tb) {
        set_error("output LUTs need %zu, %zu and %zu bytes", sz.nt * sizeof(float4),
                  sz.ni * sizeof(float4), sz.ns * tb);
        return FB_EINVAL;
    }
    // Scratch layout: the transient LUTs and tables of fb_pending_atmosphere,
    // each 256-byte aligned.
    const size_t bytes[9] = {sz.ni * sizeof(float4),  sz.ns * tb, sz.ns * tb, sz.ns * tb,
                             sz.ns * tb,              sz.nmr * sizeof(float4),
                             sz.nt * sizeof(float4),  sz.ndg * sizeof(float4),
                             sz.nsr * sizeof(float4)};
    size_t off[9], total = 0;
    for (int k = 0; k < 9; ++k) {
        off[k] = total;
        total += align_up(bytes[k]);
    }
    if (total > kScratchBytes) {
        set_error("these LUT sizes need %zu bytes of scratch; fb_atmosphere_build_into holds "
                  "at most %zu (use fb_atmosphere_build)",
                  total, kScratchBytes);
        return FB_EINVAL;
    }
    const int dev = b->device;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d has no scratch", dev);
        return FB_EINVAL;
    }
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (!g_done[dev] || !g_ptr[dev]) {
            set_error("static scratch of device %d unavailable: %s", dev,
                      g_done[dev] ? hipGetErrorString(g_err[dev]) : "not set up");
            return FB_EHIP;
        }
        base = g_ptr[dev];
    }
    FB_HIP_TRY(hipSetDevice(dev));
    fb_atmosphere a;
    a.builder = b;
    a.params = *params;
    a.dev = make_dev_params(*params);
    a.T = Buf{static_cast<float4*>(transmittance), sz.nt * sizeof(float4)};
    a.E = Buf{static_cast<float4*>(irradiance), sz.ni * sizeof(float4)};
    a.S = Buf{static_cast<float4*>(scattering), sz.ns * tb};
    fb_pending_atmosphere pa;
    pa.builder = b;
    pa.atm = &a;
    Buf* slots[9] = {&pa.dE, &pa.dR, &pa.dM, &pa.dMS, &pa.J, &pa.MR, &pa.OD, &pa.DG, &pa.SR};
    for (int k = 0; k < 9; ++k)
        *slots[k] = Buf{reinterpret_cast<float4*>(base + off[k]), bytes[k]};
    st = enqueue_precompute(&a, &pa, static_cast<hipStream_t>(stream));
    pa.atm = nullptr;
    return st;
}

}  // extern "C"
