// runtime.hpp — internal host-side objects behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fuzzyblue_hip.h"
#include "atmo_device.hpp"
#include "kernels.hpp"

namespace fbrt {

// Thread-local last error (fb_last_error).
void set_error(const char* fmt, ...);
fb_status hip_fail(hipError_t e, const char* what);

#define FB_HIP_TRY(expr)                                 \
    do {                                                 \
        hipError_t _e = (expr);                          \
        if (_e != hipSuccess) return fbrt::hip_fail(_e, #expr); \
    } while (0)

// A caching device allocator.  Blocks are returned to a per-size free list
// and reused; a block is handed back only by owners that have already
// waited for the GPU work using it, so reuse never races.
class Pool {
  public:
    explicit Pool(int device) : device_(device) {}
    int device() const { return device_; }
    ~Pool();
    fb_status alloc(size_t bytes, void** out);
    void release(void* p, size_t bytes);
    void trim();
    size_t in_use() const { return in_use_; }
    size_t cached() const { return cached_; }

  private:
    int device_;
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    size_t in_use_ = 0, cached_ = 0;
};

struct Buf {
    float4* p = nullptr;
    size_t bytes = 0;
};

fbk::DevParams make_dev_params(const fb_parameters& p);

struct Sizes {
    size_t nt, ni, ns;  // texel counts: transmittance, irradiance, scattering
    size_t nmr;         // float4s of the multiple-scattering row table
    size_t ndg;         // float4s of the density geometry table
    size_t nsr;         // float4s of the single-scattering row table
};
Sizes sizes_of(const fb_parameters& p);

}  // namespace fbrt

// Opaque C-ABI objects.
//
// The builder is reference counted like the reference's Arc<Builder>
// (src/precompute.rs:1045-1047, 1077): every atmosphere and pending
// atmosphere holds a reference, fb_builder_destroy drops the caller's, and
// the device state is released with the last one.
namespace fbrt {
// Tuning-table key: the precision flag and the eight LUT sizes.
using TuneKey = std::array<uint32_t, 9>;
TuneKey tune_key(const fb_parameters& p);
}  // namespace fbrt

struct fb_builder {
    int device = 0;
    hipDeviceProp_t prop{};
    std::unique_ptr<fbrt::Pool> pool;
    hipStream_t capture_stream = nullptr;  // private stream for graph capture
    std::mutex mu;
    std::atomic<int> refs{1};
    std::mutex tune_mu;
    std::map<fbrt::TuneKey, fbk::Tuning> tuning;  // per LUT-size class (tune.hip)
};

namespace fbrt {
void builder_retain(fb_builder* b);
void builder_release(fb_builder* b);
// The builder's block sizes for a parameter set (defaults when untuned).
fbk::Tuning builder_tuning(fb_builder* b, const fb_parameters& p);
}  // namespace fbrt

struct fb_atmosphere {
    fb_builder* builder = nullptr;
    fb_parameters params{};
    fbk::DevParams dev{};
    fbrt::Buf T, E, S;
    hipEvent_t ready = nullptr;  // precompute complete
    // Renders enqueued against it, possibly on several streams: one event
    // per render still in flight (events of finished renders are reused).
    // Destroy waits for every one of them before the LUTs go back to the pool.
    std::mutex use_mu;
    std::vector<hipEvent_t> uses;
};

struct fb_pending_atmosphere {
    fb_builder* builder = nullptr;
    fb_atmosphere* atm = nullptr;
    fbrt::Buf dE, dR, dM, dMS, J, MR, OD, DG, SR;
    hipEvent_t done = nullptr;
    hipGraphExec_t graph = nullptr;
    std::shared_ptr<void> batch;  // shared state of a batched build
};

namespace fbrt {
// Events recorded between the launches of a schedule (fb_pending_profile).
struct PassTrace {
    std::vector<hipEvent_t> events;  // events[0] before the first launch
    std::vector<fb_pass_timing> launches;
    fb_status mark(hipStream_t s, uint32_t pass, uint32_t order);
    ~PassTrace();
};
// Enqueue the whole pass schedule of one atmosphere on a stream.
fb_status enqueue_precompute(fb_atmosphere* a, fb_pending_atmosphere* pa, hipStream_t s,
                             PassTrace* trace = nullptr);
fb_status alloc_luts(fb_builder* b, const fb_parameters& p, fb_atmosphere** out_a,
                     fb_pending_atmosphere** out_pa);
void free_pending_scratch(fb_pending_atmosphere* pa);
// Batches (batch.hip).  share_pct < 100: persistent launches holding at
// most that share of the device's resident-block capacity.  direct: launch
// once without a graph and complete each member after its own last pass
// (no relaunch); otherwise capture a graph and complete all members at the
// end.
fb_status build_batch(fb_builder* b, hipStream_t s, const fb_parameters* params, size_t n,
                      fb_pending_atmosphere** out, unsigned share_pct, bool direct);
fb_status batch_relaunch(fb_pending_atmosphere* pa, hipStream_t s);
void batch_detach(fb_pending_atmosphere* pa);
void free_atmosphere_buffers(fb_atmosphere* a);
// Record a render of `a` enqueued on `s` (after its kernels).
fb_status atmosphere_mark_use(fb_atmosphere* a, hipStream_t s);
// True when the precompute and every recorded render of `a` have finished
// (hipEventQuery only: never blocks).
bool atmosphere_idle(fb_atmosphere* a);
// Look up this device's copy of the no-allocation build's static scratch
// (noalloc.hip); called once per device by fb_builder_create.
void noalloc_init(int device);
}  // namespace fbrt
