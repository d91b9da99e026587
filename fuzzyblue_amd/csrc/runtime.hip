// runtime.hip — C ABI implementation: parameters, builder, LUT-set
// lifecycle and the precompute pass schedule.
//
// Reference counterparts (Ralith/fuzzyblue):
//   Parameters + Default ............ src/precompute.rs:655-793, 849-935
//   Builder::new .................... src/precompute.rs:61-587
//   Atmosphere::build ............... src/precompute.rs:1077-2073
//   PendingAtmosphere ............... src/precompute.rs:2108-2212
//   Atmosphere accessors / Drop ..... src/precompute.rs:1059-1073, 2075-2101
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <vector>

#include "runtime.hpp"

namespace fbrt {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

fb_status hip_fail(hipError_t e, const char* what) {
    set_error("%s failed: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? FB_ENOMEM : FB_EHIP;
}

// ------------------------------------------------------------------ Pool
Pool::~Pool() { trim(); }

fb_status Pool::alloc(size_t bytes, void** out) {
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = free_.find(bytes);
        if (it != free_.end()) {
            *out = it->second;
            free_.erase(it);
            cached_ -= bytes;
            in_use_ += bytes;
            return FB_OK;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // Drop cached blocks and retry once.
        trim();
        e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
            return FB_ENOMEM;
        }
    }
    std::lock_guard<std::mutex> g(mu_);
    in_use_ += bytes;
    *out = p;
    return FB_OK;
}

void Pool::release(void* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    free_.emplace(bytes, p);
    in_use_ -= bytes;
    cached_ += bytes;
}

void Pool::trim() {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached_ = 0;
}

void builder_retain(fb_builder* b) { b->refs.fetch_add(1, std::memory_order_relaxed); }

void builder_release(fb_builder* b) {
    if (b->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    b->pool.reset();
    delete b;
}

// ------------------------------------------------------------ parameters
Sizes sizes_of(const fb_parameters& p) {
    Sizes s;
    s.nt = size_t(p.transmittance_mu_size) * p.transmittance_r_size;
    s.ni = size_t(p.irradiance_mu_s_size) * p.irradiance_r_size;
    s.ns = size_t(p.scattering_nu_size) * p.scattering_mu_s_size * p.scattering_mu_size *
           p.scattering_r_size;
    s.nmr = size_t(p.scattering_mu_size) * p.scattering_r_size * fbk::kMsRowFloat4s;
    s.ndg = fbk::dg_count(int(p.scattering_r_size), int(p.scattering_mu_s_size),
                          int(p.scattering_nu_size));
    s.nsr = size_t(p.scattering_mu_size) * p.scattering_r_size * fbk::kSsRowFloat4s;
    return s;
}

fbk::DevParams make_dev_params(const fb_parameters& p) {
    fbk::DevParams d;
    memset(&d, 0, sizeof(d));
    for (int c = 0; c < 3; ++c) {
        d.solar[c] = p.solar_irradiance[c];
        d.rayleigh_scattering[c] = p.rayleigh_scattering[c];
        d.mie_scattering[c] = p.mie_scattering[c];
        d.mie_extinction[c] = p.mie_extinction[c];
        d.albedo[c] = p.ground_albedo[c];
        d.absorption_extinction[c] = p.absorption_extinction[c];
    }
    d.sun_angular_radius = p.sun_angular_radius;
    d.bottom = p.bottom_radius;
    d.top = p.top_radius;
    d.g = p.mie_phase_function_g;
    d.mu_s_min = p.mu_s_min;
    d.flags = int(p.flags);
    d.t_mu = int(p.transmittance_mu_size);
    d.t_r = int(p.transmittance_r_size);
    d.s_r = int(p.scattering_r_size);
    d.s_mu = int(p.scattering_mu_size);
    d.s_mus = int(p.scattering_mu_s_size);
    d.s_nu = int(p.scattering_nu_size);
    d.i_mus = int(p.irradiance_mu_s_size);
    d.i_r = int(p.irradiance_r_size);
    const size_t texels = size_t(p.scattering_r_size) * p.scattering_mu_size *
                          p.scattering_mu_s_size * p.scattering_nu_size;
    d.den_split = texels < fbk::kDenSplitTexels ? fbk::kDenTheta : 1;
    const fb_density_profile* src[3] = {&p.rayleigh_density, &p.mie_density,
                                        &p.absorption_density};
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 2; ++l) {
            d.prof[k][l].width = src[k]->layers[l].width;
            d.prof[k][l].exp_term = src[k]->layers[l].exp_term;
            d.prof[k][l].exp_scale = src[k]->layers[l].exp_scale;
            d.prof[k][l].linear_term = src[k]->layers[l].linear_term;
            d.prof[k][l].constant_term = src[k]->layers[l].constant_term;
        }
    return d;
}

// ----------------------------------------------------------- allocation
static_assert(fbk::kFlagF16Storage == int(FB_FLAG_F16_STORAGE), "storage flag mirror");
// Bytes per texel of the scattering LUT and the 3D scratch LUTs.
static size_t s_texel_bytes(const fb_parameters& p) {
    return (p.flags & FB_FLAG_F16_STORAGE) ? 8 : sizeof(float4);
}
static fb_status alloc_buf(fb_builder* b, size_t texels, Buf* out,
                           size_t texel_bytes = sizeof(float4)) {
    out->bytes = texels * texel_bytes;
    void* p = nullptr;
    fb_status st = b->pool->alloc(out->bytes, &p);
    out->p = static_cast<float4*>(p);
    return st;
}
static void free_buf(fb_builder* b, Buf* buf) {
    if (buf->p) b->pool->release(buf->p, buf->bytes);
    buf->p = nullptr;
    buf->bytes = 0;
}

void free_pending_scratch(fb_pending_atmosphere* pa) {
    fb_builder* b = pa->builder;
    free_buf(b, &pa->dE);
    free_buf(b, &pa->dR);
    free_buf(b, &pa->dM);
    free_buf(b, &pa->dMS);
    free_buf(b, &pa->J);
    free_buf(b, &pa->MR);
    free_buf(b, &pa->OD);
    free_buf(b, &pa->DG);
    free_buf(b, &pa->SR);
}

void free_atmosphere_buffers(fb_atmosphere* a) {
    fb_builder* b = a->builder;
    free_buf(b, &a->T);
    free_buf(b, &a->E);
    free_buf(b, &a->S);
}

static bool event_done(hipEvent_t e) {
    hipError_t q = hipEventQuery(e);
    if (q == hipErrorNotReady) (void)hipGetLastError();
    return q != hipErrorNotReady;  // an error also ends the wait for it
}

fb_status atmosphere_mark_use(fb_atmosphere* a, hipStream_t s) {
    std::lock_guard<std::mutex> g(a->use_mu);
    // Reuse the event of a render that has finished; otherwise add one, so
    // renders in flight on different streams each keep their own event.
    hipEvent_t ev = nullptr;
    for (hipEvent_t e : a->uses)
        if (event_done(e)) {
            ev = e;
            break;
        }
    if (!ev) {
        FB_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        a->uses.push_back(ev);
    }
    FB_HIP_TRY(hipEventRecord(ev, s));
    return FB_OK;
}

bool atmosphere_idle(fb_atmosphere* a) {
    if (!event_done(a->ready)) return false;
    std::lock_guard<std::mutex> g(a->use_mu);
    for (hipEvent_t e : a->uses)
        if (!event_done(e)) return false;
    return true;
}

// A finished atmosphere's outputs (T, E, S) and completion event, without
// the scratch of a build: for LUT sets read from files.
fb_status alloc_atmosphere(fb_builder* b, const fb_parameters& p, fb_atmosphere** out) {
    Sizes s = sizes_of(p);
    auto* a = new fb_atmosphere();
    a->builder = b;
    builder_retain(b);
    a->params = p;
    a->dev = make_dev_params(p);
    fb_status st = alloc_buf(b, s.nt, &a->T);
    if (st == FB_OK) st = alloc_buf(b, s.ni, &a->E);
    if (st == FB_OK) st = alloc_buf(b, s.ns, &a->S, s_texel_bytes(p));
    if (st == FB_OK) {
        hipError_t e = hipEventCreateWithFlags(&a->ready, hipEventDisableTiming);
        if (e != hipSuccess) st = hip_fail(e, "hipEventCreate");
    }
    if (st != FB_OK) {
        free_atmosphere_buffers(a);
        delete a;
        builder_release(b);
        return st;
    }
    *out = a;
    return FB_OK;
}

fb_status alloc_luts(fb_builder* b, const fb_parameters& p, fb_atmosphere** out_a,
                     fb_pending_atmosphere** out_pa) {
    Sizes s = sizes_of(p);
    auto* a = new fb_atmosphere();
    auto* pa = new fb_pending_atmosphere();
    a->builder = b;
    builder_retain(b);
    a->params = p;
    a->dev = make_dev_params(p);
    pa->builder = b;
    builder_retain(b);
    pa->atm = a;
    fb_status st = FB_OK;
    if (st == FB_OK) st = alloc_buf(b, s.nt, &a->T);
    if (st == FB_OK) st = alloc_buf(b, s.ni, &a->E);
    const size_t tb = s_texel_bytes(p);
    if (st == FB_OK) st = alloc_buf(b, s.ns, &a->S, tb);
    if (st == FB_OK) st = alloc_buf(b, s.ni, &pa->dE);
    if (st == FB_OK) st = alloc_buf(b, s.ns, &pa->dR, tb);
    if (st == FB_OK) st = alloc_buf(b, s.ns, &pa->dM, tb);
    if (st == FB_OK) st = alloc_buf(b, s.ns, &pa->dMS, tb);
    if (st == FB_OK) st = alloc_buf(b, s.ns, &pa->J, tb);
    if (st == FB_OK) st = alloc_buf(b, s.nmr, &pa->MR);
    if (st == FB_OK) st = alloc_buf(b, s.nt, &pa->OD);
    if (st == FB_OK) st = alloc_buf(b, s.ndg, &pa->DG);
    if (st == FB_OK) st = alloc_buf(b, s.nsr, &pa->SR);
    hipError_t e = hipSuccess;
    if (st == FB_OK) e = hipEventCreateWithFlags(&pa->done, hipEventDisableTiming);
    if (st == FB_OK && e == hipSuccess) e = hipEventCreateWithFlags(&a->ready, hipEventDisableTiming);
    if (st == FB_OK && e != hipSuccess) st = hip_fail(e, "hipEventCreate");
    if (st != FB_OK) {
        free_pending_scratch(pa);
        free_atmosphere_buffers(a);
        if (pa->done) (void)hipEventDestroy(pa->done);
        if (a->ready) (void)hipEventDestroy(a->ready);
        delete pa;
        delete a;
        builder_release(b);
        builder_release(b);
        return st;
    }
    *out_a = a;
    *out_pa = pa;
    return FB_OK;
}

// ------------------------------------------------------- pass schedule
// src/precompute.rs:1726-1999: transmittance -> direct irradiance -> single
// scattering -> clear irradiance -> for order n in 2..=order:
// density(n), indirect(n-1), multiple.  Stream order replaces the
// reference's pipeline barriers.
PassTrace::~PassTrace() {
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
}

fb_status PassTrace::mark(hipStream_t s, uint32_t pass, uint32_t order) {
    hipEvent_t e = nullptr;
    FB_HIP_TRY(hipEventCreate(&e));
    events.push_back(e);
    FB_HIP_TRY(hipEventRecord(e, s));
    if (events.size() > 1) launches.push_back(fb_pass_timing{pass, order, 0.0f});
    return FB_OK;
}

#define FB_MARK(pass, order)                                              \
    do {                                                                  \
        if (trace) {                                                      \
            fb_status _st = trace->mark(s, (pass), uint32_t(order));      \
            if (_st != FB_OK) return _st;                                 \
        }                                                                 \
    } while (0)

fb_status enqueue_precompute(fb_atmosphere* a, fb_pending_atmosphere* pa, hipStream_t s,
                             PassTrace* trace) {
    const fbk::DevParams& P = a->dev;
    Sizes sz = sizes_of(a->params);
    const fbk::Tuning tu = builder_tuning(a->builder, a->params);
    FB_MARK(0, 0);
    // Precision mode keeps the optical depths beside T (extended-range
    // segment transmittance); FP32 mode follows the reference's quotient.
    float4* OD = (a->params.flags & FB_FLAG_PRECISE) ? pa->OD.p : nullptr;
    FB_HIP_TRY(fbk::launch_transmittance(P, a->T.p, OD, s));
    FB_MARK(FB_PASS_TRANSMITTANCE, 1);
    FB_HIP_TRY(fbk::launch_geometry(P, a->T.p, OD, pa->DG.p, pa->SR.p, pa->MR.p, s));
    FB_MARK(FB_PASS_GEOMETRY, 1);
    FB_HIP_TRY(fbk::launch_direct_irradiance(P, a->T.p, pa->dE.p, s));
    FB_MARK(FB_PASS_DIRECT_IRRADIANCE, 1);
    FB_HIP_TRY(
        fbk::launch_single_scattering(P, a->T.p, pa->SR.p, pa->dR.p, pa->dM.p, a->S.p, tu, s));
    FB_MARK(FB_PASS_SINGLE_SCATTERING, 1);
    FB_HIP_TRY(fbk::launch_fill_zero(a->E.p, sz.ni, s));
    FB_MARK(FB_PASS_CLEAR_IRRADIANCE, 1);
    for (int order = 2; order <= int(a->params.order); ++order) {
        FB_HIP_TRY(fbk::launch_scattering_density(P, order, pa->DG.p, pa->dR.p, pa->dM.p,
                                                  pa->dMS.p, pa->dE.p, pa->J.p, tu, s));
        FB_MARK(FB_PASS_SCATTERING_DENSITY, order);
        FB_HIP_TRY(fbk::launch_indirect_irradiance(P, order - 1, pa->dR.p, pa->dM.p, pa->dMS.p,
                                                   pa->dE.p, a->E.p, tu, s));
        FB_MARK(FB_PASS_INDIRECT_IRRADIANCE, order);
        FB_HIP_TRY(
            fbk::launch_multiple_scattering(P, pa->MR.p, pa->J.p, pa->dMS.p, a->S.p, tu, s));
        FB_MARK(FB_PASS_MULTIPLE_SCATTERING, order);
    }
    return FB_OK;
}

}  // namespace fbrt

using namespace fbrt;

// ====================================================================== ABI
extern "C" {

const char* fb_status_string(fb_status s) {
    switch (s) {
        case FB_OK: return "FB_OK";
        case FB_EINVAL: return "FB_EINVAL: invalid argument";
        case FB_ENOMEM: return "FB_ENOMEM: out of memory";
        case FB_EBUDGET: return "FB_EBUDGET: exceeds memory budget";
        case FB_EHIP: return "FB_EHIP: HIP runtime error";
        case FB_EBUSY: return "FB_EBUSY: object in use by in-flight work";
        case FB_ENOTREADY: return "FB_ENOTREADY: work not complete";
        case FB_ENOENT: return "FB_ENOENT: no such object";
        case FB_ENODEV: return "FB_ENODEV: no HIP device";
    }
    return "unknown fb_status";
}

const char* fb_last_error(void) { return g_last_error.c_str(); }

uint32_t fb_abi_version(void) { return FB_ABI_VERSION; }

// Parameters::default — src/precompute.rs:849-935.
void fb_parameters_default(fb_parameters* out) {
    if (!out) return;
    memset(out, 0, sizeof(*out));
    out->order = 4;
    out->transmittance_mu_size = 256;
    out->transmittance_r_size = 64;
    out->scattering_r_size = 32;
    out->scattering_mu_size = 128;
    out->scattering_mu_s_size = 32;
    out->scattering_nu_size = 8;
    out->irradiance_mu_s_size = 64;
    out->irradiance_r_size = 16;
    out->solar_irradiance[0] = 1.474f;
    out->solar_irradiance[1] = 1.850f;
    out->solar_irradiance[2] = 1.91198f;
    out->sun_angular_radius = 0.004675f;
    out->bottom_radius = 6360.0f;
    out->top_radius = 6420.0f;
    out->rayleigh_density.layers[1].exp_term = 1.0f;
    out->rayleigh_density.layers[1].exp_scale = -0.125f;
    out->rayleigh_scattering[0] = 0.005802f;
    out->rayleigh_scattering[1] = 0.013558f;
    out->rayleigh_scattering[2] = 0.033100f;
    out->mie_density.layers[1].exp_term = 1.0f;
    out->mie_density.layers[1].exp_scale = -0.833333f;
    for (int c = 0; c < 3; ++c) {
        out->mie_scattering[c] = 0.003996f;
        out->mie_extinction[c] = 0.004440f;
        out->ground_albedo[c] = 0.1f;
    }
    out->mie_phase_function_g = 0.8f;
    out->absorption_density.layers[0].width = 25.0f;
    out->absorption_density.layers[0].linear_term = 0.066667f;
    out->absorption_density.layers[0].constant_term = -0.666667f;
    out->absorption_density.layers[1].linear_term = -0.066667f;
    out->absorption_density.layers[1].constant_term = 2.666667f;
    out->absorption_extinction[0] = 6.5e-4f;
    out->absorption_extinction[1] = 1.881e-3f;
    out->absorption_extinction[2] = 8.5e-5f;
    out->mu_s_min = -0.207912f;
    out->flags = 0;
}

fb_status fb_parameters_validate(const fb_parameters* p) {
    if (!p) {
        set_error("parameters pointer is NULL");
        return FB_EINVAL;
    }
    const uint32_t sizes[8] = {p->transmittance_mu_size, p->transmittance_r_size,
                               p->scattering_r_size,     p->scattering_mu_size,
                               p->scattering_mu_s_size,  p->scattering_nu_size,
                               p->irradiance_mu_s_size,  p->irradiance_r_size};
    static const char* names[8] = {"transmittance_mu_size", "transmittance_r_size",
                                   "scattering_r_size",     "scattering_mu_size",
                                   "scattering_mu_s_size",  "scattering_nu_size",
                                   "irradiance_mu_s_size",  "irradiance_r_size"};
    for (int i = 0; i < 8; ++i) {
        // util.h:37 divides by size-1; sizes of 1 are meaningless.
        if (sizes[i] < 2 || sizes[i] > 65536) {
            set_error("%s = %u out of range [2, 65536]", names[i], sizes[i]);
            return FB_EINVAL;
        }
    }
    // scattering.h:36,45 use MU/2 as a texture size (needs MU/2 >= 2).
    if (p->scattering_mu_size < 4) {
        set_error("scattering_mu_size = %u must be >= 4", p->scattering_mu_size);
        return FB_EINVAL;
    }
    if (p->scattering_nu_size > uint32_t(fbk::kMaxNu)) {
        set_error("scattering_nu_size = %u exceeds %d", p->scattering_nu_size, fbk::kMaxNu);
        return FB_EINVAL;
    }
    if (p->scattering_nu_size * p->scattering_mu_s_size > uint32_t(fbk::kMaxRowTexels)) {
        set_error("scattering_nu_size * scattering_mu_s_size = %u exceeds %d",
                  p->scattering_nu_size * p->scattering_mu_s_size, fbk::kMaxRowTexels);
        return FB_EINVAL;
    }
    if (p->irradiance_mu_s_size > uint32_t(fbk::kMaxIrrMuS)) {
        set_error("irradiance_mu_s_size = %u exceeds %d", p->irradiance_mu_s_size,
                  fbk::kMaxIrrMuS);
        return FB_EINVAL;
    }
    uint64_t ns = uint64_t(p->scattering_nu_size) * p->scattering_mu_s_size *
                  p->scattering_mu_size * p->scattering_r_size;
    // The kernels address one scattering-sized LUT, and the transmittance
    // LUT, with 32-bit byte offsets (buffer resources, int texel offsets).
    const uint64_t s_bytes = ns * ((p->flags & FB_FLAG_F16_STORAGE) ? 8u : 16u);
    if (s_bytes >= (uint64_t(1) << 31)) {
        set_error("scattering LUT of %llu texels (%llu bytes) exceeds 2 GiB",
                  (unsigned long long)ns, (unsigned long long)s_bytes);
        return FB_EINVAL;
    }
    const uint64_t t_bytes = uint64_t(p->transmittance_mu_size) * p->transmittance_r_size * 16u;
    if (t_bytes >= (uint64_t(1) << 31)) {
        set_error("transmittance LUT of %llu bytes exceeds 2 GiB", (unsigned long long)t_bytes);
        return FB_EINVAL;
    }
    if (p->order < 1 || p->order > 64) {
        set_error("order = %u out of range [1, 64]", p->order);
        return FB_EINVAL;
    }
    if (!(p->bottom_radius > 0.0f) || !(p->top_radius > p->bottom_radius) ||
        !std::isfinite(p->top_radius)) {
        set_error("need 0 < bottom_radius (%g) < top_radius (%g)", p->bottom_radius,
                  p->top_radius);
        return FB_EINVAL;
    }
    // scattering.h:53-54 divides by mu_s_min.
    if (!(p->mu_s_min < 0.0f) || !(p->mu_s_min >= -1.0f)) {
        set_error("mu_s_min = %g must lie in [-1, 0)", p->mu_s_min);
        return FB_EINVAL;
    }
    if (!(std::fabs(p->mie_phase_function_g) < 1.0f)) {
        set_error("mie_phase_function_g = %g must satisfy |g| < 1", p->mie_phase_function_g);
        return FB_EINVAL;
    }
    if (!(p->sun_angular_radius > 0.0f) || !(p->sun_angular_radius < 0.1f)) {
        set_error("sun_angular_radius = %g must lie in (0, 0.1)", p->sun_angular_radius);
        return FB_EINVAL;
    }
    const float* vec[6] = {p->solar_irradiance, p->rayleigh_scattering, p->mie_scattering,
                           p->mie_extinction, p->absorption_extinction, p->ground_albedo};
    for (int k = 0; k < 6; ++k)
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(vec[k][c]) || vec[k][c] < 0.0f) {
                set_error("coefficients must be finite and non-negative");
                return FB_EINVAL;
            }
    if (p->flags & ~(FB_FLAG_PRECISE | FB_FLAG_F16_STORAGE)) {
        set_error("unknown flags 0x%x", p->flags);
        return FB_EINVAL;
    }
    return FB_OK;
}

// Parameters::*_extent — src/precompute.rs:771-793.
void fb_parameters_extents(const fb_parameters* p, uint32_t transmittance[2],
                           uint32_t irradiance[2], uint32_t scattering[3]) {
    if (!p) return;
    if (transmittance) {
        transmittance[0] = p->transmittance_mu_size;
        transmittance[1] = p->transmittance_r_size;
    }
    if (irradiance) {
        irradiance[0] = p->irradiance_mu_s_size;
        irradiance[1] = p->irradiance_r_size;
    }
    if (scattering) {
        scattering[0] = p->scattering_nu_size * p->scattering_mu_s_size;
        scattering[1] = p->scattering_mu_size;
        scattering[2] = p->scattering_r_size;
    }
}

size_t fb_parameters_footprint(const fb_parameters* p, size_t* resident, size_t* scratch) {
    if (!p) return 0;
    Sizes s = sizes_of(*p);
    const size_t tb = s_texel_bytes(*p);
    size_t res = (s.nt + s.ni) * sizeof(float4) + s.ns * tb;
    size_t scr = (s.nt + s.ni + s.nmr + s.ndg + s.nsr) * sizeof(float4) + 4 * s.ns * tb;
    if (resident) *resident = res;
    if (scratch) *scratch = scr;
    return res + scr;
}

// ---------------------------------------------------------------- builder
fb_status fb_builder_create(int hip_device, fb_builder** out) {
    if (!out) {
        set_error("out is NULL");
        return FB_EINVAL;
    }
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        set_error("no HIP device available (%s)", hipGetErrorString(e));
        return FB_ENODEV;
    }
    if (hip_device < 0 || hip_device >= n) {
        set_error("device %d out of range (%d devices)", hip_device, n);
        return FB_EINVAL;
    }
    FB_HIP_TRY(hipSetDevice(hip_device));
    auto* b = new fb_builder();
    b->device = hip_device;
    e = hipGetDeviceProperties(&b->prop, hip_device);
    if (e == hipSuccess) e = fbk::upload_tables();
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        fb_status st = hip_fail(e, "builder setup");
        delete b;
        return st;
    }
    b->pool = std::make_unique<Pool>(hip_device);
    noalloc_init(hip_device);
    *out = b;
    return FB_OK;
}

void fb_builder_destroy(fb_builder* b) {
    if (!b) return;
    builder_release(b);
}

// ------------------------------------------------------------- precompute
fb_status fb_atmosphere_build(fb_builder* b, fb_stream stream, const fb_parameters* params,
                              fb_pending_atmosphere** out) {
    if (!b || !params || !out) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    *out = nullptr;
    fb_status st = fb_parameters_validate(params);
    if (st != FB_OK) return st;
    FB_HIP_TRY(hipSetDevice(b->device));
    fb_atmosphere* a = nullptr;
    fb_pending_atmosphere* pa = nullptr;
    st = alloc_luts(b, *params, &a, &pa);
    if (st != FB_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    st = enqueue_precompute(a, pa, s);
    if (st == FB_OK) {
        hipError_t e = hipEventRecord(pa->done, s);
        if (e == hipSuccess) e = hipEventRecord(a->ready, s);
        if (e != hipSuccess) st = hip_fail(e, "hipEventRecord");
    }
    if (st != FB_OK) {
        (void)hipStreamSynchronize(s);
        free_pending_scratch(pa);
        free_atmosphere_buffers(a);
        (void)hipEventDestroy(pa->done);
        (void)hipEventDestroy(a->ready);
        delete pa;
        delete a;
        builder_release(b);
        builder_release(b);
        return st;
    }
    *out = pa;
    return FB_OK;
}

fb_status fb_pending_relaunch(fb_pending_atmosphere* pa, fb_stream stream) {
    if (!pa || !pa->atm) {
        set_error("pending atmosphere is NULL or consumed");
        return FB_EINVAL;
    }
    fb_builder* b = pa->builder;
    FB_HIP_TRY(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (pa->batch) return batch_relaunch(pa, s);
    if (!pa->graph) {
        // Capture the pass schedule once on the builder's private stream.
        std::lock_guard<std::mutex> g(b->mu);
        hipGraph_t graph = nullptr;
        FB_HIP_TRY(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
        fb_status st = enqueue_precompute(pa->atm, pa, b->capture_stream);
        hipError_t e = hipStreamEndCapture(b->capture_stream, &graph);
        if (st != FB_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&pa->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            pa->graph = nullptr;
            return hip_fail(e, "hipGraphInstantiate");
        }
    }
    FB_HIP_TRY(hipGraphLaunch(pa->graph, s));
    FB_HIP_TRY(hipEventRecord(pa->done, s));
    FB_HIP_TRY(hipEventRecord(pa->atm->ready, s));
    return FB_OK;
}

fb_status fb_pending_profile(fb_pending_atmosphere* pa, fb_stream stream, fb_pass_timing* out,
                             size_t cap, size_t* n) {
    if (!pa || !pa->atm || !n) {
        set_error("NULL argument or consumed pending atmosphere");
        return FB_EINVAL;
    }
    FB_HIP_TRY(hipSetDevice(pa->builder->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    PassTrace trace;
    fb_status st = enqueue_precompute(pa->atm, pa, s, &trace);
    if (st != FB_OK) return st;
    FB_HIP_TRY(hipEventRecord(pa->done, s));
    FB_HIP_TRY(hipEventRecord(pa->atm->ready, s));
    FB_HIP_TRY(hipStreamSynchronize(s));
    *n = trace.launches.size();
    for (size_t i = 0; i < trace.launches.size(); ++i) {
        float ms = 0.0f;
        FB_HIP_TRY(hipEventElapsedTime(&ms, trace.events[i], trace.events[i + 1]));
        trace.launches[i].ms = ms;
        if (out && i < cap) out[i] = trace.launches[i];
    }
    return FB_OK;
}

const fb_atmosphere* fb_pending_peek(const fb_pending_atmosphere* pa) {
    return pa ? pa->atm : nullptr;
}

fb_status fb_pending_query(const fb_pending_atmosphere* pa) {
    if (!pa) {
        set_error("pending atmosphere is NULL");
        return FB_EINVAL;
    }
    hipError_t e = hipEventQuery(pa->done);
    if (e == hipSuccess) return FB_OK;
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        return FB_ENOTREADY;
    }
    return hip_fail(e, "hipEventQuery");
}

fb_status fb_pending_wait(const fb_pending_atmosphere* pa) {
    if (!pa) {
        set_error("pending atmosphere is NULL");
        return FB_EINVAL;
    }
    FB_HIP_TRY(hipEventSynchronize(pa->done));
    return FB_OK;
}

fb_status fb_pending_stream_wait(const fb_pending_atmosphere* pa, fb_stream consumer) {
    if (!pa) {
        set_error("pending atmosphere is NULL");
        return FB_EINVAL;
    }
    FB_HIP_TRY(hipStreamWaitEvent(static_cast<hipStream_t>(consumer), pa->done, 0));
    return FB_OK;
}

fb_status fb_pending_record_ready_event(const fb_pending_atmosphere* pa, fb_event event) {
    if (!pa || !event) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    fb_builder* b = pa->builder;
    FB_HIP_TRY(hipSetDevice(b->device));
    // The builder's private stream (graph capture also runs on it, under
    // the same lock) waits for the precompute, then records the event.
    std::lock_guard<std::mutex> g(b->mu);
    FB_HIP_TRY(hipStreamWaitEvent(b->capture_stream, pa->done, 0));
    FB_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(event), b->capture_stream));
    return FB_OK;
}

fb_status fb_pending_assert_ready(fb_pending_atmosphere* pa, fb_atmosphere** out) {
    if (!pa || !out || !pa->atm) {
        set_error("NULL argument or consumed pending atmosphere");
        return FB_EINVAL;
    }
    fb_status st = fb_pending_query(pa);
    if (st == FB_ENOTREADY) {
        set_error("precompute still in flight");
        return FB_EBUSY;
    }
    if (st != FB_OK) return st;
    *out = pa->atm;
    pa->atm = nullptr;
    fb_pending_destroy(pa);
    return FB_OK;
}

void fb_pending_destroy(fb_pending_atmosphere* pa) {
    if (!pa) return;
    (void)hipSetDevice(pa->builder->device);
    (void)hipEventSynchronize(pa->done);
    if (pa->batch) batch_detach(pa);
    if (pa->graph) (void)hipGraphExecDestroy(pa->graph);
    free_pending_scratch(pa);
    (void)hipEventDestroy(pa->done);
    if (pa->atm) fb_atmosphere_destroy(pa->atm);
    fb_builder* b = pa->builder;
    delete pa;
    builder_release(b);
}

// --------------------------------------------------------------- accessors
static fb_status view_of(const fb_parameters& p, fb_lut which, const Buf& buf,
                         fb_lut_view* out) {
    out->device_ptr = buf.p;
    out->bytes = buf.bytes;
    out->texel_bytes = uint32_t(sizeof(float4));
    switch (which) {
        case FB_LUT_TRANSMITTANCE:
            out->width = p.transmittance_mu_size;
            out->height = p.transmittance_r_size;
            out->depth = 1;
            break;
        case FB_LUT_IRRADIANCE:
        case FB_LUT_DELTA_IRRADIANCE:
            out->width = p.irradiance_mu_s_size;
            out->height = p.irradiance_r_size;
            out->depth = 1;
            break;
        default:
            out->width = p.scattering_nu_size * p.scattering_mu_s_size;
            out->height = p.scattering_mu_size;
            out->depth = p.scattering_r_size;
            out->texel_bytes = uint32_t(s_texel_bytes(p));
            break;
    }
    return FB_OK;
}

fb_status fb_atmosphere_lut(const fb_atmosphere* a, fb_lut which, fb_lut_view* out) {
    if (!a || !out) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    switch (which) {
        case FB_LUT_TRANSMITTANCE: return view_of(a->params, which, a->T, out);
        case FB_LUT_IRRADIANCE: return view_of(a->params, which, a->E, out);
        case FB_LUT_SCATTERING: return view_of(a->params, which, a->S, out);
        default: set_error("LUT %d is not an atmosphere output", int(which)); return FB_EINVAL;
    }
}

fb_status fb_pending_lut(const fb_pending_atmosphere* pa, fb_lut which, fb_lut_view* out) {
    if (!pa || !out) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    const fb_parameters& p = pa->atm ? pa->atm->params : fb_parameters{};
    if (!pa->atm) {
        set_error("pending atmosphere was consumed");
        return FB_EINVAL;
    }
    switch (which) {
        case FB_LUT_DELTA_IRRADIANCE: return view_of(p, which, pa->dE, out);
        case FB_LUT_DELTA_RAYLEIGH: return view_of(p, which, pa->dR, out);
        case FB_LUT_DELTA_MIE: return view_of(p, which, pa->dM, out);
        case FB_LUT_DELTA_MULTIPLE_SCATTERING: return view_of(p, which, pa->dMS, out);
        case FB_LUT_SCATTERING_DENSITY: return view_of(p, which, pa->J, out);
        default: return fb_atmosphere_lut(pa->atm, which, out);
    }
}

fb_status fb_atmosphere_parameters(const fb_atmosphere* a, fb_parameters* out) {
    if (!a || !out) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    *out = a->params;
    return FB_OK;
}

static uint16_t f32_to_f16(float f) {
    _Float16 h = static_cast<_Float16>(f);
    uint16_t u;
    memcpy(&u, &h, 2);
    return u;
}
static float f16_to_f32(uint16_t u) {
    _Float16 h;
    memcpy(&h, &u, 2);
    return static_cast<float>(h);
}

fb_status fb_atmosphere_download(const fb_atmosphere* a, fb_lut which, uint32_t format,
                                 void* host, size_t bytes) {
    if (!a || !host) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    fb_lut_view v;
    fb_status st = fb_atmosphere_lut(a, which, &v);
    if (st != FB_OK) return st;
    size_t texels = v.bytes / v.texel_bytes;
    size_t need = texels * (format == FB_FORMAT_F16 ? 8 : sizeof(float4));
    if (format != FB_FORMAT_F32 && format != FB_FORMAT_F16) {
        set_error("unknown format %u", format);
        return FB_EINVAL;
    }
    if (bytes < need) {
        set_error("host buffer of %zu bytes is smaller than %zu", bytes, need);
        return FB_EINVAL;
    }
    FB_HIP_TRY(hipSetDevice(a->builder->device));
    FB_HIP_TRY(hipEventSynchronize(a->ready));
    const bool stored_f16 = v.texel_bytes == 8;
    if ((format == FB_FORMAT_F16) == stored_f16) {
        // Stored in the requested format: a plain copy.
        FB_HIP_TRY(hipMemcpy(host, v.device_ptr, v.bytes, hipMemcpyDeviceToHost));
    } else if (stored_f16) {
        std::vector<uint16_t> tmp(texels * 4);
        FB_HIP_TRY(hipMemcpy(tmp.data(), v.device_ptr, v.bytes, hipMemcpyDeviceToHost));
        float* dst = static_cast<float*>(host);
        for (size_t i = 0; i < texels * 4; ++i) dst[i] = f16_to_f32(tmp[i]);
    } else {
        std::vector<float> tmp(texels * 4);
        FB_HIP_TRY(hipMemcpy(tmp.data(), v.device_ptr, v.bytes, hipMemcpyDeviceToHost));
        uint16_t* dst = static_cast<uint16_t*>(host);
        for (size_t i = 0; i < texels * 4; ++i) dst[i] = f32_to_f16(tmp[i]);
    }
    return FB_OK;
}

void fb_atmosphere_destroy(fb_atmosphere* a) {
    if (!a) return;
    (void)hipSetDevice(a->builder->device);
    (void)hipEventSynchronize(a->ready);
    {
        std::lock_guard<std::mutex> g(a->use_mu);
        for (hipEvent_t e : a->uses) {
            (void)hipEventSynchronize(e);
            (void)hipEventDestroy(e);
        }
        a->uses.clear();
    }
    free_atmosphere_buffers(a);
    (void)hipEventDestroy(a->ready);
    fb_builder* b = a->builder;
    delete a;
    builder_release(b);
}

// ----------------------------------------------------------------- plumbing
fb_status fb_device_count(int* out) {
    if (!out) return FB_EINVAL;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *out = 0;
        return FB_OK;
    }
    *out = n;
    return FB_OK;
}
fb_status fb_device_alloc(size_t bytes, void** out) {
    if (!out) return FB_EINVAL;
    FB_HIP_TRY(hipMalloc(out, bytes ? bytes : 1));
    return FB_OK;
}
fb_status fb_device_free(void* p) {
    FB_HIP_TRY(hipFree(p));
    return FB_OK;
}
fb_status fb_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    FB_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return FB_OK;
}
fb_status fb_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    FB_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return FB_OK;
}
fb_status fb_memset_d(void* dst, int value, size_t bytes) {
    FB_HIP_TRY(hipMemset(dst, value, bytes));
    return FB_OK;
}
fb_status fb_device_copy(void* dst, const void* src, size_t bytes, fb_stream stream) {
    FB_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                              static_cast<hipStream_t>(stream)));
    return FB_OK;
}
fb_status fb_stream_create(fb_stream* out) {
    if (!out) return FB_EINVAL;
    hipStream_t s = nullptr;
    FB_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = s;
    return FB_OK;
}
fb_status fb_stream_destroy(fb_stream s) {
    FB_HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(s)));
    return FB_OK;
}
fb_status fb_stream_synchronize(fb_stream s) {
    FB_HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(s)));
    return FB_OK;
}
fb_status fb_device_synchronize(void) {
    FB_HIP_TRY(hipDeviceSynchronize());
    return FB_OK;
}
fb_status fb_event_create(fb_event* out) {
    if (!out) return FB_EINVAL;
    hipEvent_t e = nullptr;
    FB_HIP_TRY(hipEventCreate(&e));
    *out = e;
    return FB_OK;
}
fb_status fb_event_destroy(fb_event e) {
    FB_HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(e)));
    return FB_OK;
}
fb_status fb_event_record(fb_event e, fb_stream s) {
    FB_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(e), static_cast<hipStream_t>(s)));
    return FB_OK;
}
fb_status fb_event_synchronize(fb_event e) {
    FB_HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(e)));
    return FB_OK;
}
fb_status fb_event_elapsed_ms(fb_event start, fb_event stop, float* ms) {
    if (!ms) return FB_EINVAL;
    FB_HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    FB_HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                   static_cast<hipEvent_t>(stop)));
    return FB_OK;
}

}  // extern "C"
