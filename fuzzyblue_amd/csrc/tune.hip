// tune.hip — block-size autotuner for the precompute passes.
//
// The reference dispatches every pass with fixed 8x8 / 4x4x4 workgroups
// (shaders/*.comp local_size, src/precompute.rs:1742-1744).  Here the block
// shapes of the four row/tile passes are chosen per LUT-size class by
// timing each candidate on the device: single and multiple scattering
// (threads per (r, mu) row), scattering density (threads per (r, mu_s)
// tile, order 2 and order >= 3 separately) and indirect irradiance
// (wavefronts per block).  A candidate changes only which thread evaluates
// which texel, never the order of a texel's sums, so every choice gives
// bit-identical LUTs; plain and batched builds read the same table.
#include <string.h>

#include "runtime.hpp"

namespace fbrt {

TuneKey tune_key(const fb_parameters& p) {
    return TuneKey{p.flags & (FB_FLAG_PRECISE | FB_FLAG_F16_STORAGE), p.transmittance_mu_size, p.transmittance_r_size,
                   p.scattering_r_size,       p.scattering_mu_size,    p.scattering_mu_s_size,
                   p.scattering_nu_size,      p.irradiance_mu_s_size,  p.irradiance_r_size};
}

fbk::Tuning builder_tuning(fb_builder* b, const fb_parameters& p) {
    std::lock_guard<std::mutex> g(b->tune_mu);
    auto it = b->tuning.find(tune_key(p));
    return it == b->tuning.end() ? fbk::Tuning{} : it->second;
}

}  // namespace fbrt

using namespace fbrt;

namespace {

// The resolved block sizes of a tuning (defaults filled in).
fb_tuning to_public(const fbk::DevParams& P, const fbk::Tuning& t) {
    fb_tuning o;
    memset(&o, 0, sizeof(o));
    o.single_scattering_threads = fbk::tuned_threads(P, fbk::kPassSingle, 1, t);
    o.density_order2_threads = fbk::tuned_threads(P, fbk::kPassDensity, 2, t);
    o.density_threads = fbk::tuned_threads(P, fbk::kPassDensity, 3, t);
    o.indirect_order1_waves = fbk::tuned_threads(P, fbk::kPassIndirect, 2, t) / 64;
    o.indirect_waves = fbk::tuned_threads(P, fbk::kPassIndirect, 3, t) / 64;
    o.multiple_scattering_threads = fbk::tuned_threads(P, fbk::kPassMultiple, 2, t);
    o.xcd_remap = t.remap;
    return o;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

extern "C" {

fb_status fb_builder_autotune(fb_builder* b, fb_stream stream, const fb_parameters* p,
                              fb_tuning* out) {
    if (!b || !p) {
        set_error("NULL builder or parameters");
        return FB_EINVAL;
    }
    fb_status st = fb_parameters_validate(p);
    if (st != FB_OK) return st;
    FB_HIP_TRY(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);

    // A scratch LUT set of these sizes, precomputed once so that every pass
    // reads the data it would read in a real build.
    fb_atmosphere* a = nullptr;
    fb_pending_atmosphere* pa = nullptr;
    if ((st = alloc_luts(b, *p, &a, &pa)) != FB_OK) return st;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);
        fb_pending_destroy(pa);  // also releases the unconsumed atmosphere
    };
    if ((st = enqueue_precompute(a, pa, s)) != FB_OK) {
        cleanup();
        return st;
    }
    const fbk::DevParams P = a->dev;  // a copy: the set is released before the result
    EventPair ev;
    hipError_t e = hipEventCreate(&ev.a);
    if (e == hipSuccess) e = hipEventCreate(&ev.b);
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "hipEventCreate");
    }

    constexpr int kWarm = 2, kReps = 5;
    // (Only multiple scattering accepts more than 256 threads.)
    const unsigned candidates[] = {64, 128, 256, 512, 1024};
    fbk::Tuning best = builder_tuning(b, *p);
    float best_ms[6] = {0, 0, 0, 0, 0, 0};

    // Mean ms per launch of `launch(t)` after warm-up.
    auto time = [&](const fbk::Tuning& t, auto launch, float* ms) -> fb_status {
        for (int k = 0; k < kWarm; ++k) FB_HIP_TRY(launch(t));
        FB_HIP_TRY(hipEventRecord(ev.a, s));
        for (int k = 0; k < kReps; ++k) FB_HIP_TRY(launch(t));
        FB_HIP_TRY(hipEventRecord(ev.b, s));
        FB_HIP_TRY(hipEventSynchronize(ev.b));
        FB_HIP_TRY(hipEventElapsedTime(ms, ev.a, ev.b));
        *ms /= kReps;
        return FB_OK;
    };
    // Time each valid block size of one knob, keep the fastest in
    // best.*field; then, for the passes with a block-order choice
    // (`remap_pass` >= 0), time the XCD-contiguous order against the plain one.
    auto tune = [&](int knob, int pass, int order, unsigned fbk::Tuning::*field, unsigned unit,
                    int remap_pass, auto launch) -> fb_status {
        float best_t = -1.0f;
        unsigned best_v = 0;
        for (unsigned c : candidates) {
            if (!fbk::tuning_valid(P, pass, order, c)) continue;
            fbk::Tuning t = best;
            t.*field = c / unit;
            float ms = 0.0f;
            fb_status st2 = time(t, launch, &ms);
            if (st2 != FB_OK) return st2;
            if (best_t < 0.0f || ms < best_t) {
                best_t = ms;
                best_v = c / unit;
            }
        }
        if (!best_v) return FB_OK;
        best.*field = best_v;
        if (remap_pass >= 0) {
            const unsigned bit = 1u << remap_pass;
            fbk::Tuning t = best;
            t.remap = best.remap ^ bit;
            float ms = 0.0f;
            fb_status st2 = time(t, launch, &ms);
            if (st2 != FB_OK) return st2;
            if (ms < best_t) {
                best = t;
                best_t = ms;
            }
        }
        best_ms[knob] = best_t;
        return FB_OK;
    };
    float4 *T = a->T.p, *E = a->E.p, *S = a->S.p;
    float4 *dE = pa->dE.p, *dR = pa->dR.p, *dM = pa->dM.p, *dMS = pa->dMS.p, *J = pa->J.p,
           *MR = pa->MR.p;
    const float4* DG = pa->DG.p;
    const float4* SR = pa->SR.p;
    // (The repeated launches rewrite the scratch set's LUTs, and the
    // indirect and multiple passes accumulate into E and S; the set is
    // discarded afterwards.)
    st = tune(0, fbk::kPassSingle, 1, &fbk::Tuning::single, 1, fbk::kPassSingle,
              [&](const fbk::Tuning& t) {
        return fbk::launch_single_scattering(P, T, SR, dR, dM, S, t, s);
    });
    if (st == FB_OK)
        st = tune(1, fbk::kPassDensity, 2, &fbk::Tuning::density2, 1, -1,
                  [&](const fbk::Tuning& t) {
            return fbk::launch_scattering_density(P, 2, DG, dR, dM, dMS, dE, J, t, s);
        });
    if (st == FB_OK)
        st = tune(2, fbk::kPassDensity, 3, &fbk::Tuning::density, 1, fbk::kPassDensity,
                  [&](const fbk::Tuning& t) {
            return fbk::launch_scattering_density(P, 3, DG, dR, dM, dMS, dE, J, t, s);
        });
    if (st == FB_OK)
        st = tune(3, fbk::kPassIndirect, 2, &fbk::Tuning::indirect1, 64, -1,
                  [&](const fbk::Tuning& t) {
                      return fbk::launch_indirect_irradiance(P, 1, dR, dM, dMS, dE, E, t, s);
                  });
    if (st == FB_OK)
        st = tune(4, fbk::kPassIndirect, 3, &fbk::Tuning::indirect, 64, -1,
                  [&](const fbk::Tuning& t) {
                      return fbk::launch_indirect_irradiance(P, 2, dR, dM, dMS, dE, E, t, s);
                  });
    if (st == FB_OK)
        st = tune(5, fbk::kPassMultiple, 2, &fbk::Tuning::multiple, 1, fbk::kPassMultiple,
                  [&](const fbk::Tuning& t) {
            return fbk::launch_multiple_scattering(P, MR, J, dMS, S, t, s);
        });
    cleanup();
    if (st != FB_OK) return st;
    {
        std::lock_guard<std::mutex> g(b->tune_mu);
        b->tuning[tune_key(*p)] = best;
    }
    if (out) {
        *out = to_public(P, best);
        memcpy(out->best_ms, best_ms, sizeof(best_ms));
    }
    return FB_OK;
}

fb_status fb_builder_get_tuning(fb_builder* b, const fb_parameters* p, fb_tuning* out) {
    if (!b || !p || !out) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    fb_status st = fb_parameters_validate(p);
    if (st != FB_OK) return st;
    bool found;
    fbk::Tuning t;
    {
        std::lock_guard<std::mutex> g(b->tune_mu);
        auto it = b->tuning.find(tune_key(*p));
        found = it != b->tuning.end();
        if (found) t = it->second;
    }
    *out = to_public(make_dev_params(*p), t);
    if (!found) {
        set_error("no tuning for these LUT sizes: defaults returned");
        return FB_ENOENT;
    }
    return FB_OK;
}

fb_status fb_builder_set_tuning(fb_builder* b, const fb_parameters* p, const fb_tuning* t) {
    if (!b || !p || !t) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    fb_status st = fb_parameters_validate(p);
    if (st != FB_OK) return st;
    const fbk::DevParams P = make_dev_params(*p);
    struct Knob {
        uint32_t v;
        int pass, order;
        unsigned unit;
        unsigned fbk::Tuning::*field;
    } knobs[] = {
        {t->single_scattering_threads, fbk::kPassSingle, 1, 1, &fbk::Tuning::single},
        {t->density_order2_threads, fbk::kPassDensity, 2, 1, &fbk::Tuning::density2},
        {t->density_threads, fbk::kPassDensity, 3, 1, &fbk::Tuning::density},
        {t->indirect_order1_waves, fbk::kPassIndirect, 2, 64, &fbk::Tuning::indirect1},
        {t->indirect_waves, fbk::kPassIndirect, 3, 64, &fbk::Tuning::indirect},
        {t->multiple_scattering_threads, fbk::kPassMultiple, 2, 1, &fbk::Tuning::multiple},
    };
    fbk::Tuning tu;
    const uint32_t remappable = (1u << fbk::kPassSingle) | (1u << fbk::kPassDensity) |
                                (1u << fbk::kPassMultiple);
    if (t->xcd_remap & ~remappable) {
        set_error("xcd_remap bits other than single, density and multiple scattering");
        return FB_EINVAL;
    }
    tu.remap = t->xcd_remap;
    for (const Knob& k : knobs) {
        if (k.v == 0) continue;  // default
        if (k.v > 1024 || !fbk::tuning_valid(P, k.pass, k.order, k.v * k.unit)) {
            set_error("block size %u is not valid for pass %d at these sizes", k.v, k.pass);
            return FB_EINVAL;
        }
        tu.*k.field = k.v;
    }
    std::lock_guard<std::mutex> g(b->tune_mu);
    b->tuning[tune_key(*p)] = tu;
    return FB_OK;
}

}  // extern "C"
