// batch.hip — many atmospheres in one build call.
#include "runtime.hpp"

using namespace fbrt;

extern "C" fb_status fb_atmosphere_build_batch(fb_builder* b, fb_stream stream,
                                               const fb_parameters* params, size_t n,
                                               fb_pending_atmosphere** out) {
    if (!b || (!params && n) || (!out && n)) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    for (size_t i = 0; i < n; ++i) {
        fb_status st = fb_parameters_validate(&params[i]);
        if (st != FB_OK) return st;
    }
    for (size_t i = 0; i < n; ++i) out[i] = nullptr;
    for (size_t i = 0; i < n; ++i) {
        fb_status st = fb_atmosphere_build(b, stream, &params[i], &out[i]);
        if (st != FB_OK) {
            for (size_t k = 0; k < i; ++k) {
                fb_pending_destroy(out[k]);
                out[k] = nullptr;
            }
            return st;
        }
    }
    return FB_OK;
}
