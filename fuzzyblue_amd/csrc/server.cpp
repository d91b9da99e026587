// server.cpp — placeholder
#include "../../include/fuzzyblue_hip.h"
extern "C" {
fb_status fb_server_create(fb_builder*, const fb_server_config*, fb_server** out) { if (out) *out = nullptr; return FB_EINVAL; }
void fb_server_destroy(fb_server*) {}
fb_status fb_server_submit_precompute(fb_server*, const fb_parameters*, uint64_t*, uint64_t*) { return FB_EINVAL; }
fb_status fb_server_submit_render(fb_server*, uint64_t, float, float, uint32_t, uint32_t, uint64_t, uint64_t*) { return FB_EINVAL; }
fb_status fb_server_poll(fb_server*, fb_request_info*, size_t, size_t* n) { if (n) *n = 0; return FB_EINVAL; }
fb_status fb_server_drain(fb_server*) { return FB_EINVAL; }
fb_status fb_server_download(fb_server*, uint64_t, fb_lut, void*, size_t) { return FB_EINVAL; }
fb_status fb_server_render_result(fb_server*, uint64_t, void*, size_t) { return FB_EINVAL; }
fb_status fb_server_get_stats(fb_server*, fb_server_stats*) { return FB_EINVAL; }
}
