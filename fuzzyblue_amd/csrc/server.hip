// server.hip — a request loop that serves precompute and sky-render
// requests as they arrive (no reference counterpart: the reference records
// one precompute into a caller's command buffer and never schedules).
//
// Continuous batching: one worker thread owns every HIP call.  Each turn it
//   1. retires finished work (HIP events / pending queries): a finished
//      precompute becomes a resident LUT set, a finished render unpins its
//      LUT set and releases (or keeps) its output;
//   2. admits waiting renders whose LUT set is resident (a render of a set
//      still being built waits in the queue until the set is resident);
//   3. admits waiting precomputes, in arrival order, while precompute slots
//      and the HBM budget allow — evicting least-recently-used resident LUT
//      sets that no render pins — and launches every admitted group of the
//      same precision as ONE batch (one launch per pass over all of them).
// A render pins its LUT set from submission until it finishes, so a set that
// a queued render needs is never evicted under it.  An evicted set leaves
// the resident map at once; its memory stays counted until its last use on
// the GPU has finished (polled with hipEventQuery in step 1), so an eviction
// never blocks the loop.
// Renders have a device stream of their own and precomputes rotate over the
// others.  With precompute_share_pct below 100 every precompute launch is
// persistent and sized to that share of the device's resident-block slots,
// so a render arriving mid-precompute finds free slots on every CU instead
// of waiting for the precompute's blocks to drain.  Memory admission counts
// resident LUT sets, in-flight scratch and render outputs against the budget.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>
#include <vector>

#include "runtime.hpp"

namespace fbk {
hipError_t launch_render_frame(const DevParams& P, const float4* T, const float4* S, float alt,
                               float zs, uint32_t pixels, uint64_t seed, float4* rad,
                               hipStream_t s);
}

using namespace fbrt;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    uint64_t id = 0;
    uint32_t kind = 0;
    fb_parameters params{};
    uint64_t lut_id = 0;
    float alt = 0.0f, zen = 0.0f;
    uint32_t w = 0, h = 0;
    uint64_t seed = 0;
    double t_submit = 0.0, t_start = 0.0;
    // in flight
    fb_pending_atmosphere* pending = nullptr;
    hipEvent_t ev = nullptr, ev0 = nullptr, ev_adm = nullptr;
    float4* out = nullptr;
    size_t out_bytes = 0;
    size_t reserved = 0;
};

struct Resident {
    fb_atmosphere* atm = nullptr;
    size_t bytes = 0;
    uint64_t last_use = 0;
};

struct Evicted {
    fb_atmosphere* atm = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct fb_server {
    fb_builder* b = nullptr;
    fb_server_config cfg{};
    std::vector<hipStream_t> streams;
    hipStream_t clock_stream = nullptr;  // idle stream: admission timestamps
    size_t next_stream = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_idle;
    bool stop = false;
    std::deque<Req*> waiting;
    std::vector<Req*> inflight;
    std::vector<fb_request_info> finished;
    std::unordered_map<uint64_t, Resident> resident;
    std::unordered_map<uint64_t, Req*> building;  // lut_id -> in-flight precompute
    std::unordered_map<uint64_t, int> pins;       // lut_id -> renders queued or in flight
    std::vector<Evicted> evicted;                 // destroyed once idle on the GPU
    size_t evicted_bytes = 0;
    std::unordered_map<uint64_t, std::pair<float4*, size_t>> outputs;
    size_t in_use = 0, peak = 0;
    uint64_t next_id = 1, next_lut = 1, clock = 0;
    fb_server_stats stats{};
    std::thread worker;

    // Stream 0 carries renders only, so a render is never queued behind a
    // long precompute; precomputes rotate over the other streams.
    hipStream_t render_stream() { return streams[0]; }
    hipStream_t precompute_stream() {
        if (streams.size() == 1) return streams[0];
        return streams[1 + (next_stream++ % (streams.size() - 1))];
    }

    void account(size_t add, size_t sub) {
        in_use = in_use + add - sub;
        peak = std::max(peak, in_use);
    }

    void finish(Req* r, fb_status st, float gpu_ms = 0.0f, float gpu_wait_ms = 0.0f) {
        fb_request_info info{};
        info.gpu_ms = gpu_ms;
        info.gpu_wait_ms = gpu_wait_ms;
        info.id = r->id;
        info.kind = r->kind;
        info.status = int32_t(st);
        info.lut_id = r->lut_id;
        info.t_submit = r->t_submit;
        info.t_start = r->t_start;
        info.t_done = now_s();
        finished.push_back(info);
        if (r->kind == FB_REQUEST_RENDER) unpin(r->lut_id);
        if (st == FB_OK)
            stats.completed++;
        else
            stats.rejected++;
        delete r;
    }

    void unpin(uint64_t lut) {
        auto it = pins.find(lut);
        if (it != pins.end() && --it->second <= 0) pins.erase(it);
    }
    bool pinned(uint64_t lut) const {
        auto it = pins.find(lut);
        return it != pins.end() && it->second > 0;
    }

    // Destroy evicted sets whose GPU work has finished (never waits).
    void reap_evicted() {
        std::vector<Evicted> still;
        for (const Evicted& e : evicted) {
            if (!atmosphere_idle(e.atm)) {
                still.push_back(e);
                continue;
            }
            fb_atmosphere_destroy(e.atm);  // idle: returns without waiting
            evicted_bytes -= e.bytes;
            account(0, e.bytes);
        }
        evicted.swap(still);
    }

    // Evict LRU resident sets (not pinned, not `keep`) until `need` will fit
    // once the evicted sets are idle; true when it fits now.
    bool make_room(size_t need, uint64_t keep) {
        reap_evicted();
        while (in_use - evicted_bytes + need > cfg.hbm_budget_bytes) {
            uint64_t victim = 0, oldest = UINT64_MAX;
            for (auto& kv : resident) {
                if (kv.first == keep || pinned(kv.first)) continue;
                if (kv.second.last_use < oldest) {
                    oldest = kv.second.last_use;
                    victim = kv.first;
                }
            }
            if (!victim) break;
            Resident res = resident[victim];
            resident.erase(victim);
            evicted.push_back(Evicted{res.atm, res.bytes});
            evicted_bytes += res.bytes;
            stats.evictions++;
        }
        reap_evicted();
        return in_use + need <= cfg.hbm_budget_bytes;
    }

    void retire() {
        reap_evicted();
        std::vector<Req*> still;
        for (Req* r : inflight) {
            if (r->kind == FB_REQUEST_PRECOMPUTE) {
                fb_status st = fb_pending_query(r->pending);
                if (st == FB_ENOTREADY) {
                    still.push_back(r);
                    continue;
                }
                building.erase(r->lut_id);
                if (st == FB_OK) {
                    fb_atmosphere* a = nullptr;
                    st = fb_pending_assert_ready(r->pending, &a);
                    if (st == FB_OK) {
                        size_t res = 0;
                        fb_parameters_footprint(&r->params, &res, nullptr);
                        resident[r->lut_id] = Resident{a, res, ++clock};
                        account(res, r->reserved);
                    }
                } else {
                    fb_pending_destroy(r->pending);
                    account(0, r->reserved);
                }
                finish(r, st);
            } else {
                hipError_t e = hipEventQuery(r->ev);
                if (e == hipErrorNotReady) {
                    (void)hipGetLastError();
                    still.push_back(r);
                    continue;
                }
                float gpu_ms = 0.0f, gpu_wait_ms = 0.0f;
                if (e == hipSuccess && hipEventElapsedTime(&gpu_ms, r->ev0, r->ev) != hipSuccess) {
                    (void)hipGetLastError();
                    gpu_ms = 0.0f;
                }
                if (e == hipSuccess &&
                    hipEventElapsedTime(&gpu_wait_ms, r->ev_adm, r->ev0) != hipSuccess) {
                    (void)hipGetLastError();
                    gpu_wait_ms = 0.0f;
                }
                (void)hipEventDestroy(r->ev);
                (void)hipEventDestroy(r->ev0);
                (void)hipEventDestroy(r->ev_adm);
                if (cfg.keep_outputs && e == hipSuccess) {
                    outputs[r->id] = {r->out, r->out_bytes};
                } else {
                    b->pool->release(r->out, r->out_bytes);
                    account(0, r->out_bytes);
                }
                finish(r, e == hipSuccess ? FB_OK : FB_EHIP, gpu_ms, gpu_wait_ms);
            }
        }
        inflight.swap(still);
    }

    void admit_renders() {
        std::deque<Req*> keep;
        for (Req* r : waiting) {
            if (r->kind != FB_REQUEST_RENDER) {
                keep.push_back(r);
                continue;
            }
            fb_atmosphere* atm = nullptr;
            auto it = resident.find(r->lut_id);
            if (it != resident.end()) {
                atm = it->second.atm;
                it->second.last_use = ++clock;
            } else {
                if (building.count(r->lut_id)) {
                    // Deferred until the set is resident: a stream wait on the
                    // precompute would stall every later render on the render
                    // stream behind it.
                    keep.push_back(r);
                    continue;
                } else {
                    bool queued = false;  // its precompute has not been admitted yet
                    for (Req* q : waiting)
                        if (q->kind == FB_REQUEST_PRECOMPUTE && q->lut_id == r->lut_id) queued = true;
                    if (queued) {
                        keep.push_back(r);
                        continue;
                    }
                    finish(r, FB_ENOENT);
                    continue;
                }
            }
            size_t bytes = size_t(r->w) * r->h * sizeof(float4);
            if (in_use + bytes > cfg.hbm_budget_bytes && !make_room(bytes, r->lut_id)) {
                keep.push_back(r);
                continue;
            }
            void* out = nullptr;
            if (b->pool->alloc(bytes, &out) != FB_OK) {
                keep.push_back(r);
                continue;
            }
            hipStream_t s = render_stream();
            hipError_t e = hipEventCreate(&r->ev);
            if (e == hipSuccess) e = hipEventCreate(&r->ev0);
            if (e == hipSuccess) e = hipEventCreate(&r->ev_adm);
            if (e == hipSuccess) e = hipEventRecord(r->ev_adm, clock_stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, atm->ready, 0);
            if (e == hipSuccess) e = hipEventRecord(r->ev0, s);
            if (e == hipSuccess)
                e = fbk::launch_render_frame(atm->dev, atm->T.p, atm->S.p, r->alt, r->zen,
                                             r->w * r->h, r->seed, static_cast<float4*>(out), s);
            if (e == hipSuccess) e = hipEventRecord(r->ev, s);
            fb_status mst = e == hipSuccess ? atmosphere_mark_use(atm, s) : FB_OK;
            if (e == hipSuccess && mst != FB_OK) e = hipErrorUnknown;
            if (e != hipSuccess) {
                (void)hipStreamSynchronize(s);
                if (r->ev) (void)hipEventDestroy(r->ev);
                if (r->ev0) (void)hipEventDestroy(r->ev0);
                if (r->ev_adm) (void)hipEventDestroy(r->ev_adm);
                b->pool->release(out, bytes);
                finish(r, hip_fail(e, "render launch"));
                continue;
            }
            r->out = static_cast<float4*>(out);
            r->out_bytes = bytes;
            r->t_start = now_s();
            account(bytes, 0);
            inflight.push_back(r);
        }
        waiting.swap(keep);
    }

    void admit_precomputes() {
        size_t slots = cfg.max_inflight;
        for (Req* r : inflight)
            if (r->kind == FB_REQUEST_PRECOMPUTE && slots) --slots;
        // Hysteresis (admit_min_free): with precomputes in flight, wait for
        // enough free slots to make the next group a batch worth launching.
        if (slots < cfg.max_inflight && slots < std::min(cfg.admit_min_free, cfg.max_inflight))
            return;
        std::vector<Req*> group[2];
        std::deque<Req*> keep;
        bool blocked = false;
        for (Req* r : waiting) {
            if (r->kind != FB_REQUEST_PRECOMPUTE || blocked || slots == 0) {
                keep.push_back(r);
                continue;
            }
            size_t need = fb_parameters_footprint(&r->params, nullptr, nullptr);
            if (need > cfg.hbm_budget_bytes) {
                finish(r, FB_EBUDGET);
                continue;
            }
            if (in_use + need > cfg.hbm_budget_bytes && !make_room(need, 0)) {
                blocked = true;  // keep arrival order: later requests wait too
                keep.push_back(r);
                continue;
            }
            r->reserved = need;
            account(need, 0);
            group[(r->params.flags & FB_FLAG_PRECISE) ? 1 : 0].push_back(r);
            --slots;
        }
        waiting.swap(keep);
        for (auto& g : group) {
            if (g.empty()) continue;
            std::vector<fb_parameters> ps;
            for (Req* r : g) ps.push_back(r->params);
            std::vector<fb_pending_atmosphere*> outs(g.size(), nullptr);
            hipStream_t s = precompute_stream();
            const unsigned share = cfg.precompute_share_pct;
            fb_status st = (g.size() == 1 && (share == 0 || share >= 100))
                               ? fb_atmosphere_build(b, s, &ps[0], &outs[0])
                               : build_batch(b, s, ps.data(), ps.size(), outs.data(),
                                             share == 0 || share >= 100 ? 100u : share, true);
            double t = now_s();
            for (size_t i = 0; i < g.size(); ++i) {
                Req* r = g[i];
                if (st != FB_OK) {
                    account(0, r->reserved);
                    finish(r, st);
                    continue;
                }
                r->pending = outs[i];
                r->t_start = t;
                building[r->lut_id] = r;
                inflight.push_back(r);
            }
        }
    }

    void run() {
        std::unique_lock<std::mutex> lk(mu);
        (void)hipSetDevice(b->device);
        while (true) {
            if (waiting.empty() && inflight.empty()) {
                cv_idle.notify_all();
                if (stop) break;
                cv_work.wait(lk, [&] { return stop || !waiting.empty(); });
                continue;
            }
            double t0 = now_s();
            retire();
            double t1 = now_s();
            admit_renders();
            double t2 = now_s();
            admit_precomputes();
            double t3 = now_s();
            const double steps[3] = {t1 - t0, t2 - t1, t3 - t2};
            for (int k = 0; k < 3; ++k) {
                stats.step_max_ms[k] = std::max(stats.step_max_ms[k], float(1e3 * steps[k]));
                if (steps[k] > 2e-3) stats.slow_steps++;
            }
            if (inflight.empty() && evicted.empty() && !waiting.empty()) {
                // Nothing runs, so nothing will free memory: what is still
                // waiting cannot fit the budget.
                for (Req* r : waiting) {
                    if (r->reserved) account(0, r->reserved);
                    finish(r, FB_EBUDGET);
                }
                waiting.clear();
            }
            if (!inflight.empty() || !evicted.empty()) {
                // Poll completions; new submissions wake the loop early.
                cv_work.wait_for(lk, std::chrono::microseconds(50));
            }
        }
    }
};

extern "C" {

fb_status fb_server_create(fb_builder* b, const fb_server_config* cfg, fb_server** out) {
    if (!b || !cfg || !out) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    *out = nullptr;
    if (cfg->hbm_budget_bytes == 0 || cfg->max_inflight == 0) {
        set_error("budget and max_inflight must be positive");
        return FB_EINVAL;
    }
    auto* s = new fb_server();
    s->b = b;
    builder_retain(b);
    s->cfg = *cfg;
    uint32_t ns = cfg->n_streams ? std::min(cfg->n_streams, 16u) : 4u;
    hipError_t e = hipSetDevice(b->device);
    for (uint32_t i = 0; i < ns && e == hipSuccess; ++i) {
        hipStream_t st = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) s->streams.push_back(st);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->clock_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        fb_status st = hip_fail(e, "server streams");
        for (auto st2 : s->streams) (void)hipStreamDestroy(st2);
        if (s->clock_stream) (void)hipStreamDestroy(s->clock_stream);
        builder_release(b);
        delete s;
        return st;
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return FB_OK;
}

void fb_server_destroy(fb_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->worker.join();  // the loop drains every submitted request first
    (void)hipSetDevice(s->b->device);
    for (auto& kv : s->resident) fb_atmosphere_destroy(kv.second.atm);
    for (auto& e : s->evicted) fb_atmosphere_destroy(e.atm);
    for (auto& kv : s->outputs) s->b->pool->release(kv.second.first, kv.second.second);
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    (void)hipStreamDestroy(s->clock_stream);
    builder_release(s->b);
    delete s;
}

fb_status fb_server_submit_precompute(fb_server* s, const fb_parameters* p, uint64_t* request_id,
                                      uint64_t* lut_id) {
    if (!s || !p) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    fb_status st = fb_parameters_validate(p);
    if (st != FB_OK) return st;
    auto* r = new Req();
    r->kind = FB_REQUEST_PRECOMPUTE;
    r->params = *p;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        r->id = s->next_id++;
        r->lut_id = s->next_lut++;
        s->waiting.push_back(r);
        s->stats.submitted++;
        if (request_id) *request_id = r->id;
        if (lut_id) *lut_id = r->lut_id;
    }
    s->cv_work.notify_all();
    return FB_OK;
}

fb_status fb_server_submit_precomputes(fb_server* s, const fb_parameters* ps, size_t n,
                                       uint64_t* request_ids, uint64_t* lut_ids) {
    if (!s || (!ps && n)) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    for (size_t i = 0; i < n; ++i) {
        fb_status st = fb_parameters_validate(&ps[i]);
        if (st != FB_OK) return st;  // nothing is enqueued
    }
    std::vector<Req*> rs(n);
    const double t = now_s();
    for (size_t i = 0; i < n; ++i) {
        rs[i] = new Req();
        rs[i]->kind = FB_REQUEST_PRECOMPUTE;
        rs[i]->params = ps[i];
        rs[i]->t_submit = t;
    }
    {
        // One critical section: the worker sees all n together and can admit
        // them as one batch.
        std::lock_guard<std::mutex> g(s->mu);
        for (size_t i = 0; i < n; ++i) {
            Req* r = rs[i];
            r->id = s->next_id++;
            r->lut_id = s->next_lut++;
            s->waiting.push_back(r);
            s->stats.submitted++;
            if (request_ids) request_ids[i] = r->id;
            if (lut_ids) lut_ids[i] = r->lut_id;
        }
    }
    s->cv_work.notify_all();
    return FB_OK;
}

fb_status fb_server_submit_render(fb_server* s, uint64_t lut_id, float altitude_km,
                                  float sun_zenith_rad, uint32_t width, uint32_t height,
                                  uint64_t seed, uint64_t* request_id) {
    if (!s || width == 0 || height == 0) {
        set_error("invalid render request");
        return FB_EINVAL;
    }
    auto* r = new Req();
    r->kind = FB_REQUEST_RENDER;
    r->lut_id = lut_id;
    r->alt = altitude_km;
    r->zen = sun_zenith_rad;
    r->w = width;
    r->h = height;
    r->seed = seed;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        r->id = s->next_id++;
        s->pins[lut_id]++;  // until the render finishes, whatever its outcome
        s->waiting.push_back(r);
        s->stats.submitted++;
        if (request_id) *request_id = r->id;
    }
    s->cv_work.notify_all();
    return FB_OK;
}

fb_status fb_server_poll(fb_server* s, fb_request_info* out, size_t cap, size_t* n) {
    if (!s || !n || (!out && cap)) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    std::lock_guard<std::mutex> g(s->mu);
    size_t k = std::min(cap, s->finished.size());
    for (size_t i = 0; i < k; ++i) out[i] = s->finished[i];
    s->finished.erase(s->finished.begin(), s->finished.begin() + k);
    *n = k;
    return FB_OK;
}

fb_status fb_server_drain(fb_server* s) {
    if (!s) return FB_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_work.notify_all();
    s->cv_idle.wait(lk, [&] { return s->waiting.empty() && s->inflight.empty(); });
    return FB_OK;
}

fb_status fb_server_download(fb_server* s, uint64_t lut_id, fb_lut which, void* host,
                             size_t bytes) {
    if (!s) return FB_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->resident.find(lut_id);
    if (it == s->resident.end()) {
        set_error("LUT set %llu is not resident", (unsigned long long)lut_id);
        return FB_ENOENT;
    }
    it->second.last_use = ++s->clock;
    return fb_atmosphere_download(it->second.atm, which, FB_FORMAT_F32, host, bytes);
}

fb_status fb_server_render_result(fb_server* s, uint64_t request_id, void* host, size_t bytes) {
    if (!s || !host) return FB_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->outputs.find(request_id);
    if (it == s->outputs.end()) {
        set_error("no kept output for request %llu", (unsigned long long)request_id);
        return FB_ENOENT;
    }
    if (bytes < it->second.second) {
        set_error("host buffer too small");
        return FB_EINVAL;
    }
    FB_HIP_TRY(hipMemcpy(host, it->second.first, it->second.second, hipMemcpyDeviceToHost));
    s->b->pool->release(it->second.first, it->second.second);
    s->account(0, it->second.second);
    s->outputs.erase(it);
    return FB_OK;
}

fb_status fb_server_get_stats(fb_server* s, fb_server_stats* out) {
    if (!s || !out) return FB_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->stats;
    out->hbm_in_use = s->in_use;
    out->hbm_peak = s->peak;
    out->resident_sets = uint32_t(s->resident.size());
    return FB_OK;
}

fb_status fb_server_resident(fb_server* s, uint64_t* ids, size_t cap, size_t* n) {
    if (!s || !n || (!ids && cap)) return FB_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    size_t k = 0;
    for (auto& kv : s->resident) {
        if (k < cap) ids[k] = kv.first;
        ++k;
    }
    *n = std::min(k, cap);
    return FB_OK;
}

}  // extern "C"
