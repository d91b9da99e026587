// batch.hip — many atmospheres of uneven LUT resolution and scattering
// order in one launch per pass.
//
// The reference builds one atmosphere per recorded command buffer
// (src/precompute.rs:1077-2073); here a batch plan enqueues, for every pass
// of the schedule (src/precompute.rs:1726-1999), ONE launch whose blocks
// cover every participating atmosphere: block b reads tile word b
// (atmosphere slot << 24 | block index inside that atmosphere's grid) and
// finds the atmosphere's parameters and LUT pointers in HBM arrays.  The
// order loop runs to the largest order in the batch; an atmosphere of order
// k contributes no tiles to the launches of orders above k.  The whole
// schedule is captured once as a hipGraph and replayed.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <utility>
#include <vector>

#include "runtime.hpp"

using namespace fbrt;

namespace {

struct BatchLaunch {
    int pass;
    int order;
    uint32_t* d_tiles = nullptr;
    unsigned ntiles = 0;
    unsigned threads = 0;
    size_t lds = 0;
    std::vector<uint32_t> host_tiles;  // until uploaded
    size_t offset = 0;                 // of d_tiles in the plan's device block
    unsigned grid = 0;                 // persistent grid size (0: one block per tile)
    unsigned tpt = 1;                  // kernel variant: multiple scattering texels per
                                       // thread (or chunked), single scattering kernel,
                                       // density texels per thread
};

// Shared state of one batch: device arrays, the launch list and the graph.
// Members (pending atmospheres) hold it through shared_ptr; it is destroyed
// with the last member, after that member waited for the batch's work.
struct BatchPlan {
    fb_builder* builder = nullptr;
    bool precise = false;
    bool f16 = false;  // FB_FLAG_F16_STORAGE
    std::mutex mu;
    size_t n_total = 0;
    std::vector<fb_pending_atmosphere*> members;
    fbk::DevParams* d_params = nullptr;
    fbk::LutPtrs* d_luts = nullptr;
    std::vector<BatchLaunch> launches;
    // Per member: its multiple-scattering rows in cache-friendly order
    // (ms_row_order), computed for the first launch of that pass.
    std::vector<std::vector<uint32_t>> ms_rows;
    // Percentage of the device's resident-block capacity each launch may
    // hold (100: one block per tile).  Below 100 every launch is persistent,
    // leaving the remaining slots on every CU to other streams.
    unsigned share_pct = 100;
    hipGraphExec_t graph = nullptr;
    // One block from the builder's pool holds the parameter and LUT-pointer
    // arrays and every tile list: no hipMalloc / hipFree (which synchronises
    // the whole device) while other work runs.
    void* block = nullptr;
    size_t block_bytes = 0;
    std::vector<uint8_t> staging;  // host image of `block`, alive until the plan dies

    ~BatchPlan() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (block) builder->pool->release(block, block_bytes);
    }

    fb_status enqueue(hipStream_t s) const {
        for (const auto& l : launches)
            FB_HIP_TRY(fbk::launch_batch(l.pass, l.order, precise, f16, d_params, d_luts,
                                         l.d_tiles, l.ntiles, l.threads, l.lds, l.grid, l.tpt, s));
        return FB_OK;
    }
};

// Order the batch's tiles so that each atmosphere's blocks are contiguous
// (its LUTs stay hot in L2 while its blocks run) and large atmospheres come
// first (the small ones fill the tail of the launch).
std::vector<int> tile_order(const std::vector<fb_parameters>& ps) {
    std::vector<int> idx(ps.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = int(i);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
        return sizes_of(ps[a]).ns > sizes_of(ps[b]).ns;
    });
    return idx;
}

// Texels per thread of a multiple-scattering row of W texels at `threads`
// threads per block (the kernel variants 1, 2, 4), or kMsChunkedTpt for the
// chunked kernel of narrow rows.
unsigned ms_tpt(const fbk::DevParams& P, unsigned threads) {
    if (fbk::ms_chunked(P)) return fbk::kMsChunkedTpt;
    const unsigned W = unsigned(P.s_nu * P.s_mus);
    return W <= threads ? 1u : (W <= 2 * threads ? 2u : 4u);
}

// The rows (mu index y, r index z) of one atmosphere's multiple-scattering
// pass in an order that keeps the density rows they read in L2.  The row's
// ray starts at radius r in direction mu; its 51 samples read the density
// LUT rows of the (r, mu) cells the ray crosses (multiple_scattering.comp:
// 30-54).  Rays with the same impact parameter p = r sqrt(1 - mu^2) lie on
// one straight line and cross the same cells, so rows ordered by p put rays
// that share cells next to each other: consecutive tiles run together on
// one XCD (contiguous eighths) or on neighbouring XCDs at the same time.
// The order only steers caching, never a result, so it is approximate and
// cheap (it is built on the serving loop's thread): r and mu of the texel
// centres (scattering.h:96-131) in binary32, p in 4096 buckets (a counting
// sort), rows of one bucket by increasing r.
std::vector<uint32_t> ms_row_order(const fbk::DevParams& P) {
    constexpr int kBuckets = 4096;
    const int MU = P.s_mu, R = P.s_r;
    const float top = P.top, bottom = P.bottom;
    const float H = sqrtf(top * top - bottom * bottom);
    auto unit = [](float u, int n) { return (u - 0.5f / n) / (1.0f - 1.0f / n); };
    std::vector<uint16_t> bucket(size_t(MU) * R);
    std::vector<uint32_t> start(kBuckets + 1, 0u);
    for (int z = 0; z < R; ++z) {
        const float rho = H * unit((z + 0.5f) / R, R);
        const float r = sqrtf(rho * rho + bottom * bottom);
        for (int y = 0; y < MU; ++y) {
            const float uz = (y + 0.5f) / MU;
            float mu;
            if (uz < 0.5f) {
                const float d_min = r - bottom, d_max = rho;
                const float d = d_min + (d_max - d_min) * unit(1.0f - 2.0f * uz, MU / 2);
                mu = d == 0.0f ? -1.0f : -(rho * rho + d * d) / (2.0f * r * d);
            } else {
                const float d_min = top - r, d_max = rho + H;
                const float d = d_min + (d_max - d_min) * unit(2.0f * uz - 1.0f, MU / 2);
                mu = d == 0.0f ? 1.0f : (H * H - rho * rho - d * d) / (2.0f * r * d);
            }
            const float p = r * sqrtf(std::max(0.0f, 1.0f - mu * mu));
            const int k = std::min(kBuckets - 1, std::max(0, int(p / top * kBuckets)));
            bucket[size_t(z) * MU + y] = uint16_t(k);
            ++start[k + 1];
        }
    }
    for (int k = 0; k < kBuckets; ++k) start[k + 1] += start[k];
    std::vector<uint32_t> order(bucket.size());
    for (uint32_t i = 0; i < uint32_t(bucket.size()); ++i) order[start[bucket[i]]++] = i;
    return order;
}

// XCD-contiguous order of a tile list: hardware block i runs on XCD i % 8,
// so tile i of the permuted list is logical tile xcd_remap(i), and XCD k
// covers the k-th contiguous eighth of the logical list.
std::vector<uint32_t> xcd_order(const std::vector<uint32_t>& tiles) {
    const size_t n = tiles.size(), per = n / 8, rem = n % 8;
    std::vector<uint32_t> out(n);
    for (size_t i = 0; i < n; ++i) {
        const size_t xcd = i % 8, local = i / 8;
        out[i] = tiles[xcd * per + std::min(xcd, rem) + local];
    }
    return out;
}

fb_status add_launch(BatchPlan& plan, const std::vector<fb_parameters>& ps,
                     const std::vector<fbk::DevParams>& dps, const std::vector<fbk::Tuning>& tus,
                     const std::vector<int>& order_idx, int pass, int order) {
    auto member = [&](int i) { return pass < fbk::kPassDensity || int(ps[i].order) >= order; };
    // One launch per block shape.  Multiple scattering: each atmosphere runs
    // its own tuned shape (rows of very different widths want different
    // block sizes), one launch per distinct (threads, texels per thread).
    // The other tunable passes: one launch, with the tuning of its largest
    // atmosphere when that suits every member, else a size that suits all.
    struct Group {
        unsigned threads, tpt;
        bool remap;
        std::vector<int> members;
    };
    std::vector<Group> groups;
    const bool tunable = pass == fbk::kPassSingle || pass == fbk::kPassDensity ||
                         pass == fbk::kPassIndirect || pass == fbk::kPassMultiple;
    for (int i : order_idx) {
        if (!member(i)) continue;
        if (pass == fbk::kPassSingle) {
            // The staged kernel (transmittance rows in LDS), the direct one,
            // or the chunked direct one for narrow rows: one launch per kernel.
            const unsigned c = fbk::ss_chunked(dps[i])  ? fbk::kSsChunked
                               : fbk::ss_staged(dps[i]) ? fbk::kSsStaged
                                                        : fbk::kSsDirect;
            auto g = std::find_if(groups.begin(), groups.end(),
                                  [&](const Group& x) { return x.tpt == c; });
            if (g == groups.end())
                groups.push_back(Group{c == fbk::kSsChunked
                                           ? fbk::kSsChunkThreads
                                           : fbk::tuned_threads(dps[i], pass, order, tus[i]),
                                       c, bool((tus[i].remap >> pass) & 1u), {i}});
            else
                g->members.push_back(i);
        } else if (pass == fbk::kPassMultiple) {
            const unsigned t = fbk::ms_chunked(dps[i])
                                   ? fbk::ms_chunk_threads(dps[i])
                                   : fbk::tuned_threads(dps[i], pass, order, tus[i]);
            const unsigned c = ms_tpt(dps[i], t);
            auto g = std::find_if(groups.begin(), groups.end(),
                                  [&](const Group& x) { return x.threads == t && x.tpt == c; });
            if (g == groups.end())
                groups.push_back(Group{t, c, bool((tus[i].remap >> pass) & 1u), {i}});
            else
                g->members.push_back(i);
        } else if (pass == fbk::kPassDensity) {
            // One launch per texels-per-thread variant (den_tpx).
            const unsigned c = unsigned(fbk::den_tpx(dps[i]));
            auto g = std::find_if(groups.begin(), groups.end(),
                                  [&](const Group& x) { return x.tpt == c; });
            if (g == groups.end())
                groups.push_back(Group{fbk::tuned_threads(dps[i], pass, order, tus[i]), c,
                                       bool((tus[i].remap >> pass) & 1u), {i}});
            else
                g->members.push_back(i);
        } else if (groups.empty()) {
            groups.push_back(Group{tunable ? fbk::tuned_threads(dps[i], pass, order, tus[i]) : 0u,
                                   1u, tunable && ((tus[i].remap >> pass) & 1u), {i}});
        } else {
            groups[0].members.push_back(i);
        }
    }
    if (tunable && pass != fbk::kPassMultiple)
        for (Group& g : groups)
            for (int i : g.members)
                if (!fbk::tuning_valid(dps[i], pass, order, g.threads))
                    g.threads = pass == fbk::kPassIndirect ? 64 : 256;
    for (Group& g : groups) {
        BatchLaunch l;
        l.pass = pass;
        l.order = order;
        l.tpt = g.tpt;
        std::vector<uint32_t> tiles;
        for (int i : g.members) {
            fbk::PassShape sh = fbk::batch_pass_shape(dps[i], pass, order, g.threads);
            if (sh.blocks >= (1u << 24)) {
                set_error("atmosphere %d needs %u blocks in pass %d (limit 2^24)", i, sh.blocks,
                          pass);
                return FB_EINVAL;
            }
            if (pass == fbk::kPassMultiple) {
                if (plan.ms_rows.size() < dps.size()) plan.ms_rows.resize(dps.size());
                if (plan.ms_rows[i].empty()) plan.ms_rows[i] = ms_row_order(dps[i]);
                // block index = z * MU + y
                for (uint32_t b : plan.ms_rows[i]) tiles.push_back((uint32_t(i) << 24) | b);
            } else {
                for (unsigned b = 0; b < sh.blocks; ++b) tiles.push_back((uint32_t(i) << 24) | b);
            }
            l.threads = std::max(l.threads, sh.threads);
            l.lds = std::max(l.lds, sh.lds);
        }
        if (tiles.empty()) continue;
        if (l.lds > 65536) {
            set_error("pass %d needs %zu bytes of LDS per block", pass, l.lds);
            return FB_EINVAL;
        }
        l.ntiles = unsigned(tiles.size());
        if (plan.share_pct < 100) {
            const unsigned resident =
                fbk::batch_resident_blocks(pass, order, plan.precise, plan.f16, l.threads, l.lds,
                                           l.tpt);
            if (resident) l.grid = std::max(1u, resident * plan.share_pct / 100);
        }
        l.host_tiles = g.remap ? xcd_order(tiles) : std::move(tiles);
        plan.launches.push_back(std::move(l));
    }
    return FB_OK;
}

// Place the parameter / LUT-pointer arrays and the tile lists in one pool
// block and upload them with one asynchronous copy on the launch stream,
// ahead of the batch's first launch: the host waits for nothing (a
// synchronous copy, or a wait on another stream, can sit behind unrelated
// work that shares the stream's hardware queue).
fb_status upload_plan(fb_builder* b, BatchPlan& plan, const std::vector<fbk::DevParams>& dps,
                      const std::vector<fbk::LutPtrs>& luts, hipStream_t s) {
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t n = dps.size();
    const size_t off_luts = align(n * sizeof(fbk::DevParams));
    size_t off = off_luts + align(n * sizeof(fbk::LutPtrs));
    for (auto& l : plan.launches) {
        l.offset = off;
        off += align(l.host_tiles.size() * sizeof(uint32_t));
    }
    constexpr size_t kGrain = size_t(256) << 10;  // size classes the pool can reuse
    plan.block_bytes = (off + kGrain - 1) / kGrain * kGrain;
    fb_status st = b->pool->alloc(plan.block_bytes, &plan.block);
    if (st != FB_OK) {
        plan.block = nullptr;
        return st;
    }
    std::vector<uint8_t>& staging = plan.staging;
    staging.assign(off, 0);
    memcpy(staging.data(), dps.data(), n * sizeof(fbk::DevParams));
    memcpy(staging.data() + off_luts, luts.data(), n * sizeof(fbk::LutPtrs));
    for (const auto& l : plan.launches)
        memcpy(staging.data() + l.offset, l.host_tiles.data(),
               l.host_tiles.size() * sizeof(uint32_t));
    uint8_t* base = static_cast<uint8_t*>(plan.block);
    FB_HIP_TRY(hipMemcpyAsync(base, staging.data(), off, hipMemcpyHostToDevice, s));
    plan.d_params = reinterpret_cast<fbk::DevParams*>(base);
    plan.d_luts = reinterpret_cast<fbk::LutPtrs*>(base + off_luts);
    for (auto& l : plan.launches) {
        l.d_tiles = reinterpret_cast<uint32_t*>(base + l.offset);
        l.host_tiles = std::vector<uint32_t>();
    }
    return FB_OK;
}

// Build (and launch) one group: same precision flags, at most 256 members.
fb_status build_group(fb_builder* b, hipStream_t s, const std::vector<fb_parameters>& ps,
                      fb_pending_atmosphere** out, unsigned share_pct, bool direct) {
    auto plan = std::make_shared<BatchPlan>();
    plan->builder = b;
    plan->share_pct = share_pct;
    plan->precise = (ps[0].flags & FB_FLAG_PRECISE) != 0;
    plan->f16 = (ps[0].flags & FB_FLAG_F16_STORAGE) != 0;
    const size_t n = ps.size();
    std::vector<fbk::DevParams> dps(n);
    std::vector<fbk::LutPtrs> luts(n);
    fb_status st = FB_OK;
    for (size_t i = 0; i < n && st == FB_OK; ++i) {
        fb_atmosphere* a = nullptr;
        fb_pending_atmosphere* pa = nullptr;
        st = alloc_luts(b, ps[i], &a, &pa);
        if (st != FB_OK) break;
        plan->members.push_back(pa);
        dps[i] = a->dev;
        luts[i] = fbk::LutPtrs{a->T.p,   a->E.p,   a->S.p, pa->dE.p,
                               pa->dR.p,  pa->dM.p,  pa->dMS.p,
                               pa->J.p,   pa->MR.p,
                               (ps[i].flags & FB_FLAG_PRECISE) ? pa->OD.p : nullptr,
                               pa->DG.p,  pa->SR.p};
    }
    bool uploaded = false;  // an asynchronous copy into plan->block is enqueued on s
    auto fail = [&](fb_status code) {
        if (uploaded) (void)hipStreamSynchronize(s);  // before block and staging go away
        for (auto* pa : plan->members) fb_pending_destroy(pa);
        plan->members.clear();
        return code;
    };
    if (st != FB_OK) return fail(st);

    const std::vector<int> order_idx = tile_order(ps);
    int max_order = 1;
    for (const auto& p : ps) max_order = std::max(max_order, int(p.order));
    std::vector<fbk::Tuning> tus(n);
    for (size_t i = 0; i < n; ++i) tus[i] = builder_tuning(b, ps[i]);
    for (int pass : {fbk::kPassTransmittance, fbk::kPassGeometry, fbk::kPassDirect,
                     fbk::kPassSingle, fbk::kPassClear})
        if ((st = add_launch(*plan, ps, dps, tus, order_idx, pass, 1)) != FB_OK) return fail(st);
    for (int order = 2; order <= max_order; ++order)
        for (int pass : {fbk::kPassDensity, fbk::kPassIndirect, fbk::kPassMultiple})
            if ((st = add_launch(*plan, ps, dps, tus, order_idx, pass, order)) != FB_OK)
                return fail(st);
    plan->ms_rows.clear();
    plan->ms_rows.shrink_to_fit();
    st = upload_plan(b, *plan, dps, luts, s);
    uploaded = plan->block != nullptr;
    if (st != FB_OK) return fail(st);

    hipError_t e = hipSuccess;
    if (direct) {
        // Launch pass by pass and complete each member right after the last
        // launch it takes part in: a small or low-order atmosphere is
        // resident long before the batch's largest member.  (No graph: the
        // batch runs once.)
        std::vector<size_t> last(n, 0);
        for (size_t li = 0; li < plan->launches.size(); ++li)
            for (size_t i = 0; i < n; ++i)
                if (plan->launches[li].pass < fbk::kPassDensity ||
                    int(ps[i].order) >= plan->launches[li].order)
                    last[i] = li;
        for (size_t li = 0; li < plan->launches.size(); ++li) {
            const BatchLaunch& l = plan->launches[li];
            e = fbk::launch_batch(l.pass, l.order, plan->precise, plan->f16, plan->d_params,
                                  plan->d_luts, l.d_tiles, l.ntiles, l.threads, l.lds, l.grid,
                                  l.tpt, s);
            for (size_t i = 0; i < n && e == hipSuccess; ++i) {
                if (last[i] != li) continue;
                e = hipEventRecord(plan->members[i]->done, s);
                if (e == hipSuccess) e = hipEventRecord(plan->members[i]->atm->ready, s);
            }
            if (e != hipSuccess) return fail(hip_fail(e, "batch launch"));
        }
        plan->n_total = plan->members.size();
        for (auto* pa : plan->members) pa->batch = plan;
        for (size_t i = 0; i < n; ++i) out[i] = plan->members[i];
        return FB_OK;
    }

    // Capture the schedule once; replays reuse it (fb_pending_relaunch).
    {
        std::lock_guard<std::mutex> g(b->mu);
        hipGraph_t graph = nullptr;
        e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return fail(hip_fail(e, "hipStreamBeginCapture"));
        st = plan->enqueue(b->capture_stream);
        e = hipStreamEndCapture(b->capture_stream, &graph);
        if (st != FB_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(st);
        }
        if (e != hipSuccess) return fail(hip_fail(e, "hipStreamEndCapture"));
        e = hipGraphInstantiate(&plan->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            plan->graph = nullptr;
            return fail(hip_fail(e, "hipGraphInstantiate"));
        }
    }
    e = hipGraphLaunch(plan->graph, s);
    if (e != hipSuccess) return fail(hip_fail(e, "hipGraphLaunch"));
    plan->n_total = plan->members.size();
    for (auto* pa : plan->members) {
        e = hipEventRecord(pa->done, s);
        if (e == hipSuccess) e = hipEventRecord(pa->atm->ready, s);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(s);
            return fail(hip_fail(e, "hipEventRecord"));
        }
        pa->batch = plan;
    }
    for (size_t i = 0; i < n; ++i) out[i] = plan->members[i];
    return FB_OK;
}

}  // namespace

namespace fbrt {
// Replay a batch member's captured schedule (the whole batch).  Only while
// every member is alive and unconsumed: the graph writes every member's
// scratch and output LUTs.
fb_status batch_relaunch(fb_pending_atmosphere* pa, hipStream_t s) {
    auto plan = std::static_pointer_cast<BatchPlan>(pa->batch);
    std::lock_guard<std::mutex> g(plan->mu);
    if (!plan->graph) {
        set_error("this batch was launched once without a graph and cannot be relaunched");
        return FB_EINVAL;
    }
    if (plan->members.size() != plan->n_total) {
        set_error("batch has released members; it can no longer be relaunched");
        return FB_EBUSY;
    }
    for (auto* m : plan->members)
        if (!m->atm) {
            set_error("a batch member was consumed by assert_ready");
            return FB_EBUSY;
        }
    FB_HIP_TRY(hipGraphLaunch(plan->graph, s));
    for (auto* m : plan->members) {
        FB_HIP_TRY(hipEventRecord(m->done, s));
        if (m->atm) FB_HIP_TRY(hipEventRecord(m->atm->ready, s));
    }
    return FB_OK;
}
// Called by fb_pending_destroy: the member leaves the batch.
void batch_detach(fb_pending_atmosphere* pa) {
    auto plan = std::static_pointer_cast<BatchPlan>(pa->batch);
    std::lock_guard<std::mutex> g(plan->mu);
    plan->members.erase(std::remove(plan->members.begin(), plan->members.end(), pa),
                        plan->members.end());
}
}  // namespace fbrt

namespace fbrt {
fb_status build_batch(fb_builder* b, hipStream_t s, const fb_parameters* params, size_t n,
                      fb_pending_atmosphere** out, unsigned share_pct, bool direct) {
    if (!b || (!params && n) || (!out && n)) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    for (size_t i = 0; i < n; ++i) {
        fb_status st = fb_parameters_validate(&params[i]);
        if (st != FB_OK) return st;
        out[i] = nullptr;
    }
    FB_HIP_TRY(hipSetDevice(b->device));
    // Groups: one per precision and storage mode, at most 256 atmospheres
    // each (the tile word's 8-bit atmosphere slot).
    std::vector<size_t> built;
    const uint32_t kModes = FB_FLAG_PRECISE | FB_FLAG_F16_STORAGE;
    for (uint32_t flags : {0u, FB_FLAG_PRECISE, FB_FLAG_F16_STORAGE, kModes}) {
        std::vector<size_t> idx;
        for (size_t i = 0; i < n; ++i)
            if ((params[i].flags & kModes) == flags) idx.push_back(i);
        for (size_t g0 = 0; g0 < idx.size(); g0 += 256) {
            size_t g1 = std::min(idx.size(), g0 + 256);
            std::vector<fb_parameters> ps;
            for (size_t k = g0; k < g1; ++k) ps.push_back(params[idx[k]]);
            std::vector<fb_pending_atmosphere*> outs(ps.size(), nullptr);
            fb_status st = build_group(b, s, ps, outs.data(), share_pct, direct);
            if (st != FB_OK) {
                for (size_t i : built) {
                    fb_pending_destroy(out[i]);
                    out[i] = nullptr;
                }
                return st;
            }
            for (size_t k = g0; k < g1; ++k) {
                out[idx[k]] = outs[k - g0];
                built.push_back(idx[k]);
            }
        }
    }
    return FB_OK;
}
}  // namespace fbrt

extern "C" fb_status fb_atmosphere_build_batch(fb_builder* b, fb_stream stream,
                                               const fb_parameters* params, size_t n,
                                               fb_pending_atmosphere** out) {
    return build_batch(b, static_cast<hipStream_t>(stream), params, n, out, 100, false);
}
