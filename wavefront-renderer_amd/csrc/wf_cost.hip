// Shading-cost tracking: CostTracker (include/render/CostTracker.h:8-63) restated on the host,
// and the live tracking of shadeAll (src/render/WavefrontRenderer.cpp:612-627, 746-758,
// 800-844): the shading kernel adds the wall-clock time of a quarter of its hits' scatter to
// per-material and per-texture device sums; after each bounce the host turns them into one mean
// per item and records it, and before each bounce the scene's relative-cost tables are refreshed
// from the trackers.
#include <vector>

#include "wf_internal.h"

namespace wf {

constexpr double kEmaWeight = 0.05;        // CostTracker.h:20
constexpr int kMinSamplesForConfidence = 50;   // CostTracker.h:44

static void tracker_record(wf_cost_tracker* t, int id, double cost_ns) {
    t->samples[id] += 1;
    if (!t->initialized[id]) {
        t->average[id] = cost_ns;
        t->initialized[id] = 1;
        t->tracked += 1;
        t->total += cost_ns;
    } else {
        const double old_cost = t->average[id];
        const double new_cost = (1.0 - kEmaWeight) * old_cost + kEmaWeight * cost_ns;
        t->average[id] = new_cost;
        t->total += (new_cost - old_cost);
    }
}

static double tracker_relative(const wf_cost_tracker* t, int id) {
    if (t->samples[id] < kMinSamplesForConfidence || t->tracked == 0) return 1.0;
    const double global_average = t->total / t->tracked;
    if (global_average <= 0.0) return 1.0;
    return t->average[id] / global_average;
}

wf_status costs_refresh(wf_scene* s) {
    if (!s->track_mat) return WF_OK;
    const int nm = s->dev.n_materials, nt = s->dev.n_textures;
    std::vector<double> mc(nm), tc(nt);
    for (int i = 0; i < nm; ++i) mc[i] = tracker_relative(s->track_mat, i);
    for (int i = 0; i < nt; ++i) tc[i] = tracker_relative(s->track_tex, i);
    return scene_upload_costs(s, mc.data(), tc.data());
}

wf_status costs_clear(wf_scene* s) {
    if (!s->track_mat) return WF_OK;
    const size_t nm = (size_t)s->dev.n_materials, nt = (size_t)s->dev.n_textures;
    if (nm) {
        WF_TRY(wf_fill_u32(s->ctx, s->probe.mat_sum, 0u, 2 * nm));
        WF_TRY(wf_fill_u32(s->ctx, s->probe.mat_cnt, 0u, nm));
    }
    if (nt) {
        WF_TRY(wf_fill_u32(s->ctx, s->probe.tex_sum, 0u, 2 * nt));
        WF_TRY(wf_fill_u32(s->ctx, s->probe.tex_cnt, 0u, nt));
    }
    return WF_OK;
}

// One mean per item timed this bounce (the combine + record loops, cpp:800-844).
wf_status costs_record(wf_scene* s) {
    if (!s->track_mat) return WF_OK;
    const int nm = s->dev.n_materials, nt = s->dev.n_textures;
    std::vector<unsigned long long> sum((size_t)std::max(nm, nt));
    std::vector<uint32_t> cnt((size_t)std::max(nm, nt));
    auto record = [&](wf_cost_tracker* t, int n, const unsigned long long* dsum,
                      const uint32_t* dcnt) -> wf_status {
        if (n == 0) return WF_OK;
        WF_TRY(readback(s->ctx, sum.data(), dsum, sizeof(unsigned long long) * n));
        WF_TRY(readback(s->ctx, cnt.data(), dcnt, sizeof(uint32_t) * n));
        for (int i = 0; i < n; ++i)
            if (cnt[i] > 0)
                tracker_record(t, i, (double)sum[i] * s->ns_per_tick / (double)cnt[i]);
        return WF_OK;
    };
    WF_TRY(record(s->track_mat, nm, s->probe.mat_sum, s->probe.mat_cnt));
    WF_TRY(record(s->track_tex, nt, s->probe.tex_sum, s->probe.tex_cnt));
    return WF_OK;
}

}  // namespace wf

using namespace wf;

extern "C" {

wf_status wf_cost_tracker_create(int32_t items, wf_cost_tracker** out) {
    clear_error();
    WF_REQUIRE(out, "wf_cost_tracker_create: NULL out");
    WF_REQUIRE(items >= 0, "wf_cost_tracker_create: negative item count %d", items);
    auto* t = new wf_cost_tracker();
    t->average.assign((size_t)items, 0.0);
    t->initialized.assign((size_t)items, 0);
    t->samples.assign((size_t)items, 0);
    *out = t;
    return WF_OK;
}

wf_status wf_cost_tracker_destroy(wf_cost_tracker* t) {
    delete t;
    return WF_OK;
}

wf_status wf_cost_tracker_record(wf_cost_tracker* t, int32_t id, double cost_ns) {
    clear_error();
    WF_REQUIRE(t, "wf_cost_tracker_record: NULL tracker");
    WF_REQUIRE(id >= 0 && id < (int32_t)t->average.size(),
               "wf_cost_tracker_record: item %d of %zu", id, t->average.size());
    tracker_record(t, id, cost_ns);
    return WF_OK;
}

wf_status wf_cost_tracker_query(const wf_cost_tracker* t, double* average_ns, int32_t* samples,
                                double* relative) {
    clear_error();
    WF_REQUIRE(t, "wf_cost_tracker_query: NULL tracker");
    for (size_t i = 0; i < t->average.size(); ++i) {
        if (average_ns) average_ns[i] = t->average[i];
        if (samples) samples[i] = t->samples[i];
        if (relative) relative[i] = tracker_relative(t, (int)i);
    }
    return WF_OK;
}

wf_status wf_scene_track_costs(wf_scene* s, wf_cost_tracker* materials,
                               wf_cost_tracker* textures) {
    clear_error();
    WF_REQUIRE(s, "wf_scene_track_costs: NULL scene");
    WF_REQUIRE((materials == nullptr) == (textures == nullptr),
               "wf_scene_track_costs: attach both trackers or neither");
    if (!materials) {
        s->track_mat = s->track_tex = nullptr;
        return WF_OK;
    }
    const int nm = s->dev.n_materials, nt = s->dev.n_textures;
    WF_REQUIRE((int)materials->average.size() == nm,
               "wf_scene_track_costs: material tracker has %zu items, the scene %d materials",
               materials->average.size(), nm);
    WF_REQUIRE((int)textures->average.size() == nt,
               "wf_scene_track_costs: texture tracker has %zu items, the scene %d textures",
               textures->average.size(), nt);
    if (!s->probe.mat_cnt) {
        void* p = nullptr;
        WF_TRY(scene_alloc(s, sizeof(unsigned long long) * std::max(nm, 1), &p));
        s->probe.mat_sum = (unsigned long long*)p;
        WF_TRY(scene_alloc(s, sizeof(uint32_t) * std::max(nm, 1), &p));
        s->probe.mat_cnt = (uint32_t*)p;
        WF_TRY(scene_alloc(s, sizeof(unsigned long long) * std::max(nt, 1), &p));
        s->probe.tex_sum = (unsigned long long*)p;
        WF_TRY(scene_alloc(s, sizeof(uint32_t) * std::max(nt, 1), &p));
        s->probe.tex_cnt = (uint32_t*)p;
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device) ==
                hipSuccess && khz > 0)
            s->ns_per_tick = 1e6 / (double)khz;
    }
    s->track_mat = materials;
    s->track_tex = textures;
    return WF_OK;
}

}  // extern "C"
