// WavefrontRenderer::renderScene (src/render/WavefrontRenderer.cpp:94-399) on the device: per sample
// pass, primary rays for the active pixels, then per bounce ray sort + materialize, closest hit +
// stable hit/miss split, shading-queue policy sort + materialize, shade (emission, NEE with shadow
// rays, scatter, Russian roulette, next-ray compaction) and the miss loop; adaptive sampling
// every pass and the per-pixel resolve at the end.
#include <cmath>
#include <utility>

#include "wf_device.h"
#include "wf_internal.h"

namespace wf {

wf_status carve_ray_soa(wf_context* ctx, const char* name, size_t cap, wf_ray_soa* r) {
    void* p = nullptr;
    WF_TRY(workspace(ctx, name, cap * 48, &p));
    float* f = (float*)p;
    r->ox = f; r->oy = f + cap; r->oz = f + 2 * cap;
    r->dx = f + 3 * cap; r->dy = f + 4 * cap; r->dz = f + 5 * cap;
    r->tr = f + 6 * cap; r->tg = f + 7 * cap; r->tb = f + 8 * cap;
    r->pixel = (int32_t*)(f + 9 * cap);
    r->depth = (int32_t*)(f + 10 * cap);
    r->pdf = f + 11 * cap;
    return WF_OK;
}

wf_status carve_shade_soa(wf_context* ctx, const char* name, size_t cap, wf_shade_soa* q) {
    void* p = nullptr;
    WF_TRY(workspace(ctx, name, cap * 104, &p));
    float* f = (float*)p;
    float** fl[19] = {&q->px, &q->py, &q->pz, &q->nx, &q->ny, &q->nz, &q->t, &q->u, &q->v,
                      &q->ox, &q->oy, &q->oz, &q->dx, &q->dy, &q->dz, &q->tr, &q->tg, &q->tb,
                      &q->pdf};
    for (int k = 0; k < 19; ++k) *fl[k] = f + k * cap;
    int32_t** in[6] = {&q->tri, &q->mat, &q->tex, &q->pixel, &q->depth, &q->geom};
    for (int k = 0; k < 6; ++k) *in[k] = (int32_t*)(f + (19 + k) * cap);
    q->front = (uint8_t*)(f + 25 * cap);
    return WF_OK;
}

// generatePrimaryRays (cpp:401-457) with the Camera basis (Camera.h:13-40) computed on the host.
__global__ void k_raygen(CamBasis c, int width, int height, const int32_t* __restrict__ pixels,
                         int n, uint32_t seed, int job, int sample, wf_ray_soa r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = pixels ? pixels[i] : i;
    PathRng rng{seed, (uint32_t)job, (uint32_t)p, (uint32_t)sample, kRaygenDepth};
    const V3 d = primary_dir(c.v, p % width, p / width, width, height, rng(0), rng(1));
    st3(r.ox, r.oy, r.oz, i, V3(c.v[0], c.v[1], c.v[2]));
    st3(r.dx, r.dy, r.dz, i, d);
    r.tr[i] = 1.0f;
    r.tg[i] = 1.0f;
    r.tb[i] = 1.0f;
    r.pixel[i] = p;
    r.depth[i] = 0;
    r.pdf[i] = -1.0f;
}

// AdaptiveSampler::addSample on luminance(accum - accumBefore) for active pixels.
__global__ void k_adaptive_add(const float* __restrict__ accum, const float* __restrict__ before,
                               const uint32_t* __restrict__ active, int P, float* __restrict__ mean,
                               float* __restrict__ m2, int32_t* __restrict__ count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !active[p]) return;
    const V3 c(accum[p * 3] - before[p * 3], accum[p * 3 + 1] - before[p * 3 + 1],
               accum[p * 3 + 2] - before[p * 3 + 2]);
    const float x = luminance(c);
    const int cnt = count[p] + 1;
    count[p] = cnt;
    const float mb = mean[p];
    const float dev = x - mb;
    const float mn = mb + dev / (float)cnt;
    mean[p] = mn;
    const float dev2 = x - mn;
    m2[p] = m2[p] + dev * dev2;
}

// AdaptiveSampler::updateConvergence; active_count receives the number still active.
__global__ void k_adaptive_update(const float* __restrict__ mean, const float* __restrict__ m2,
                                  const int32_t* __restrict__ count, uint32_t* __restrict__ active,
                                  int P, float rel, int min_samples, int32_t* __restrict__ active_count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool act = false;
    if (p < P && active[p]) {
        act = true;
        const int c = count[p];
        if (c >= min_samples) {
            const float var = m2[p] / (float)(c - 1);
            const float se = __builtin_sqrtf(var / (float)c);
            const float thr = rel * (mean[p] + 0.01f);
            if (se < thr) {
                active[p] = 0u;
                act = false;
            }
        }
    }
    const uint64_t b = __ballot(act);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(active_count, (int32_t)__popcll(b));
}

__global__ void k_resolve(const float* __restrict__ accum, const int32_t* __restrict__ counts,
                          int samples_run, int P, float* __restrict__ image,
                          int32_t* __restrict__ counts_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int c = counts ? counts[p] : samples_run;
    c = c > 1 ? c : 1;
    image[p * 3] = accum[p * 3] / (float)c;
    image[p * 3 + 1] = accum[p * 3 + 1] / (float)c;
    image[p * 3 + 2] = accum[p * 3 + 2] / (float)c;
    if (counts_out) counts_out[p] = c;
}


__global__ void k_copy_f32(const float* __restrict__ a, float* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_active_list(const uint32_t* __restrict__ active, const uint32_t* __restrict__ prefix,
                              int P, int32_t* __restrict__ list) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P && active[p]) list[prefix[p]] = p;
}

struct PhaseTimer {
    wf_context* ctx;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> marks[5];
    hipEvent_t open = nullptr;
    int open_phase = -1;
    void begin(int phase) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, ctx->stream);
        open = e;
        open_phase = phase;
    }
    void end() {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, ctx->stream);
        marks[open_phase].push_back({open, e});
        open_phase = -1;
    }
    double total(int phase) {
        double t = 0.0;
        for (auto& m : marks[phase]) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, m.first, m.second);
            t += ms;
        }
        return t;
    }
    ~PhaseTimer() {
        for (auto& v : marks)
            for (auto& m : v) {
                (void)hipEventDestroy(m.first);
                (void)hipEventDestroy(m.second);
            }
    }
};

}  // namespace wf

using namespace wf;

extern "C" wf_status wf_render(wf_scene* s, const wf_render_params* rp, float* image_rgb,
                               int32_t* sample_counts, wf_render_stats* stats) {
    clear_error();
    WF_REQUIRE(s && rp && image_rgb, "wf_render: NULL argument");
    WF_REQUIRE(rp->width >= 2 && rp->height >= 2, "wf_render: image must be at least 2x2 (got %dx%d)",
               rp->width, rp->height);
    WF_REQUIRE(rp->spp >= 0 && rp->max_depth >= 0, "wf_render: negative spp/depth");
    WF_REQUIRE(rp->first_sample >= 0, "wf_render: negative first_sample");
    WF_REQUIRE(rp->policy >= WF_POLICY_NONE && rp->policy <= WF_POLICY_COST_BENEFIT,
               "wf_render: unknown policy %d", rp->policy);
    WF_REQUIRE(!rp->adaptive || rp->adaptive_interval > 0, "wf_render: adaptive_interval must be > 0");
    WF_REQUIRE(rp->schedule == WF_SCHEDULE_REGENERATE || rp->schedule == WF_SCHEDULE_PASSES,
               "wf_render: unknown schedule %d", rp->schedule);
    WF_REQUIRE(rp->coherence_stats >= WF_COHERENCE_OFF &&
                   rp->coherence_stats <= WF_COHERENCE_CACHE_LINES,
               "wf_render: unknown coherence_stats %d", rp->coherence_stats);
    WF_TRY(scene_ready(s));
    WF_REQUIRE((int64_t)rp->width * rp->height < (1ll << 30), "wf_render: image too large");
    if (rp->schedule == WF_SCHEDULE_REGENERATE) return render_regenerate(s, rp, image_rgb, sample_counts, stats);
    wf_context* ctx = s->ctx;
    const int W = rp->width, H = rp->height;
    const int64_t P64 = (int64_t)W * H;
    WF_REQUIRE(P64 < (1ll << 30), "wf_render: image too large");
    const int P = (int)P64;
    const size_t cap = (size_t)P;

    wf_ray_soa rayA, rayB, miss;
    wf_shade_soa shA, shB;
    WF_TRY(carve_ray_soa(ctx, "r_rayA", cap, &rayA));
    WF_TRY(carve_ray_soa(ctx, "r_rayB", cap, &rayB));
    WF_TRY(carve_ray_soa(ctx, "r_miss", cap, &miss));
    WF_TRY(carve_shade_soa(ctx, "r_shA", cap, &shA));
    WF_TRY(carve_shade_soa(ctx, "r_shB", cap, &shB));
    void *pidx, *pacc, *pbefore, *pactive, *pmean, *pm2, *pcount, *plist, *pprefix, *pcounts;
    WF_TRY(workspace(ctx, "r_idx", cap * 4, &pidx));
    WF_TRY(workspace(ctx, "r_accum", cap * 12, &pacc));
    WF_TRY(workspace(ctx, "r_before", rp->adaptive ? cap * 12 : 4, &pbefore));
    WF_TRY(workspace(ctx, "r_active", cap * 4, &pactive));
    WF_TRY(workspace(ctx, "r_mean", cap * 4, &pmean));
    WF_TRY(workspace(ctx, "r_m2", cap * 4, &pm2));
    WF_TRY(workspace(ctx, "r_count", cap * 4, &pcount));
    WF_TRY(workspace(ctx, "r_list", cap * 4, &plist));
    WF_TRY(workspace(ctx, "r_prefix", cap * 4, &pprefix));
    WF_TRY(workspace(ctx, "r_counts", 64, &pcounts));
    void* pcoh = nullptr;
    WF_TRY(workspace(ctx, "r_coherence", sizeof(double) * 9, &pcoh));
    double* coh = (double*)pcoh;
    WF_TRY(wf_fill_u32(ctx, pcoh, 0u, 18));
    float* accum = (float*)pacc;
    int32_t* counts = (int32_t*)pcounts;
    WF_TRY(wf_fill_u32(ctx, pacc, 0u, cap * 3));
    WF_TRY(wf_fill_u32(ctx, pactive, 1u, cap));
    WF_TRY(wf_fill_u32(ctx, pmean, 0u, cap));
    WF_TRY(wf_fill_u32(ctx, pm2, 0u, cap));
    WF_TRY(wf_fill_u32(ctx, pcount, 0u, cap));

    const CamBasis cam = camera_basis(rp->camera, W, H);
    wf_render_stats st{};
    PhaseTimer pt{ctx};
    hipEvent_t f0, f1;
    WF_HIP(hipEventCreate(&f0));
    WF_HIP(hipEventCreate(&f1));
    WF_HIP(hipEventRecord(f0, ctx->stream));
    int samples_run = rp->spp;
    int active_left = P;
    for (int sample = 0; sample < rp->spp; ++sample) {
        int n = P;
        const int32_t* pixels = nullptr;
        if (rp->adaptive) {
            k_copy_f32<<<grid_for(P64 * 3, 256), 256, 0, ctx->stream>>>(accum, (float*)pbefore, cap * 3);
            WF_HIP(hipGetLastError());
            WF_TRY(exclusive_scan_u32(ctx, (uint32_t*)pactive, (uint32_t*)pprefix, host_count(P), counts + 8,
                                      "compact"));
            k_active_list<<<grid_for(P, 256), 256, 0, ctx->stream>>>(
                (uint32_t*)pactive, (uint32_t*)pprefix, P, (int32_t*)plist);
            WF_HIP(hipGetLastError());
            WF_TRY(readback(ctx, &n, counts + 8, sizeof(int32_t)));
            pixels = (int32_t*)plist;
        }
        wf_ray_soa* cur = &rayA;
        wf_ray_soa* alt = &rayB;
        if (n > 0) {
            Timed tm(ctx, "raygen", 48.0 * n);
            k_raygen<<<grid_for(n, 256), 256, 0, ctx->stream>>>(cam, W, H, pixels, n, rp->shade.seed,
                                                                rp->shade.job,
                                                                rp->first_sample + sample, *cur);
            WF_HIP(hipGetLastError());
        }
        wf_shade_params sp = rp->shade;
        sp.sample_index = rp->first_sample + sample;
        for (int depth = 0; depth < rp->max_depth && n > 0; ++depth) {
            st.rays_traced += n;
            if (rp->ray_sort) {
                pt.begin(0);
                WF_TRY(ray_schedule(ctx, cur, host_count(n), s->dev.bounds_min, s->dev.bounds_max,
                                    (int32_t*)pidx, nullptr));
                WF_TRY(gather_rays(ctx, cur, (int32_t*)pidx, host_count(n), alt));
                std::swap(cur, alt);
                pt.end();
            }
            pt.begin(1);
            WF_TRY(intersect_split(ctx, s, cur, host_count(n), &shA, &miss, counts));
            pt.end();
            int32_t nh = 0;
            WF_TRY(readback(ctx, &nh, counts, sizeof(int32_t)));
            const int nm = n - nh;
            st.shaded_hits += nh;
            st.misses += nm;
            wf_shade_soa* q = &shA;
            // live cost tracking: this bounce's schedule uses the trackers as they stand
            WF_TRY(costs_refresh(s));
            WF_TRY(costs_clear(s));
            if (rp->policy != WF_POLICY_NONE && nh > 0) {
                pt.begin(2);
                WF_TRY(shading_schedule(ctx, s, &shA, host_count(nh), rp->policy, (int32_t*)pidx));
                WF_TRY(gather_shading(ctx, &shA, (int32_t*)pidx, host_count(nh), &shB));
                q = &shB;
                pt.end();
            }
            if (rp->coherence_stats && nh > 0)
                WF_TRY(coherence_accumulate(ctx, q->mat, q->tex, host_count(nh),
                                            rp->coherence_stats >= WF_COHERENCE_CACHE_LINES,
                                            coh));
            pt.begin(3);
            WF_TRY(shade_bounce(ctx, s, q, host_count(nh), &sp, accum, alt, counts + 4, ShadeJobs{},
                                nullptr, s->track_mat ? &s->probe : nullptr));
            WF_TRY(costs_record(s));
            pt.end();
            pt.begin(4);
            WF_TRY(shade_misses(ctx, s, &miss, host_count(nm), rp->shade.physics, accum));
            pt.end();
            int32_t c[2] = {0, 0};
            WF_TRY(readback(ctx, c, counts + 4, sizeof(c)));
            st.shadow_rays += c[1];
            n = c[0];
            std::swap(cur, alt);
        }
        if (rp->adaptive) {
            k_adaptive_add<<<grid_for(P, 256), 256, 0, ctx->stream>>>(
                accum, (float*)pbefore, (uint32_t*)pactive, P, (float*)pmean, (float*)pm2,
                (int32_t*)pcount);
            WF_HIP(hipGetLastError());
            if ((sample + 1) % rp->adaptive_interval == 0) {
                WF_TRY(wf_fill_u32(ctx, counts + 12, 0u, 1));
                k_adaptive_update<<<grid_for(P, 256), 256, 0, ctx->stream>>>(
                    (float*)pmean, (float*)pm2, (int32_t*)pcount, (uint32_t*)pactive, P,
                    rp->adaptive_threshold, rp->adaptive_min_samples, counts + 12);
                WF_HIP(hipGetLastError());
                WF_TRY(readback(ctx, &active_left, counts + 12, sizeof(int32_t)));
                if (active_left == 0) {
                    samples_run = sample + 1;
                    break;
                }
            }
        }
    }
    k_resolve<<<grid_for(P, 256), 256, 0, ctx->stream>>>(accum, rp->adaptive ? (int32_t*)pcount : nullptr,
                                                         samples_run, P, image_rgb, sample_counts);
    WF_HIP(hipGetLastError());
    WF_HIP(hipEventRecord(f1, ctx->stream));
    WF_HIP(hipEventSynchronize(f1));
    float fms = 0.0f;
    (void)hipEventElapsedTime(&fms, f0, f1);
    (void)hipEventDestroy(f0);
    (void)hipEventDestroy(f1);
    st.samples_run = samples_run;
    st.active_pixels_end = rp->adaptive ? active_left : 0;
    st.frame_ms = fms;
    st.sort_ms = pt.total(0);
    st.intersect_ms = pt.total(1);
    st.schedule_ms = pt.total(2);
    st.shade_ms = pt.total(3);
    st.miss_ms = pt.total(4);
    if (rp->coherence_stats) WF_TRY(coherence_read(ctx, coh, &st));
    WF_TRY(check_fault(ctx));
    if (stats) *stats = st;
    return WF_OK;
}
