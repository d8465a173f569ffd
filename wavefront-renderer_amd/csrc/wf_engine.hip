// Regenerating, job-tagged wavefront over one device-resident scene: the batched and served
// forms, and the default schedule of wf_render.
//
// The reference renders sample pass by sample pass, and inside a pass bounce by bounce
// (WavefrontRenderer::renderScene, src/render/WavefrontRenderer.cpp:141-312), so every bounce
// after the first runs on a shrinking queue. Here a fixed-capacity ray queue is refilled at every
// iteration: when a path ends, its pixel goes back to a ready queue and starts its next sample at
// the next iteration. A pixel has L sample lanes (L = 1 for adaptive jobs): lane l runs samples
// l, l+L, l+2L, ... one after another into an accumulator of its own, and the lanes are summed
// in lane order when the job resolves. With L = 1 the accumulator receives exactly the additions
// of the pass-by-pass loop, in the same order, and with the counter-based RNG the image is
// identical; with L > 1 only the order of the float additions differs. More lanes keep the
// wavefront full while pixels with long paths finish their samples. Rays carry the accumulator
// index in their pixel field and a job tag (job slot, sample index), so paths of many jobs of
// different sizes share one wavefront.
//
// One iteration (all counts stay on the device, so the sequence can be captured in a graph):
//   regenerate -> [radix sort of the rays' sort words (rays then visited through the sort
//   index)] -> closest hit, in the same launch as the previous bounce's shadow rays ->
//   [shading policy: hits counted per material, one range per group key, in key order] ->
//   split (hits into the shading queue, each into its key's range under a policy; misses into
//   the miss queue) -> shade (emission, NEE, scatter, RR, depth limit; next
//   rays appended to the ray queue, shadow rays left for the next closest-hit launch) -> misses
//   -> per-pixel completion (adaptive sampling, sample budget, requeue).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "wf_device.h"
#include "wf_internal.h"
#include "wf_rays.h"

namespace wf {

// Camera constructor (include/core/Camera.h:13-33) in host float arithmetic.
CamBasis camera_basis(const wf_camera& c, int width, int height) {
    struct H {
        float x, y, z;
        H operator+(const H& o) const { return {x + o.x, y + o.y, z + o.z}; }
        H operator-(const H& o) const { return {x - o.x, y - o.y, z - o.z}; }
        H operator*(float s) const { return {x * s, y * s, z * s}; }
        H operator/(float s) const { return {x / s, y / s, z / s}; }
        H cross(const H& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
        H normalized() const { return *this / std::sqrt(x * x + y * y + z * z); }
    };
    const H from{c.eye[0], c.eye[1], c.eye[2]}, at{c.target[0], c.target[1], c.target[2]};
    const H up{c.up[0], c.up[1], c.up[2]};
    const float aspect = static_cast<float>(width) / height;
    const float theta = c.vfov_deg * 3.14159265f / 180.0f;
    const float vh = 2.0f * std::tan(theta / 2.0f);
    const float vw = aspect * vh;
    const H fwd = (from - at).normalized();
    const H right = up.cross(fwd).normalized();
    const H upv = fwd.cross(right);
    const H horizontal = right * vw, vertical = upv * vh;
    const H tl = from - horizontal / 2.0f - vertical / 2.0f - fwd;
    return CamBasis{{from.x, from.y, from.z, tl.x, tl.y, tl.z, horizontal.x, horizontal.y,
                     horizontal.z, vertical.x, vertical.y, vertical.z}};
}

// Device counters of an engine (int32 slots).
enum : int {
    C_LIVE = 0,        // rays in the queue at the start of an iteration
    C_HITS = 1,        // hits of the current iteration
    C_MISSES = 2,      // misses of the current iteration
    C_NEXT = 4,        // next rays (shade_bounce counts_dev[0..2])
    C_SHADOW = 5,
    C_SHADOW_ENTRIES = 6,
    C_TAKEN = 7,       // new paths started by the current iteration's regeneration
    C_COUNT = 16,
};
// 64-bit device statistics.
enum : int { S_RAYS = 0, S_SHADOW, S_HITS, S_MISSES, S_PATHS, S_HEAD, S_TAIL, S_COUNT = 8 };

// Adaptive jobs run their samples in rounds of at most this many lanes (a lane count that divides
// the convergence-check interval), so the accumulator copies they need stay bounded.
constexpr int kMaxAdaptiveLanes = 16;

// Host-visible status block, copied back once per iteration: live rays, ready-queue cursors,
// then the work left of every job slot (pixels of an adaptive job, pixel lanes of a job with a
// fixed sample count).
constexpr int kStatusHeader = 8;   // int32: [0] live rays, [2..3] head, [4..5] tail (u64),
                                   // [7] fault word of the engine's context
constexpr int kStatusFault = 7;

struct EngineView {
    int32_t cap;               // ray slots
    int32_t px;                // framebuffer pixels (accumulator lane stride)
    unsigned long long ring;   // ready-queue entries: px * most lanes
    int32_t* counters;
    unsigned long long* stats; // S_HEAD / S_TAIL: ready-queue cursors (monotonic)
    int2* ready;               // (accumulator index, job slot << kTagJobShift | next sample)
    float* accum;              // 3 per pixel and lane (lane-major)
    int32_t* lanes_done;       // per pixel of an adaptive job: lanes that ended the round
    float* before;             // adaptive jobs: 3 per pixel and lane (lane-major, at most
                               // kMaxAdaptiveLanes lanes), the lane accumulator at the start of
                               // its current sample
    float* mean;
    float* m2;
    int32_t* pixel_count;      // samples of a retired pixel (written when it retires)
    const JobDev* jobs;
    int32_t* status;           // kStatusHeader + job slots
    int32_t* job_left;         // = status + kStatusHeader: units of each job not finished yet
                               // (adaptive: pixels; fixed sample count: pixel lanes)
    int32_t* job_max_count;    // most samples any pixel of the job ran
    int32_t* job_unconverged;  // adaptive: pixels that ran out of budget without converging
};

// Admits pixels [base, base + n) of job j: clears their state and queues their first sample of
// every lane (lane-major, so that lane 0 of every pixel starts first).
__global__ void k_admit(EngineView e, int job, int base, int n, int lanes, int queue) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int p = base + k;
    for (int l = 0; l < lanes; ++l) {
        const size_t a = (size_t)p + (size_t)l * (size_t)e.px;
        e.accum[3 * a] = 0.0f;
        e.accum[3 * a + 1] = 0.0f;
        e.accum[3 * a + 2] = 0.0f;
        if (queue)
            e.ready[(e.stats[S_TAIL] + (unsigned long long)l * n + (unsigned long long)k) % e.ring] =
                make_int2((int)a, (job << kTagJobShift) | l);
    }
    for (int l = 0; l < lanes && l < kMaxAdaptiveLanes; ++l) {
        const size_t a = (size_t)p + (size_t)l * (size_t)e.px;
        e.before[3 * a] = 0.0f;
        e.before[3 * a + 1] = 0.0f;
        e.before[3 * a + 2] = 0.0f;
    }
    e.mean[p] = 0.0f;
    e.m2[p] = 0.0f;
    e.pixel_count[p] = 0;
    e.lanes_done[p] = 0;
}

__global__ void k_admit_commit(EngineView e, int job, int n, int lanes, int queue, int left) {
    if (queue) e.stats[S_TAIL] += (unsigned long long)n * (unsigned long long)lanes;
    e.job_left[job] = queue ? left : 0;
    e.job_max_count[job] = 0;
    e.job_unconverged[job] = 0;
}

__device__ __forceinline__ int regen_take(const EngineView& e) {
    const int live = e.counters[C_LIVE];
    const unsigned long long avail = e.stats[S_TAIL] - e.stats[S_HEAD];
    const int room = e.cap - live;
    return (unsigned long long)room < avail ? room : (int)avail;
}

// Primary rays for the next samples of queued pixels, appended after the live rays
// (generatePrimaryRays, cpp:401-457, per pixel instead of per pass).
__global__ void k_regen(EngineView e, RayPacked r, int32_t* __restrict__ tags,
                        uint64_t* __restrict__ keys, RayKeyFrame f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int take = regen_take(e);
    if (i >= take) return;
    const int2 entry = e.ready[(e.stats[S_HEAD] + (unsigned long long)i) % e.ring];
    const int a = entry.x;
    const int32_t pt = entry.y;
    const int j = pt >> kTagJobShift;
    const JobDev& J = e.jobs[j];
    const int s = pt & kTagSampleMask;
    if (J.adaptive) {   // the lane's accumulator before this sample
        e.before[3 * (size_t)a] = e.accum[3 * (size_t)a];
        e.before[3 * (size_t)a + 1] = e.accum[3 * (size_t)a + 1];
        e.before[3 * (size_t)a + 2] = e.accum[3 * (size_t)a + 2];
    }
    const int local = lane_pixel(J, a, s) - J.pixel_base;
    PathRng rng{J.seed, (uint32_t)J.tag, (uint32_t)local, (uint32_t)(s + J.first_sample),
                kRaygenDepth};
    const V3 d = primary_dir(J.cam, local % J.width, local / J.width, J.width, J.height, rng(0),
                             rng(1));
    const int k = e.counters[C_LIVE] + i;
    RayVal ray;
    ray.o = V3(J.cam[0], J.cam[1], J.cam[2]);
    ray.d = d;
    ray.t = V3(1.0f, 1.0f, 1.0f);
    ray.pixel = a;
    ray.depth = 0;
    ray.pdf = -1.0f;
    st_ray(r, k, ray);
    tags[k] = (j << kTagJobShift) | s;
    if (keys) keys[k] = ray_sort_word(ray.o, ray.d, f, k);
}

__global__ void k_regen_commit(EngineView e) {
    const int take = regen_take(e);
    e.counters[C_TAKEN] = take;
    e.stats[S_HEAD] += (unsigned long long)take;
    const int live = e.counters[C_LIVE] + take;
    e.counters[C_LIVE] = live;
    e.stats[S_RAYS] += (unsigned long long)live;
    e.stats[S_PATHS] += (unsigned long long)take;
}

constexpr int kFinishBlock = 1024;
constexpr int kFinishItems = 4;

// What an ended path retired: one unit of job_left (a pixel of an adaptive job, a pixel lane of
// a job with a fixed sample count) with its pixel's sample count.
struct Retired {
    int32_t job, samples;
    bool on, unconverged;
};

// A path of accumulator a ended: count the sample, then requeue the lane's next sample or end the
// lane; the pixel retires when its last lane ends or it converged. Returns the number of ready
// entries to requeue: entry q is (first.x + q * stride, first.y + q).
//
// Non-adaptive lanes are independent: lane l runs samples l, l + L, ... to the budget.
// Adaptive jobs (AdaptiveSampler::addSample/updateConvergence,
// include/render/AdaptiveSampler.h:21-105) run their samples in rounds of L: samples rL ..
// rL + L - 1 on lanes 0 .. L - 1. The lane that ends a round last updates the pixel's running
// mean and M2 with the round's samples in sample order and tests convergence where the sample
// count is a multiple of the interval, as the pass loop does (L divides the interval, so such
// counts only fall on round ends); then the next round is queued, or the pixel retires.
// valid = this thread has an ended path; tag = its job slot and sample index. The job counters
// of what it retired are left to finish_retire.
__device__ __forceinline__ int finish_eval(const EngineView& e, int a, int32_t tag, bool valid,
                                           int2* first, int32_t* stride, Retired* out) {
    int requeue = 0;
    bool retire = false, unconverged = false;
    int32_t job = 0, samples = 0;
    if (valid) {
        const int j = tag >> kTagJobShift;
        const JobDev& J = e.jobs[j];
        const int s = tag & kTagSampleMask;
        const int L = J.lanes;
        const int p = lane_pixel(J, a, s);
        if (J.adaptive) {
            const int r0 = s - (s & (L - 1));              // first sample of the round
            const int round_n = min(L, J.spp - r0);
            const bool last = L == 1 || atomicAdd(&e.lanes_done[p], 1) == round_n - 1;
            if (last) {
                if (L > 1) e.lanes_done[p] = 0;
                float mn = e.mean[p], m2 = e.m2[p];
                bool converged = false;
                int cnt = r0;
                for (int q = 0; q < round_n; ++q) {
                    const size_t b = 3 * ((size_t)p + (size_t)q * (size_t)J.lane_stride);
                    if (converged) {   // samples past the converged count are taken back
                        e.accum[b] = e.before[b];
                        e.accum[b + 1] = e.before[b + 1];
                        e.accum[b + 2] = e.before[b + 2];
                        continue;
                    }
                    cnt = r0 + q + 1;
                    const V3 c(e.accum[b] - e.before[b], e.accum[b + 1] - e.before[b + 1],
                               e.accum[b + 2] - e.before[b + 2]);
                    const float x = luminance(c);
                    const float dev = x - mn;
                    mn = mn + dev / (float)cnt;
                    const float dev2 = x - mn;
                    m2 = m2 + dev * dev2;
                    if (cnt % J.adaptive_interval == 0 && cnt >= J.adaptive_min_samples) {
                        const float var = m2 / (float)(cnt - 1);
                        const float se = __builtin_sqrtf(var / (float)cnt);
                        converged = se < J.adaptive_threshold * (mn + 0.01f);
                    }
                }
                e.mean[p] = mn;
                e.m2[p] = m2;
                const int next = r0 + L;
                if (converged || next >= J.spp) {
                    e.pixel_count[p] = cnt;
                    retire = true;
                    unconverged = !converged;
                    job = j;
                    samples = cnt;
                } else {
                    requeue = min(L, J.spp - next);
                    *first = make_int2(p, (j << kTagJobShift) | next);
                    *stride = J.lane_stride;
                }
            }
        } else {
            const int ns = s + L;
            if (ns < J.spp) {
                requeue = 1;
                *first = make_int2(a, (j << kTagJobShift) | ns);
                *stride = 0;
            } else {
                // the lane ran its last sample; the job ends when all its pixel lanes have (every
                // lane of the pixel writes the same count, so no per-pixel counter is needed)
                e.pixel_count[p] = J.spp;
                retire = true;
                job = j;
                samples = J.spp;
            }
        }
    }
    *out = Retired{job, samples, retire, unconverged};
    return requeue;
}

// Which ended paths a completion pass handles. A lane of a job with a fixed sample count that
// goes on with its next sample reads no accumulator, and its next path adds to the accumulator
// only after the ended path's deferred shadow rays have (those run in the next closest-hit
// launch, before any shading of that iteration), so it is requeued at once (kFinishNext); ends of
// lanes and adaptive rounds read or publish the accumulators and wait for the shadow rays
// (kFinishRest). Misses have no pending shadow rays and are finished in one pass.
enum FinishMode : int { kFinishNext = 1, kFinishRest = 2 };

__device__ __forceinline__ bool finish_selected(const EngineView& e, int32_t tag, int mode) {
    const JobDev& J = e.jobs[tag >> kTagJobShift];
    const bool next = !J.adaptive && (tag & kTagSampleMask) + J.lanes < J.spp;
    return (mode == kFinishNext) == next;
}

// Job counters of the block's retirements: those of the job that one of them claims first are
// summed in LDS and applied with one set of atomics per block; those of other jobs (blocks that
// mix jobs) with one set per wave and job. Every thread of the block calls it.
__device__ __forceinline__ void finish_retire(const EngineView& e,
                                              const Retired (&r)[kFinishItems]) {
    __shared__ int32_t s_job, s_left, s_max, s_unconverged;
    if (threadIdx.x == 0) {
        s_job = -1;
        s_left = s_max = s_unconverged = 0;
    }
    __syncthreads();
    int32_t mine = -1;
#pragma unroll
    for (int j = 0; j < kFinishItems; ++j)
        if (mine < 0 && r[j].on) mine = r[j].job;
    const uint64_t claim = __ballot(mine >= 0);
    if (claim && (int)__lane_id() == __ffsll((unsigned long long)claim) - 1)
        atomicCAS(&s_job, -1, mine);
    __syncthreads();
    const int32_t jb = s_job;
    auto wave_sum = [](int& c, int& x, int& u) {
        for (int o = 32; o >= 1; o >>= 1) {
            c += __shfl_xor(c, o, 64);
            x = max(x, __shfl_xor(x, o, 64));
            u += __shfl_xor(u, o, 64);
        }
    };
    int left = 0, mx = 0, un = 0;
    bool other[kFinishItems];
#pragma unroll
    for (int j = 0; j < kFinishItems; ++j) {
        other[j] = r[j].on && r[j].job != jb;
        if (r[j].on && r[j].job == jb) {
            left += 1;
            mx = max(mx, r[j].samples);
            un += r[j].unconverged ? 1 : 0;
        }
    }
    // the other jobs, one at a time: the job of the wave's first pending retirement (wave-uniform
    // loop, one iteration per distinct job)
    for (;;) {
        int32_t mine = -1;
#pragma unroll
        for (int j = 0; j < kFinishItems; ++j)
            if (mine < 0 && other[j]) mine = r[j].job;
        const uint64_t m = __ballot(mine >= 0);
        if (m == 0) break;
        const int leader = __ffsll((unsigned long long)m) - 1;
        const int32_t jl = __shfl(mine, leader, 64);
        int c = 0, x = 0, u = 0;
#pragma unroll
        for (int j = 0; j < kFinishItems; ++j)
            if (other[j] && r[j].job == jl) {
                c += 1;
                x = max(x, r[j].samples);
                u += r[j].unconverged ? 1 : 0;
                other[j] = false;
            }
        wave_sum(c, x, u);
        if ((int)__lane_id() == leader) {
            atomicSub(&e.job_left[jl], c);
            atomicMax(&e.job_max_count[jl], x);
            if (u) atomicAdd(&e.job_unconverged[jl], u);
        }
    }
    wave_sum(left, mx, un);
    if (__lane_id() == 0 && left) {
        atomicAdd(&s_left, left);
        atomicMax(&s_max, mx);
        if (un) atomicAdd(&s_unconverged, un);
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_left) {
        atomicSub(&e.job_left[jb], s_left);
        atomicMax(&e.job_max_count[jb], s_max);
        if (s_unconverged) atomicAdd(&e.job_unconverged[jb], s_unconverged);
    }
}

// Requeues of the kFinishItems paths of every thread, appended to the ready ring with one cursor
// atomic per block.
__device__ __forceinline__ void finish_requeue(const EngineView& e, const int (&cnt)[kFinishItems],
                                               const int2 (&first)[kFinishItems],
                                               const int32_t (&stride)[kFinishItems],
                                               uint32_t* lds) {
    uint32_t total = 0;
#pragma unroll
    for (int j = 0; j < kFinishItems; ++j) total += (uint32_t)cnt[j];
    unsigned long long pos = block_append_counts<kFinishBlock>(&e.stats[S_TAIL], total, lds);
#pragma unroll
    for (int j = 0; j < kFinishItems; ++j)
        for (int q = 0; q < cnt[j]; ++q)
            e.ready[pos++ % e.ring] = make_int2(first[j].x + q * stride[j], first[j].y + q);
}

__global__ __launch_bounds__(kFinishBlock) void k_finish_hits(EngineView e,
                                                              const int32_t* __restrict__ pixel,
                                                              const int32_t* __restrict__ tags,
                                                              const uint32_t* __restrict__ alive,
                                                              int mode) {
    __shared__ uint32_t lds[kFinishBlock / 64];
    const int n = e.counters[C_HITS];
    const int base = blockIdx.x * kFinishBlock * kFinishItems;
    if (base >= n) return;
    int cnt[kFinishItems];
    int2 first[kFinishItems];
    int32_t stride[kFinishItems];
    Retired ret[kFinishItems];
#pragma unroll
    for (int j = 0; j < kFinishItems; ++j) {
        const int i = base + j * kFinishBlock + (int)threadIdx.x;
        const int32_t tag = i < n && !alive[i] ? tags[i] : -1;
        const bool valid = tag >= 0 && finish_selected(e, tag, mode);
        cnt[j] = finish_eval(e, valid ? pixel[i] : 0, valid ? tag : 0, valid, &first[j],
                             &stride[j], &ret[j]);
    }
    finish_retire(e, ret);
    finish_requeue(e, cnt, first, stride, lds);
}

__global__ __launch_bounds__(kFinishBlock) void k_finish_misses(EngineView e,
                                                                const int32_t* __restrict__ pixel,
                                                                const int32_t* __restrict__ tags) {
    __shared__ uint32_t lds[kFinishBlock / 64];
    const int n = e.counters[C_MISSES];
    const int base = blockIdx.x * kFinishBlock * kFinishItems;
    if (base >= n) return;
    int cnt[kFinishItems];
    int2 first[kFinishItems];
    int32_t stride[kFinishItems];
    Retired ret[kFinishItems];
#pragma unroll
    for (int j = 0; j < kFinishItems; ++j) {
        const int i = base + j * kFinishBlock + (int)threadIdx.x;
        const bool valid = i < n;
        cnt[j] = finish_eval(e, valid ? pixel[i] : 0, valid ? tags[i] : 0, valid, &first[j],
                             &stride[j], &ret[j]);
    }
    finish_retire(e, ret);
    finish_requeue(e, cnt, first, stride, lds);
}

__global__ void k_iteration_end(EngineView e) {
    e.stats[S_SHADOW] += (unsigned long long)e.counters[C_SHADOW];
    e.stats[S_HITS] += (unsigned long long)e.counters[C_HITS];
    e.stats[S_MISSES] += (unsigned long long)e.counters[C_MISSES];
    const int live = e.counters[C_NEXT];
    e.counters[C_LIVE] = live;
    e.status[0] = live;
    unsigned long long* cur = (unsigned long long*)(e.status + 2);
    cur[0] = e.stats[S_HEAD];
    cur[1] = e.stats[S_TAIL];
}

// Image of a job: the lane accumulators summed in lane order, divided by the sample count.
__global__ void k_resolve_job(const float* __restrict__ accum, const int32_t* __restrict__ pixel_count,
                              int base, int n, int lanes, int stride, float* __restrict__ image,
                              int32_t* __restrict__ counts_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int p = base + k;
    const int count = pixel_count[p];
    int c = count;
    c = c > 1 ? c : 1;
    float r = accum[3 * (size_t)p], g = accum[3 * (size_t)p + 1], b = accum[3 * (size_t)p + 2];
    for (int l = 1; l < lanes; ++l) {
        const size_t a = 3 * ((size_t)p + (size_t)l * (size_t)stride);
        r = r + accum[a];
        g = g + accum[a + 1];
        b = b + accum[a + 2];
    }
    image[3 * (size_t)k] = r / (float)c;
    image[3 * (size_t)k + 1] = g / (float)c;
    image[3 * (size_t)k + 2] = b / (float)c;
    if (counts_out) counts_out[k] = count;
}

}  // namespace wf

using namespace wf;

struct wf_engine {
    wf_context* parent = nullptr;
    wf_context* ctx = nullptr;          // own stream and workspaces
    wf_scene* scene = nullptr;
    wf_engine_config cfg{};
    int32_t cap = 0, px = 0, max_jobs = 0, lanes = 1;
    std::vector<void*> allocations;
    size_t bytes = 0;
    // device state
    // live rays (packed; visited in sorted order in place) and their job tags, two buffers: an
    // iteration reads buffer par and appends the next bounce's rays to the other one
    RayPacked rays[2]{};
    int32_t* tags[2] = {nullptr, nullptr};
    int par = 0;
    int32_t* hit_list = nullptr;        // index mode: traced-ray index of every hit
    HitRecords records{};               // index mode under a policy: the hits in range order (in
                                        // the memory of s1, which only the sorted fallback uses)
    uint64_t* keys = nullptr;           // sort word of each live ray (written by its producer)
    uint64_t* keys2 = nullptr;          // radix sort scratch
    // coherence metrics of every iteration's shading queue (set by wf_render; 0 = off)
    int coherence_mode = 0;
    double* coherence = nullptr;        // 9 doubles, see coherence_accumulate
    wf_ray_soa miss{};
    wf_shade_soa s0{}, s1{};
    int32_t *st0 = nullptr, *st1 = nullptr, *mt = nullptr;
    // the last shaded bounce: its shadow walks run in the next iteration's closest-hit launch,
    // after which its paths that ended at a hit are finished (pending_tags: their job tags)
    DeferredShadow pending;
    // what the last enqueued iteration shaded: its shading queue, or (index mode) its hits
    const wf_shade_soa* last_q = nullptr;
    bool last_index = false;
    HitSource last_hits{};
    int32_t* idx = nullptr;
    EngineView view{};
    JobDev* jobs_dev = nullptr;
    // host mirrors
    std::vector<JobDev> jobs;
    std::vector<int32_t> job_state;     // 0 free, 1 running, 2 finished (result not collected)
    std::vector<int64_t> job_iter_start;
    // running jobs whose lanes take more than one sample (fixed count, fewer lanes than samples):
    // only they need the end-of-iteration requeue pass; the captured iteration records whether
    // it has that pass
    std::vector<uint8_t> job_reuse;
    int reuse_jobs = 0;
    bool graph_reuse = false;
    std::vector<std::pair<int32_t, int32_t>> free_ranges;   // [begin, end) of free pixels
    int32_t* status_host = nullptr;     // pinned: counters + job_left per slot (two buffers)
    hipEvent_t status_ev[2] = {nullptr, nullptr};
    hipEvent_t phase_ev[2][8] = {};    // per status buffer: marks between the phases
    int phase_set = 0;                 // event set of the iteration being enqueued
    hipGraph_t graph[2] = {nullptr, nullptr};          // the captured iteration of each parity
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    bool warm = false;                  // a direct iteration has sized every workspace since the
                                        // last change of the iteration's shape
    bool launched = false;              // wf_engine_launch issued, wf_engine_wait pending
    bool launch_empty = false;          // ... with nothing enqueued
    int launch_buf = 0;
    int next_buf = 0;
    int64_t iterations = 0;
    int32_t running = 0;
    double phase_ms[5] = {0, 0, 0, 0, 0};
    bool phase_timing = false;
};

namespace {

wf_status engine_alloc(wf_engine* e, size_t bytes, void** out) {
    void* p = nullptr;
    hipError_t err = hipMalloc(&p, bytes < 16 ? 16 : bytes);
    if (err != hipSuccess) {
        (void)hipGetLastError();
        return set_error(WF_ERR_OUT_OF_MEMORY, "engine: hipMalloc(%zu) failed: %s", bytes,
                         hipGetErrorString(err));
    }
    e->allocations.push_back(p);
    e->bytes += bytes;
    *out = p;
    return WF_OK;
}

wf_status carve_rays(wf_engine* e, size_t cap, wf_ray_soa* r) {
    void* p = nullptr;
    WF_TRY(engine_alloc(e, cap * 48, &p));
    float* f = (float*)p;
    r->ox = f; r->oy = f + cap; r->oz = f + 2 * cap;
    r->dx = f + 3 * cap; r->dy = f + 4 * cap; r->dz = f + 5 * cap;
    r->tr = f + 6 * cap; r->tg = f + 7 * cap; r->tb = f + 8 * cap;
    r->pixel = (int32_t*)(f + 9 * cap);
    r->depth = (int32_t*)(f + 10 * cap);
    r->pdf = f + 11 * cap;
    return WF_OK;
}

wf_status carve_packed(wf_engine* e, size_t cap, RayPacked* r) {
    void* p = nullptr;
    WF_TRY(engine_alloc(e, cap * 48, &p));
    float4* f = (float4*)p;
    r->a = f;
    r->b = f + cap;
    r->c = f + 2 * cap;
    return WF_OK;
}

wf_status carve_shading(wf_engine* e, size_t cap, wf_shade_soa* q) {
    void* p = nullptr;
    WF_TRY(engine_alloc(e, cap * 104, &p));
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

template <typename T>
wf_status engine_array(wf_engine* e, size_t count, T** out) {
    void* p = nullptr;
    WF_TRY(engine_alloc(e, count * sizeof(T), &p));
    *out = (T*)p;
    return WF_OK;
}

// Bytes of device memory per ray slot, per framebuffer pixel and per further sample lane of a
// pixel (admission accounting).
// two ray buffers and the miss queue, two shading queues, tags (two ray buffers, two shading
// queues, misses), sort index, hit list, workspaces
constexpr size_t kBytesPerSlot = 48 * 3 + 104 * 2 + 4 * 5 + 4 + 4 + 400;
constexpr size_t kBytesPerPixel = 12 + 12 + 4 + 4 + 4 + 4 + 8;   // accum, start copy, mean, M2,
                                                                  // count, lanes done, queue
constexpr size_t kBytesPerLane = 12 + 8;                          // accum, queue
constexpr size_t kBytesPerAdaptiveLane = 12;                      // start copy
constexpr int kDefaultLanes = 64;
constexpr int kAutoPathsPerPixel = 8;
constexpr int64_t kAutoWavefront = 1 << 24;
constexpr int kMaxLanes = 64;

int lanes_of(int32_t sample_lanes) { return sample_lanes == 0 ? kDefaultLanes : sample_lanes; }
bool lanes_ok(int32_t sample_lanes) {
    return sample_lanes >= 0 && sample_lanes <= kMaxLanes && (sample_lanes & (sample_lanes - 1)) == 0;
}

void phase_mark(wf_engine* e, int k) {
    if (e->phase_timing) (void)hipEventRecord(e->phase_ev[e->phase_set][k], e->ctx->stream);
}

// Adds the phase times of the iteration recorded in event set b (already complete).
void accumulate_phases(wf_engine* e, int b) {
    if (!e->phase_timing) return;
    const int phase_of[7] = {-1, 0, 1, 2, 3, 4, -1};
    for (int k = 0; k < 7; ++k) {
        if (phase_of[k] < 0) continue;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e->phase_ev[b][k], e->phase_ev[b][k + 1]);
        e->phase_ms[phase_of[k]] += ms;
    }
}

// One iteration on the engine's stream (no host synchronisation), reading the rays of buffer
// e->par and appending the next bounce's rays to the other buffer.
//
// Index mode (the default): the split only lists the traced-ray index of every hit (into the
// range of its group key under a policy) and resolves every miss on the spot (sky or
// environment, then the accumulator index and job tag kept for completion); shading reads each
// listed ray and its closest hit where they are and maps the HitRecord in registers, so no
// shading-queue or miss-queue record is written. The shading-queue mode materialises the queue as the C-ABI
// does: the engine uses it for coherence statistics (they are taken over the queue) and for
// policies whose scenes have too many materials for ranges (their order comes from sorting the
// queue).
wf_status enqueue_iteration(wf_engine* e) {
    wf_context* ctx = e->ctx;
    const wf_scene* s = e->scene;
    const EngineView& v = e->view;
    const int cap = e->cap;
    int32_t* c = v.counters;
    const Count live{cap, c + C_LIVE}, hits{cap, c + C_HITS}, misses{cap, c + C_MISSES};
    const RayPacked cur = e->rays[e->par], nxt = e->rays[e->par ^ 1];
    int32_t* const ctag = e->tags[e->par];
    int32_t* const ntag = e->tags[e->par ^ 1];
    // the producers of live rays (regeneration, the next-ray append) write each ray's sort word
    const bool sort = e->cfg.ray_sort != 0;
    const RayKeyFrame kf = ray_key_frame(s->dev.bounds_min, s->dev.bounds_max, index_bits(cap));
    phase_mark(e, 0);
    {
        Timed tm(ctx, "regen", sort ? 64.0 : 56.0, Count{cap, c + C_TAKEN});   // bytes per new path
        k_regen<<<grid_for(cap, 256), 256, 0, ctx->stream>>>(v, cur, ctag,
                                                             sort ? e->keys : nullptr, kf);
        WF_HIP(hipGetLastError());
        k_regen_commit<<<1, 1, 0, ctx->stream>>>(v);
        WF_HIP(hipGetLastError());
    }
    phase_mark(e, 1);
    const RayPacked* rq = &cur;
    // sorted rays are visited in place through the sort order (traversal and split read each
    // ray once, so moving them first would only add a pass)
    const RayPackedOrdered ordered{cur.a, cur.b, cur.c, e->idx};
    if (sort)
        WF_TRY(radix_sort_packed(ctx, e->keys, e->keys2, e->idx, live, kf.idx_bits, 33,
                                 "ray_sort", 33 - ctx->tuning.ray_key_bits));
    phase_mark(e, 2);
    const unsigned finish_grid = grid_for(cap, kFinishBlock * kFinishItems);
    bool bucketed = false;
    int32_t* hl;
    float *ht, *hu, *hv;
    WF_TRY(hit_buffers(ctx, live, &hl, &ht, &hu, &hv));
    {
        const DeferredShadow& d = e->pending;
        if (d.armed) {
            if (sort)
                WF_TRY(trace_closest_deferred(ctx, s, &ordered, live, hl, ht, hu, hv, d));
            else
                WF_TRY(trace_closest_deferred(ctx, s, rq, live, hl, ht, hu, hv, d));
            // the rest of the previous bounce's paths that ended at a hit, now that their shadow
            // rays are in (its hit columns and hit count are still those of that bounce)
            Timed tm(ctx, "finish", 12.0, hits);
            k_finish_hits<<<finish_grid, kFinishBlock, 0, ctx->stream>>>(
                v, d.hc.pixel, d.hc.tag, d.o.alive, kFinishRest);
            WF_HIP(hipGetLastError());
        } else if (sort) {
            WF_TRY(trace_closest(ctx, s, &ordered, live, 0.001f, 1e9f, hl, ht, hu, hv, nullptr));
        } else {
            WF_TRY(trace_closest(ctx, s, rq, live, 0.001f, 1e9f, hl, ht, hu, hv, nullptr));
        }
    }
    // a shading policy's order: hits appended into the ranges of their group keys
    PolicyRanges ranges;
    if (e->cfg.policy != WF_POLICY_NONE) {
        if (sort)
            WF_TRY(policy_buckets(ctx, s, &ordered, live, hl, e->cfg.policy, c + C_HITS,
                                  &ranges, &bucketed));
        else
            WF_TRY(policy_buckets(ctx, s, rq, live, hl, e->cfg.policy, c + C_HITS, &ranges,
                                  &bucketed));
    }
    const bool index_mode = !e->coherence_mode && (e->cfg.policy == WF_POLICY_NONE || bucketed);
    ShadeJobs jobs;
    jobs.jobs = e->jobs_dev;
    jobs.next_tags = ntag;
    jobs.next_packed = nxt;
    jobs.next_keys = sort ? e->keys : nullptr;
    jobs.key_frame = kf;
    jobs.unordered = true;
    jobs.defer = &e->pending;
    const uint32_t* alive = nullptr;
    e->last_index = index_mode;
    if (index_mode) {
        HitSource h{cur.a, cur.b, cur.c, sort ? e->idx : nullptr, e->hit_list, hl, ht, hu, hv,
                    ctag};
        // (the misses are resolved by the split.) costBenefit ranks every material on its own,
        // so each of its ranges holds one material's hits, spread over the whole ray buffer:
        // the split copies them into their ranges of the records and shading reads them in
        // order (395 -> ~150 us of shading for +150 us of split per iteration on the
        // sorting-stress frame at 1080p). The other policies' ranges are few and dense (type
        // rank with cost or texture-size rank), and their hits are read where they are (the copy
        // costs more than it saves there).
        const bool records = bucketed && e->cfg.policy == WF_POLICY_COST_BENEFIT;
        WF_TRY(split_index(ctx, s, h, live, e->hit_list, e->cfg.shade.physics, v.accum,
                           e->miss.pixel, e->mt, c + C_HITS, bucketed ? &ranges : nullptr,
                           records ? &e->records : nullptr));
        if (records) {
            const HitRecords& r = e->records;
            h = HitSource{r.a, r.b, r.c, nullptr, nullptr, r.leaf, r.t, r.u, r.v, r.tag};
        }
        phase_mark(e, 3);
        e->last_hits = h;
        e->last_q = nullptr;
        phase_mark(e, 4);
        HitCols hc;
        WF_TRY(shade_hits(ctx, s, h, hits, &e->cfg.shade, v.accum, c + C_NEXT, jobs, &alive,
                          &hc));
        if (e->reuse_jobs > 0) {
            Timed tm(ctx, "finish", 12.0, hits);
            k_finish_hits<<<finish_grid, kFinishBlock, 0, ctx->stream>>>(v, hc.pixel, hc.tag,
                                                                         alive, kFinishNext);
            WF_HIP(hipGetLastError());
        }
        phase_mark(e, 5);
    } else {
        if (sort)
            WF_TRY(split_traced(ctx, s, &ordered, live, hl, ht, hu, hv, &e->s0, &e->miss,
                                c + C_HITS, ctag, e->st0, e->mt, bucketed ? &ranges : nullptr));
        else
            WF_TRY(split_traced(ctx, s, rq, live, hl, ht, hu, hv, &e->s0, &e->miss, c + C_HITS,
                                ctag, e->st0, e->mt, bucketed ? &ranges : nullptr));
        phase_mark(e, 3);
        const wf_shade_soa* q = &e->s0;
        const int32_t* qt = e->st0;
        // (a scene with more materials than the ranges support sorts and gathers instead)
        if (e->cfg.policy != WF_POLICY_NONE && !bucketed) {
            WF_TRY(shading_schedule(ctx, s, &e->s0, hits, e->cfg.policy, e->idx, false));
            WF_TRY(gather_shading(ctx, &e->s0, e->idx, hits, &e->s1, e->st0, e->st1));
            q = &e->s1;
            qt = e->st1;
        }
        e->last_q = q;
        if (e->coherence_mode)
            WF_TRY(coherence_accumulate(ctx, q->mat, q->tex, hits,
                                        e->coherence_mode >= WF_COHERENCE_CACHE_LINES,
                                        e->coherence));
        phase_mark(e, 4);
        jobs.tags = qt;
        WF_TRY(shade_bounce(ctx, s, q, hits, &e->cfg.shade, v.accum, nullptr, c + C_NEXT, jobs,
                            &alive));
        if (e->reuse_jobs > 0) {
            Timed tm(ctx, "finish", 12.0, hits);
            k_finish_hits<<<finish_grid, kFinishBlock, 0, ctx->stream>>>(v, q->pixel, qt, alive,
                                                                         kFinishNext);
            WF_HIP(hipGetLastError());
        }
        phase_mark(e, 5);
        WF_TRY(shade_misses(ctx, s, &e->miss, misses, e->cfg.shade.physics, v.accum));
    }
    phase_mark(e, 6);
    {
        Timed tm(ctx, "finish", 8.0, misses);
        k_finish_misses<<<finish_grid, kFinishBlock, 0, ctx->stream>>>(v, e->miss.pixel, e->mt);
        WF_HIP(hipGetLastError());
        k_iteration_end<<<1, 1, 0, ctx->stream>>>(v);
        WF_HIP(hipGetLastError());
    }
    phase_mark(e, 7);
    return WF_OK;
}

// Status of an iteration (one copy of the status block) into pinned buffer b.
wf_status enqueue_status(wf_engine* e, int b) {
    int32_t* dst = e->status_host + (size_t)b * (kStatusHeader + e->max_jobs);
    WF_HIP(hipMemcpyAsync(dst, e->view.status, sizeof(int32_t) * (kStatusHeader + e->max_jobs),
                          hipMemcpyDeviceToHost, e->ctx->stream));
    WF_HIP(hipEventRecord(e->status_ev[b], e->ctx->stream));
    return WF_OK;
}

void drop_graph(wf_engine* e) {
    for (int p = 0; p < 2; ++p) {
        if (e->graph_exec[p]) (void)hipGraphExecDestroy(e->graph_exec[p]);
        if (e->graph[p]) (void)hipGraphDestroy(e->graph[p]);
        e->graph_exec[p] = nullptr;
        e->graph[p] = nullptr;
    }
    e->warm = false;
}

void copy_profile_flags(wf_engine* e) {
    e->ctx->profile = e->parent->profile;
    e->ctx->profile_only = e->parent->profile_only;
    if (std::memcmp(&e->ctx->tuning, &e->parent->tuning, sizeof(wf_tuning)) != 0) {
        e->ctx->tuning = e->parent->tuning;
        // a captured iteration bakes the launch shapes in: capture again
        drop_graph(e);
    }
}

void merge_profile(wf_engine* e) {
    if (!e->ctx->profile) return;
    (void)harvest_timings(e->ctx);
    for (auto& kv : e->ctx->prof) {
        auto& d = e->parent->prof[kv.first];
        d.ms += kv.second.ms;
        d.launches += kv.second.launches;
        d.bytes += kv.second.bytes;
    }
    e->ctx->prof.clear();
}

wf_status launch_iteration(wf_engine* e, int b) {
    e->phase_set = b;
    const bool graph_ok = e->cfg.use_graph && !e->ctx->profile && !e->phase_timing;
    const int p = e->par;
    if ((e->graph_exec[0] || e->graph_exec[1]) && e->graph_reuse != (e->reuse_jobs > 0)) {
        // the captured iterations have (or lack) the requeue pass: capture again, once the
        // launches of the current graphs have completed
        WF_HIP(hipStreamSynchronize(e->ctx->stream));
        drop_graph(e);
    }
    if (graph_ok && !e->graph_exec[p] && e->warm) {
        if (!e->graph_exec[p ^ 1]) e->graph_reuse = e->reuse_jobs > 0;
        // a direct iteration of this shape sized every workspace: capture from here on (one
        // graph per ray-buffer parity, the buffers being baked into the launches)
        WF_HIP(hipStreamBeginCapture(e->ctx->stream, hipStreamCaptureModeThreadLocal));
        wf_status st = enqueue_iteration(e);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(e->ctx->stream, &g);
        if (st != WF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (ce != hipSuccess) return set_error(WF_ERR_HIP, "engine graph capture: %s", hipGetErrorString(ce));
        e->graph[p] = g;
        WF_HIP(hipGraphInstantiate(&e->graph_exec[p], g, nullptr, nullptr, 0));
    }
    if (graph_ok && e->graph_exec[p]) {
        WF_HIP(hipGraphLaunch(e->graph_exec[p], e->ctx->stream));
    } else {
        WF_TRY(enqueue_iteration(e));
        e->warm = true;
    }
    e->par ^= 1;
    ++e->iterations;
    return WF_OK;
}

}  // namespace

extern "C" {

wf_status wf_engine_create(wf_context* ctx, wf_scene* scene, const wf_engine_config* cfg,
                           wf_engine** out) {
    clear_error();
    WF_REQUIRE(ctx && scene && cfg && out, "wf_engine_create: NULL argument");
    WF_REQUIRE(cfg->policy >= WF_POLICY_NONE && cfg->policy <= WF_POLICY_COST_BENEFIT,
               "wf_engine_create: unknown policy %d", cfg->policy);
    WF_REQUIRE(cfg->shade.physics == WF_PHYSICS_REFERENCE || cfg->shade.physics == WF_PHYSICS_UNBIASED,
               "wf_engine_create: unknown physics %d", cfg->shade.physics);
    WF_REQUIRE(cfg->capacity > 0 && cfg->capacity < (1 << 29), "wf_engine_create: capacity %d out of range",
               cfg->capacity);
    WF_REQUIRE(cfg->pixel_budget > 0 && cfg->pixel_budget < (1ll << 30),
               "wf_engine_create: pixel budget %lld out of range", (long long)cfg->pixel_budget);
    WF_REQUIRE(cfg->max_jobs > 0 && cfg->max_jobs <= kMaxJobs, "wf_engine_create: max_jobs %d not in [1, %d]",
               cfg->max_jobs, kMaxJobs);
    WF_REQUIRE(lanes_ok(cfg->sample_lanes), "wf_engine_create: sample_lanes %d is not 0 or a power of two <= 64",
               cfg->sample_lanes);
    WF_REQUIRE(cfg->pixel_budget * lanes_of(cfg->sample_lanes) < (1ll << 31),
               "wf_engine_create: pixel budget x sample lanes must stay below 2^31");
    WF_TRY(scene_ready(scene));
    wf_engine* e = new wf_engine();
    e->parent = ctx;
    e->scene = scene;
    e->cfg = *cfg;
    e->cap = cfg->capacity;
    e->px = (int32_t)cfg->pixel_budget;
    e->max_jobs = cfg->max_jobs;
    e->lanes = lanes_of(cfg->sample_lanes);
    auto fail = [&](wf_status st) {
        wf_engine_destroy(e);
        return st;
    };
    wf_status st = wf_context_create(ctx->device, &e->ctx);
    if (st != WF_OK) {
        delete e;
        return st;
    }
    const size_t cap = (size_t)e->cap, px = (size_t)e->px;
#define ETRY(x)                         \
    do {                                \
        wf_status s_ = (x);             \
        if (s_ != WF_OK) return fail(s_); \
    } while (0)
    ETRY(carve_packed(e, cap, &e->rays[0]));
    ETRY(carve_packed(e, cap, &e->rays[1]));
    ETRY(carve_rays(e, cap, &e->miss));
    ETRY(carve_shading(e, cap, &e->s0));
    ETRY(carve_shading(e, cap, &e->s1));
    // columns no kernel of the engine reads: the split skips them (the ray origin of a hit, its
    // mesh-local triangle index and mesh, the origin of a miss)
    for (wf_shade_soa* q : {&e->s0, &e->s1}) {
        q->ox = q->oy = q->oz = nullptr;
        q->tri = nullptr;
        q->geom = nullptr;
    }
    e->miss.ox = e->miss.oy = e->miss.oz = nullptr;
    ETRY(engine_array(e, cap, &e->tags[0]));
    ETRY(engine_array(e, cap, &e->tags[1]));
    ETRY(engine_array(e, cap, &e->hit_list));
    {
        // 68 B per slot of s1's 104: three float4 columns, then leaf, t, u, v and tag
        float4* f4 = reinterpret_cast<float4*>(e->s1.px);
        e->records.a = f4;
        e->records.b = f4 + cap;
        e->records.c = f4 + 2 * cap;
        float* f = reinterpret_cast<float*>(f4 + 3 * cap);
        e->records.leaf = reinterpret_cast<int32_t*>(f);
        e->records.t = f + cap;
        e->records.u = f + 2 * cap;
        e->records.v = f + 3 * cap;
        e->records.tag = reinterpret_cast<int32_t*>(f + 4 * cap);
    }
    ETRY(engine_array(e, cap, &e->st0));
    ETRY(engine_array(e, cap, &e->st1));
    ETRY(engine_array(e, cap, &e->mt));
    ETRY(engine_array(e, cap, &e->idx));
    if (cfg->ray_sort) {
        ETRY(engine_array(e, cap, &e->keys));
        ETRY(engine_array(e, cap, &e->keys2));
    }
    EngineView& v = e->view;
    v.cap = e->cap;
    v.px = e->px;
    v.ring = (unsigned long long)px * (unsigned long long)e->lanes;
    ETRY(engine_array(e, (size_t)C_COUNT, &v.counters));
    ETRY(engine_array(e, (size_t)S_COUNT, &v.stats));
    ETRY(engine_array(e, px * (size_t)e->lanes, &v.ready));
    ETRY(engine_array(e, 3 * px * (size_t)e->lanes, &v.accum));
    ETRY(engine_array(e, px, &v.lanes_done));
    ETRY(engine_array(e, 3 * px * (size_t)std::min(e->lanes, kMaxAdaptiveLanes), &v.before));
    ETRY(engine_array(e, px, &v.mean));
    ETRY(engine_array(e, px, &v.m2));
    ETRY(engine_array(e, px, &v.pixel_count));
    ETRY(engine_array(e, (size_t)e->max_jobs, &e->jobs_dev));
    v.jobs = e->jobs_dev;
    ETRY(engine_array(e, (size_t)(kStatusHeader + e->max_jobs), &v.status));
    v.job_left = v.status + kStatusHeader;
    // the sort's fault word travels back with every iteration's status
    if (e->ctx->fault_owned) (void)hipFree(e->ctx->fault);
    e->ctx->fault = v.status + kStatusFault;
    e->ctx->fault_owned = false;
    ETRY(engine_array(e, (size_t)e->max_jobs, &v.job_max_count));
    ETRY(engine_array(e, (size_t)e->max_jobs, &v.job_unconverged));
    ETRY(wf_fill_u32(e->ctx, v.counters, 0u, C_COUNT));
    ETRY(wf_fill_u32(e->ctx, v.stats, 0u, 2 * S_COUNT));
    ETRY(wf_fill_u32(e->ctx, v.status, 0u, (size_t)(kStatusHeader + e->max_jobs)));
    ETRY(wf_fill_u32(e->ctx, v.pixel_count, 0u, px));
#undef ETRY
    e->jobs.assign((size_t)e->max_jobs, JobDev{});
    e->job_state.assign((size_t)e->max_jobs, 0);
    e->job_iter_start.assign((size_t)e->max_jobs, 0);
    e->job_reuse.assign((size_t)e->max_jobs, 0);
    e->free_ranges.push_back({0, e->px});
    hipError_t he = hipHostMalloc((void**)&e->status_host,
                                  sizeof(int32_t) * 2 * (kStatusHeader + (size_t)e->max_jobs));
    if (he != hipSuccess) return fail(set_error(WF_ERR_OUT_OF_MEMORY, "engine: pinned status buffer"));
    for (auto& ev : e->status_ev)
        if (hipEventCreate(&ev) != hipSuccess) return fail(set_error(WF_ERR_HIP, "engine: event"));
    for (auto& set : e->phase_ev)
        for (auto& ev : set)
            if (hipEventCreate(&ev) != hipSuccess) return fail(set_error(WF_ERR_HIP, "engine: event"));
    he = hipStreamSynchronize(e->ctx->stream);
    if (he != hipSuccess) return fail(set_error(WF_ERR_HIP, "engine init: %s", hipGetErrorString(he)));
    *out = e;
    return WF_OK;
}

wf_status wf_engine_destroy(wf_engine* e) {
    if (!e) return WF_OK;
    if (e->ctx) (void)hipStreamSynchronize(e->ctx->stream);
    for (int p = 0; p < 2; ++p) {
        if (e->graph_exec[p]) (void)hipGraphExecDestroy(e->graph_exec[p]);
        if (e->graph[p]) (void)hipGraphDestroy(e->graph[p]);
    }
    for (auto& ev : e->status_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& set : e->phase_ev)
        for (auto& ev : set)
            if (ev) (void)hipEventDestroy(ev);
    if (e->status_host) (void)hipHostFree(e->status_host);
    for (void* p : e->allocations) (void)hipFree(p);
    if (e->ctx) wf_context_destroy(e->ctx);
    delete e;
    return WF_OK;
}

wf_status wf_engine_bytes(int32_t capacity, int64_t pixel_budget, int32_t sample_lanes,
                          size_t* out) {
    WF_REQUIRE(out && capacity >= 0 && pixel_budget >= 0, "wf_engine_bytes: bad argument");
    WF_REQUIRE(lanes_ok(sample_lanes), "wf_engine_bytes: sample_lanes %d is not 0 or a power of two <= 64",
               sample_lanes);
    const int lanes = lanes_of(sample_lanes);
    *out = kBytesPerSlot * (size_t)capacity +
           (kBytesPerPixel + kBytesPerLane * (size_t)(lanes - 1) +
            kBytesPerAdaptiveLane * (size_t)(std::min(lanes, kMaxAdaptiveLanes) - 1)) *
               (size_t)pixel_budget;
    return WF_OK;
}

wf_status wf_engine_device_bytes(const wf_engine* e, size_t* out) {
    WF_REQUIRE(e && out, "wf_engine_device_bytes: NULL argument");
    size_t ws = 0;
    for (auto& kv : e->ctx->workspace) ws += kv.second.bytes;
    *out = e->bytes + ws;
    return WF_OK;
}

wf_status wf_engine_admit(wf_engine* e, const wf_job_desc* job, int32_t* job_id) {
    clear_error();
    WF_REQUIRE(e && job && job_id, "wf_engine_admit: NULL argument");
    WF_REQUIRE(job->width >= 2 && job->height >= 2, "wf_engine_admit: image must be at least 2x2 (got %dx%d)",
               job->width, job->height);
    WF_REQUIRE(job->spp >= 0 && job->spp <= kTagSampleMask, "wf_engine_admit: spp %d out of range", job->spp);
    WF_REQUIRE(job->max_depth >= 0, "wf_engine_admit: negative max_depth");
    WF_REQUIRE(!job->adaptive || job->adaptive_interval > 0, "wf_engine_admit: adaptive_interval must be > 0");
    const int64_t P64 = (int64_t)job->width * job->height;
    WF_REQUIRE(P64 <= e->px, "wf_engine_admit: %lld pixels exceed the engine's pixel budget %d",
               (long long)P64, e->px);
    const int P = (int)P64;
    int lanes = 1;
    const int most = std::min(e->lanes, lanes_of(e->cfg.sample_lanes));
    if (!job->adaptive) {
        while (lanes * 2 <= most && lanes * 2 <= job->spp) lanes *= 2;
    } else {
        // rounds of lanes that divide the convergence-check interval
        while (lanes * 2 <= std::min(most, kMaxAdaptiveLanes) && lanes * 2 <= job->spp &&
               job->adaptive_interval % (lanes * 2) == 0)
            lanes *= 2;
    }
    const long long left = job->adaptive ? (long long)P : (long long)P * lanes;
    WF_REQUIRE(left <= INT32_MAX, "wf_engine_admit: %lld pixel lanes exceed the job counter", left);
    int slot = -1;
    for (int j = 0; j < e->max_jobs; ++j)
        if (e->job_state[j] == 0) {
            slot = j;
            break;
        }
    if (slot < 0) return set_error(WF_ERR_ADMISSION, "wf_engine_admit: all %d job slots busy", e->max_jobs);
    int range = -1;
    for (size_t k = 0; k < e->free_ranges.size(); ++k)
        if (e->free_ranges[k].second - e->free_ranges[k].first >= P) {
            range = (int)k;
            break;
        }
    if (range < 0)
        return set_error(WF_ERR_ADMISSION, "wf_engine_admit: no free framebuffer range of %d pixels", P);
    const int base = e->free_ranges[range].first;
    e->free_ranges[range].first += P;
    if (e->free_ranges[range].first == e->free_ranges[range].second)
        e->free_ranges.erase(e->free_ranges.begin() + range);
    JobDev& J = e->jobs[slot];
    std::memset(&J, 0, sizeof(J));
    J.pixel_base = base;
    J.width = job->width;
    J.height = job->height;
    J.spp = job->spp;
    J.max_depth = job->max_depth;
    J.seed = job->seed;
    J.tag = job->job_tag;
    J.first_sample = job->first_sample;
    J.adaptive = job->adaptive ? 1 : 0;
    J.adaptive_threshold = job->adaptive_threshold;
    J.adaptive_min_samples = job->adaptive_min_samples;
    J.adaptive_interval = job->adaptive ? job->adaptive_interval : 1;
    J.lanes = lanes;
    J.lane_stride = e->px;
    const CamBasis cb = camera_basis(job->camera, job->width, job->height);
    std::memcpy(J.cam, cb.v, sizeof(J.cam));
    wf_context* ctx = e->ctx;
    WF_HIP(hipMemcpyAsync(e->jobs_dev + slot, &J, sizeof(JobDev), hipMemcpyHostToDevice, ctx->stream));
    const int queue = (job->spp > 0 && job->max_depth > 0) ? 1 : 0;
    k_admit<<<grid_for(P, 256), 256, 0, ctx->stream>>>(e->view, slot, base, P, lanes, queue);
    WF_HIP(hipGetLastError());
    k_admit_commit<<<1, 1, 0, ctx->stream>>>(e->view, slot, P, lanes, queue, (int)left);
    WF_HIP(hipGetLastError());
    // the host copy of J must outlive the asynchronous upload
    WF_HIP(hipStreamSynchronize(ctx->stream));
    e->job_state[slot] = queue ? 1 : 2;
    if (queue && !job->adaptive && lanes < job->spp) {
        e->job_reuse[slot] = 1;
        ++e->reuse_jobs;
    }
    e->job_iter_start[slot] = e->iterations;
    if (queue) ++e->running;
    *job_id = slot;
    return WF_OK;
}

namespace {

// Work the caller queued on the scene's context (uploads, cost tables) comes first.
wf_status wait_for_scene(wf_engine* e) {
    hipEvent_t dep;
    WF_HIP(hipEventCreateWithFlags(&dep, hipEventDisableTiming));
    WF_HIP(hipEventRecord(dep, e->scene->ctx->stream));
    WF_HIP(hipStreamWaitEvent(e->ctx->stream, dep, 0));
    (void)hipEventDestroy(dep);
    return WF_OK;
}

// Jobs finished without running (empty budget) or not reported yet.
void report_unreported(wf_engine* e, int32_t* finished, int32_t cap, int32_t* n) {
    for (int j = 0; j < e->max_jobs && *n < cap; ++j)
        if (e->job_state[j] == 2 && e->job_iter_start[j] >= 0) {
            finished[(*n)++] = j;
            e->job_iter_start[j] = -1;
        }
}

// Marks the jobs whose pixels are all done in status buffer b as finished.
// Reports finished jobs; false when the iteration's fault word was set (the word is cleared).
bool scan_status(wf_engine* e, int b, int32_t* finished, int32_t cap, int32_t* n) {
    const int32_t* st = e->status_host + (size_t)b * (kStatusHeader + e->max_jobs);
    if (st[kStatusFault] != 0) {
        (void)wf_fill_u32(e->ctx, e->view.status + kStatusFault, 0u, 1);
        return false;
    }
    for (int j = 0; j < e->max_jobs; ++j)
        if (e->job_state[j] == 1 && st[kStatusHeader + j] == 0) {
            e->job_state[j] = 2;
            if (e->job_reuse[j]) {
                e->job_reuse[j] = 0;
                --e->reuse_jobs;
            }
            --e->running;
            if (*n < cap) {
                finished[(*n)++] = j;
                e->job_iter_start[j] = -1;
            }
        }
    return true;
}

}  // namespace

wf_status wf_engine_launch(wf_engine* e, int32_t iterations) {
    clear_error();
    WF_REQUIRE(e, "wf_engine_launch: NULL engine");
    WF_REQUIRE(!e->launched, "wf_engine_launch: the previous launch was not waited for");
    WF_TRY(wait_for_scene(e));
    copy_profile_flags(e);
    e->launched = true;
    e->launch_empty = e->running == 0;
    if (e->launch_empty) return WF_OK;
    const int b = e->next_buf;
    for (int k = 0; k < (iterations > 1 ? iterations : 1); ++k) WF_TRY(launch_iteration(e, b));
    WF_TRY(enqueue_status(e, b));
    e->launch_buf = b;
    e->next_buf = b ^ 1;
    return WF_OK;
}

wf_status wf_engine_wait(wf_engine* e, int32_t* finished, int32_t finished_cap,
                         int32_t* n_finished) {
    clear_error();
    WF_REQUIRE(e && n_finished && (finished || finished_cap == 0), "wf_engine_wait: NULL argument");
    WF_REQUIRE(e->launched, "wf_engine_wait: nothing launched");
    *n_finished = 0;
    report_unreported(e, finished, finished_cap, n_finished);
    if (!e->launch_empty) {
        WF_HIP(hipEventSynchronize(e->status_ev[e->launch_buf]));
        if (!scan_status(e, e->launch_buf, finished, finished_cap, n_finished)) {
            e->launched = false;
            return fault_error();
        }
    }
    e->launched = false;
    merge_profile(e);
    return WF_OK;
}

wf_status wf_engine_run(wf_engine* e, int32_t max_iterations, int32_t* finished,
                        int32_t finished_cap, int32_t* n_finished) {
    clear_error();
    WF_REQUIRE(e && n_finished && (finished || finished_cap == 0), "wf_engine_run: NULL argument");
    WF_REQUIRE(!e->launched, "wf_engine_run: a launch is pending (call wf_engine_wait)");
    *n_finished = 0;
    WF_TRY(wait_for_scene(e));
    copy_profile_flags(e);
    // already finished jobs (spp 0) are reported at once
    report_unreported(e, finished, finished_cap, n_finished);
    if (*n_finished > 0 || e->running == 0) {
        merge_profile(e);
        return WF_OK;
    }
    // two iterations in flight: enqueue k+1 before reading the status of k
    int64_t done_iters = 0;
    int b = 0;
    WF_TRY(launch_iteration(e, b));
    WF_TRY(enqueue_status(e, b));
    for (;;) {
        const bool more = max_iterations <= 0 || done_iters + 1 < max_iterations;
        if (more) {
            WF_TRY(launch_iteration(e, b ^ 1));
            WF_TRY(enqueue_status(e, b ^ 1));
        }
        WF_HIP(hipEventSynchronize(e->status_ev[b]));
        accumulate_phases(e, b);
        ++done_iters;
        if (!scan_status(e, b, finished, finished_cap, n_finished)) {
            if (more) (void)hipEventSynchronize(e->status_ev[b ^ 1]);
            return fault_error();
        }
        if (!more || *n_finished > 0 || e->running == 0) {
            if (more) {
                WF_HIP(hipEventSynchronize(e->status_ev[b ^ 1]));
                accumulate_phases(e, b ^ 1);
            }
            break;
        }
        b ^= 1;
    }
    merge_profile(e);
    return WF_OK;
}

wf_status wf_engine_result(wf_engine* e, int32_t job_id, float* image_rgb, int32_t* sample_counts,
                           wf_job_result* result) {
    clear_error();
    WF_REQUIRE(e && image_rgb, "wf_engine_result: NULL argument");
    WF_REQUIRE(job_id >= 0 && job_id < e->max_jobs && e->job_state[job_id] == 2,
               "wf_engine_result: job %d is not finished", job_id);
    const JobDev& J = e->jobs[job_id];
    const int P = J.width * J.height;
    wf_context* ctx = e->ctx;
    k_resolve_job<<<grid_for(P, 256), 256, 0, ctx->stream>>>(e->view.accum, e->view.pixel_count,
                                                             J.pixel_base, P, J.lanes, e->px,
                                                             image_rgb, sample_counts);
    WF_HIP(hipGetLastError());
    int32_t vals[3] = {0, 0, 0};
    WF_HIP(hipMemcpyAsync(&vals[0], e->view.job_max_count + job_id, sizeof(int32_t),
                          hipMemcpyDeviceToHost, ctx->stream));
    WF_HIP(hipMemcpyAsync(&vals[1], e->view.job_unconverged + job_id, sizeof(int32_t),
                          hipMemcpyDeviceToHost, ctx->stream));
    WF_HIP(hipStreamSynchronize(ctx->stream));
    if (result) {
        result->width = J.width;
        result->height = J.height;
        result->samples_run = vals[0];
        result->active_pixels_end = J.adaptive ? vals[1] : 0;
    }
    // release the slot and its framebuffer range (merged with free neighbours)
    e->job_state[job_id] = 0;
    std::pair<int32_t, int32_t> r{J.pixel_base, J.pixel_base + P};
    e->free_ranges.push_back(r);
    std::sort(e->free_ranges.begin(), e->free_ranges.end());
    std::vector<std::pair<int32_t, int32_t>> merged;
    for (auto& x : e->free_ranges) {
        if (!merged.empty() && merged.back().second == x.first) merged.back().second = x.second;
        else merged.push_back(x);
    }
    e->free_ranges.swap(merged);
    return WF_OK;
}

wf_status wf_engine_query(wf_engine* e, wf_engine_stats* out) {
    clear_error();
    WF_REQUIRE(e && out, "wf_engine_query: NULL argument");
    unsigned long long s[S_COUNT];
    int32_t live = 0;
    WF_HIP(hipMemcpyAsync(s, e->view.stats, sizeof(s), hipMemcpyDeviceToHost, e->ctx->stream));
    WF_HIP(hipMemcpyAsync(&live, e->view.counters + C_LIVE, sizeof(int32_t), hipMemcpyDeviceToHost,
                          e->ctx->stream));
    WF_HIP(hipStreamSynchronize(e->ctx->stream));
    out->rays_traced = (int64_t)s[S_RAYS];
    out->shadow_rays = (int64_t)s[S_SHADOW];
    out->shaded_hits = (int64_t)s[S_HITS];
    out->misses = (int64_t)s[S_MISSES];
    out->paths = (int64_t)s[S_PATHS];
    out->iterations = e->iterations;
    out->live_rays = live;
    out->queued_pixels = (int64_t)(s[S_TAIL] - s[S_HEAD]);
    out->running_jobs = e->running;
    int free_px = 0;
    for (auto& r : e->free_ranges) free_px += r.second - r.first;
    out->free_pixels = free_px;
    for (int k = 0; k < 5; ++k) out->phase_ms[k] = e->phase_ms[k];
    return WF_OK;
}

wf_status wf_engine_shading_queue(wf_engine* e, wf_shade_soa* dst, int32_t capacity,
                                  int32_t* n) {
    clear_error();
    WF_REQUIRE(e && dst && n, "wf_engine_shading_queue: NULL argument");
    WF_REQUIRE(capacity >= 0, "wf_engine_shading_queue: negative capacity");
    WF_REQUIRE(!e->launched, "wf_engine_shading_queue: a launch is pending (wf_engine_wait first)");
    hipStream_t st = e->ctx->stream;
    int32_t len = 0;
    if (e->last_q || e->last_index) {
        WF_HIP(hipMemcpyAsync(&len, e->view.counters + C_HITS, sizeof(int32_t),
                              hipMemcpyDeviceToHost, st));
        WF_HIP(hipStreamSynchronize(st));
    }
    *n = len;
    const size_t k = (size_t)std::min(len, capacity);
    if (k > 0) {
        // an index-mode iteration left its hits where they were traced, or in its hit records
        // (both intact until the next iteration): map them into a queue first (s0, which index
        // mode does not use; the records live in s1's memory)
        if (e->last_index)
            WF_TRY(gather_hits(e->ctx, e->scene, e->last_hits, Count{len, nullptr}, &e->s0,
                               e->st0));
        const wf_shade_soa& q = e->last_index ? e->s0 : *e->last_q;
        auto col = [&](void* to, const void* from, size_t elem) -> wf_status {
            if (to && from) WF_HIP(hipMemcpyAsync(to, from, elem * k, hipMemcpyDeviceToDevice, st));
            return WF_OK;
        };
        const void* src[] = {q.px, q.py, q.pz, q.nx, q.ny, q.nz, q.t, q.u, q.v, q.tri, q.mat,
                             q.tex, q.ox, q.oy, q.oz, q.dx, q.dy, q.dz, q.tr, q.tg, q.tb,
                             q.pixel, q.depth, q.pdf, q.geom};
        void* to[] = {dst->px, dst->py, dst->pz, dst->nx, dst->ny, dst->nz, dst->t, dst->u,
                      dst->v, dst->tri, dst->mat, dst->tex, dst->ox, dst->oy, dst->oz, dst->dx,
                      dst->dy, dst->dz, dst->tr, dst->tg, dst->tb, dst->pixel, dst->depth,
                      dst->pdf, dst->geom};
        for (size_t c = 0; c < sizeof(src) / sizeof(src[0]); ++c) WF_TRY(col(to[c], src[c], 4));
        WF_TRY(col(dst->front, q.front, 1));
        WF_HIP(hipStreamSynchronize(st));
    }
    return check_fault(e->ctx);
}

wf_status wf_engine_phase_timing(wf_engine* e, int32_t enable) {
    WF_REQUIRE(e, "wf_engine_phase_timing: NULL engine");
    e->phase_timing = enable != 0;
    for (double& m : e->phase_ms) m = 0.0;
    return WF_OK;
}

}  // extern "C"

namespace wf {

// wf_render with WF_SCHEDULE_REGENERATE: one job on an engine cached in the scene.
wf_status render_regenerate(wf_scene* s, const wf_render_params* rp, float* image_rgb,
                            int32_t* sample_counts, wf_render_stats* stats) {
    const int P = rp->width * rp->height;
    wf_engine* e = s->render_engine;
    WF_REQUIRE(lanes_ok(rp->sample_lanes), "wf_render: sample_lanes %d is not 0 or a power of two <= 64",
               rp->sample_lanes);
    WF_REQUIRE(rp->wavefront >= 0 && rp->wavefront < (1 << 29), "wf_render: wavefront %d out of range",
               rp->wavefront);
    const int lanes = lanes_of(rp->sample_lanes);
    // automatic wavefront: up to kAutoPathsPerPixel paths per pixel, at most kAutoWavefront rays
    // (larger launches amortise the per-iteration fixed costs and the traversal tail)
    const int slots =
        rp->wavefront > 0
            ? rp->wavefront
            : (int)std::min<int64_t>((int64_t)P * std::min(lanes, kAutoPathsPerPixel),
                                     std::max<int64_t>(P, kAutoWavefront));
    if (e && (e->cap != slots || e->px < P || e->lanes < lanes)) {
        wf_engine_destroy(e);
        e = s->render_engine = nullptr;
    }
    wf_engine_config cfg{};
    cfg.capacity = slots;
    cfg.max_jobs = 1;
    cfg.pixel_budget = P;
    cfg.policy = rp->policy;
    cfg.ray_sort = rp->ray_sort;
    cfg.use_graph = 1;
    cfg.sample_lanes = lanes;
    cfg.shade = rp->shade;
    if (!e) {
        WF_TRY(wf_engine_create(s->ctx, s, &cfg, &e));
        s->render_engine = e;
    } else {
        cfg.capacity = e->cap;
        cfg.pixel_budget = e->px;
        if (std::memcmp(&cfg, &e->cfg, sizeof(cfg)) != 0) {
            // the captured iteration bakes the policy and shading settings in
            drop_graph(e);
            // an engine created without ray ordering has no sort-word buffers yet
            if (cfg.ray_sort && !e->keys) {
                WF_TRY(engine_array(e, e->cap, &e->keys));
                WF_TRY(engine_array(e, e->cap, &e->keys2));
            }
            e->cfg = cfg;
        }
    }
    if (rp->coherence_stats && !e->coherence)
        WF_TRY(engine_alloc(e, sizeof(double) * 9, (void**)&e->coherence));
    if (e->coherence_mode != rp->coherence_stats) {
        drop_graph(e);   // the captured iteration bakes the statistics launches in or out
        e->coherence_mode = rp->coherence_stats;
    }
    if (e->coherence) WF_TRY(wf_fill_u32(e->ctx, e->coherence, 0u, 18));
    // phase times only while the context times kernels (they need events between the phases,
    // which a replayed graph does not have)
    WF_TRY(wf_engine_phase_timing(e, s->ctx->profile ? 1 : 0));
    wf_engine_stats st0{}, st1{};
    WF_TRY(wf_engine_query(e, &st0));
    hipEvent_t f0, f1;
    WF_HIP(hipEventCreate(&f0));
    WF_HIP(hipEventCreate(&f1));
    WF_HIP(hipEventRecord(f0, e->ctx->stream));
    wf_job_desc job{};
    job.width = rp->width;
    job.height = rp->height;
    job.spp = rp->spp;
    job.max_depth = rp->max_depth;
    job.seed = rp->shade.seed;
    job.job_tag = rp->shade.job;
    job.adaptive = rp->adaptive;
    job.adaptive_threshold = rp->adaptive_threshold;
    job.adaptive_min_samples = rp->adaptive_min_samples;
    job.adaptive_interval = rp->adaptive_interval;
    job.first_sample = rp->first_sample;
    job.camera = rp->camera;
    int32_t id = -1, fin = -1, nf = 0;
    WF_TRY(wf_engine_admit(e, &job, &id));
    while (nf == 0) WF_TRY(wf_engine_run(e, 0, &fin, 1, &nf));
    wf_job_result res{};
    WF_TRY(wf_engine_result(e, id, image_rgb, sample_counts, &res));
    WF_HIP(hipEventRecord(f1, e->ctx->stream));
    WF_HIP(hipEventSynchronize(f1));
    float fms = 0.0f;
    (void)hipEventElapsedTime(&fms, f0, f1);
    (void)hipEventDestroy(f0);
    (void)hipEventDestroy(f1);
    WF_TRY(wf_engine_query(e, &st1));
    if (stats) {
        wf_render_stats r{};
        r.rays_traced = st1.rays_traced - st0.rays_traced;
        r.shadow_rays = st1.shadow_rays - st0.shadow_rays;
        r.shaded_hits = st1.shaded_hits - st0.shaded_hits;
        r.misses = st1.misses - st0.misses;
        r.samples_run = res.samples_run;
        r.active_pixels_end = res.active_pixels_end;
        r.frame_ms = fms;
        r.sort_ms = st1.phase_ms[0];
        r.intersect_ms = st1.phase_ms[1];
        r.schedule_ms = st1.phase_ms[2];
        r.shade_ms = st1.phase_ms[3];
        r.miss_ms = st1.phase_ms[4];
        if (rp->coherence_stats) WF_TRY(coherence_read(e->ctx, e->coherence, &r));
        *stats = r;
    }
    return WF_OK;
}

}  // namespace wf
