// Queue scheduling: RayQueue::sortKey/schedule/materialize (include/scheduling/RayQueue.h:135-215)
// and ShadingQueue::schedule/materialize for the four policies (include/scheduling/ShadingQueue.h:
// 90-141, 233-329, 437-550). Every policy becomes an integer key plus one stable radix sort:
//   none        : identity
//   material    : typeRank(3 bits) | costRank(material)          (ShadingQueue.h:112-140,306-312)
//   texture     : typeRank(3 bits) | sizeRank(texture)           (ShadingQueue.h:122-134,315-316)
//   costBenefit : type(3 bits) | priorityRank | hit-point Morton (ShadingQueue.h:437-510)
// typeRank is computed on the device from per-type hit counts / texture bytes; cost and size
// ranks are dense descending ranks of the host tables; ties keep queue order.
#include <algorithm>
#include <mutex>

#include "wf_device.h"
#include "wf_internal.h"
#include "wf_rays.h"

namespace wf {

// ---- scratch of the allocation-free ray schedule (wf_ray_schedule_noalloc) ----------------------
// Two buffers of WF_NOALLOC_MAX_RAYS packed sort words (8 bytes each: the words, then the sort's
// other ping-pong buffer) and the sort's meta block for 8-bit digits over the 33-bit key: a tile
// counter per pass (64 words), the digit totals (4 passes x 512 slots, 2048 words) and the
// look-back states (512 slots x 512 tiles of 4096 words x 4 passes). The sort checks its need
// against kFixedMetaWords and reports an error rather than overrun. Each device holds its own copy
// until the process ends.
constexpr size_t kFixedWords = (size_t)WF_NOALLOC_MAX_RAYS;
constexpr size_t kFixedMetaWords = 64 + 2048 + (size_t)512 * (kFixedWords / 4096) * 4;
constexpr size_t kFixedScratchBytes = 2 * sizeof(uint64_t) * kFixedWords +
                                      sizeof(uint32_t) * kFixedMetaWords;
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

wf_status fixed_scratch_init(wf_context* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return set_error(WF_ERR_INVALID_ARGUMENT, "device %d beyond the scratch table",
                         ctx->device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(WF_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s",
                             ctx->device, hipGetErrorString(e));
        g_scratch_ptr[ctx->device] = ptr;
    }
    ctx->fixed_scratch = g_scratch_ptr[ctx->device];
    return WF_OK;
}

RayKeyFrame ray_key_frame(const float bmin[3], const float bmax[3], int32_t idx_bits) {
    RayKeyFrame f;
    for (int k = 0; k < 3; ++k) {
        // extent = max(max - min, 1e-6) per axis (RayQueue.h:143-145), host float math
        const float e = bmax[k] - bmin[k];
        f.lo[k] = bmin[k];
        f.ext[k] = e < 1e-6f ? 1e-6f : e;
    }
    f.idx_bits = idx_bits;
    return f;
}

int32_t index_bits(int64_t n) {
    int32_t b = 1;
    while ((1ll << b) < n) ++b;
    return b;
}

// keys[i] = the sort word of ray i (the bare key when f.idx_bits is 0); raw[i] = the bare key
// when raw is given
template <class R>
__global__ void k_ray_keys(R rays, int n_host, const int32_t* __restrict__ nd, RayKeyFrame f,
                           uint64_t* __restrict__ keys, uint64_t* __restrict__ raw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= load_count(n_host, nd)) return;
    V3 o, d;
    ld_od(rays, i, o, d);
    keys[i] = ray_sort_word(o, d, f, i);
    if (raw) raw[i] = ray_sort_key(o, d, f);
}

template <class R>
static wf_status ray_keys_any(wf_context* ctx, const R& r, Count n, const float bmin[3],
                              const float bmax[3], uint64_t* keys, int idx_bits = 0,
                              uint64_t* raw = nullptr) {
    if (n.n <= 0) return WF_OK;
    Timed tm(ctx, "ray_keys", 32.0, n);
    k_ray_keys<R><<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(
        r, n.n, n.dev, ray_key_frame(bmin, bmax, idx_bits), keys, raw);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

wf_status ray_keys(wf_context* ctx, const wf_ray_soa* r, Count n, const float bmin[3],
                   const float bmax[3], uint64_t* keys) {
    return ray_keys_any(ctx, *r, n, bmin, bmax, keys);
}

// ---- materialize ---------------------------------------------------------------------------------
__global__ void k_gather_rays(wf_ray_soa s, const int32_t* __restrict__ idx, int n_host,
                              const int32_t* __restrict__ nd, wf_ray_soa d,
                              const int32_t* __restrict__ tags_in, int32_t* __restrict__ tags_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= load_count(n_host, nd)) return;
    const int j = idx[i];
    if (tags_out) tags_out[i] = tags_in[j];
    d.ox[i] = s.ox[j];
    d.oy[i] = s.oy[j];
    d.oz[i] = s.oz[j];
    d.dx[i] = s.dx[j];
    d.dy[i] = s.dy[j];
    d.dz[i] = s.dz[j];
    d.tr[i] = s.tr[j];
    d.tg[i] = s.tg[j];
    d.tb[i] = s.tb[j];
    d.pixel[i] = s.pixel[j];
    d.depth[i] = s.depth[j];
    d.pdf[i] = s.pdf[j];
}

__global__ void k_gather_shade(wf_shade_soa s, const int32_t* __restrict__ idx, int n_host,
                               const int32_t* __restrict__ nd, wf_shade_soa d,
                               const int32_t* __restrict__ tags_in,
                               int32_t* __restrict__ tags_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= load_count(n_host, nd)) return;
    const int j = idx[i];
    if (tags_out) tags_out[i] = tags_in[j];
    d.px[i] = s.px[j];
    d.py[i] = s.py[j];
    d.pz[i] = s.pz[j];
    d.nx[i] = s.nx[j];
    d.ny[i] = s.ny[j];
    d.nz[i] = s.nz[j];
    d.t[i] = s.t[j];
    d.u[i] = s.u[j];
    d.v[i] = s.v[j];
    if (d.tri) d.tri[i] = s.tri[j];
    d.mat[i] = s.mat[j];
    d.tex[i] = s.tex[j];
    if (d.ox) {
        d.ox[i] = s.ox[j];
        d.oy[i] = s.oy[j];
        d.oz[i] = s.oz[j];
    }
    d.dx[i] = s.dx[j];
    d.dy[i] = s.dy[j];
    d.dz[i] = s.dz[j];
    d.tr[i] = s.tr[j];
    d.tg[i] = s.tg[j];
    d.tb[i] = s.tb[j];
    d.pixel[i] = s.pixel[j];
    d.depth[i] = s.depth[j];
    d.pdf[i] = s.pdf[j];
    if (d.geom) d.geom[i] = s.geom ? s.geom[j] : 0;
    if (d.front) d.front[i] = s.front ? s.front[j] : 1;
}

wf_status gather_rays(wf_context* ctx, const wf_ray_soa* src, const int32_t* idx, Count n,
                      wf_ray_soa* dst, const int32_t* tags_in, int32_t* tags_out) {
    if (n.n <= 0) return WF_OK;
    Timed tm(ctx, "gather_rays", tags_out ? 104.0 : 96.0, n);
    k_gather_rays<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(*src, idx, n.n, n.dev, *dst,
                                                              tags_in, tags_out);
    WF_HIP(hipGetLastError());
    return WF_OK;
}



wf_status gather_shading(wf_context* ctx, const wf_shade_soa* src, const int32_t* idx, Count n,
                         wf_shade_soa* dst, const int32_t* tags_in, int32_t* tags_out) {
    if (n.n <= 0) return WF_OK;
    Timed tm(ctx, "gather_shading", tags_out ? 200.0 : 192.0, n);
    k_gather_shade<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(*src, idx, n.n, n.dev, *dst,
                                                               tags_in, tags_out);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

// ---- material / texture policies ---------------------------------------------------------------
struct TypeStats {
    uint32_t count[8];
    unsigned long long bytes[8];
    int32_t rank[8];
};

__global__ void k_type_stats(const int32_t* __restrict__ mat, const int32_t* __restrict__ tex,
                             int n_host, const int32_t* __restrict__ nd,
                             const DevMaterial* __restrict__ mats,
                             const uint64_t* __restrict__ tex_mem, int n_tex,
                             TypeStats* __restrict__ st) {
    const int n = load_count(n_host, nd);
    __shared__ uint32_t cnt[8];
    __shared__ unsigned long long bytes[8];
    if (threadIdx.x < 8) {
        cnt[threadIdx.x] = 0;
        bytes[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int ty = mats[mat[i]].type & 7;
        atomicAdd(&cnt[ty], 1u);
        const int t = tex[i];
        if (t >= 0 && t < n_tex) atomicAdd(&bytes[ty], (unsigned long long)tex_mem[t]);
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        atomicAdd(&st->count[threadIdx.x], cnt[threadIdx.x]);
        atomicAdd(&st->bytes[threadIdx.x], bytes[threadIdx.x]);
    }
}

// computeTypeOrder (ShadingQueue.h:277-329): material = count desc, type desc;
// texture = bytes desc (equal sums: type desc).
__device__ void type_order(int policy, TypeStats* st) {
    int order[8];
    int m = 0;
    for (int t = 0; t < 8; ++t)
        if (st->count[t] > 0) order[m++] = t;
    for (int a = 1; a < m; ++a) {
        int x = order[a];
        int b = a - 1;
        while (b >= 0) {
            int y = order[b];
            bool x_first;
            if (policy == WF_POLICY_MATERIAL)
                x_first = st->count[x] != st->count[y] ? st->count[x] > st->count[y] : x > y;
            else
                x_first = st->bytes[x] != st->bytes[y] ? st->bytes[x] > st->bytes[y] : x > y;
            if (!x_first) break;
            order[b + 1] = order[b];
            --b;
        }
        order[b + 1] = x;
    }
    for (int t = 0; t < 8; ++t) st->rank[t] = 7;
    for (int r = 0; r < m; ++r) st->rank[order[r]] = r;
}

__global__ void k_type_rank(int policy, TypeStats* st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) type_order(policy, st);
}

__global__ void k_policy_keys(int policy, const int32_t* __restrict__ mat,
                              const int32_t* __restrict__ tex, int n_host,
                              const int32_t* __restrict__ nd,
                              const DevMaterial* __restrict__ mats,
                              const TypeStats* __restrict__ st,
                              const int32_t* __restrict__ cost_rank, int cost_bits,
                              const int32_t* __restrict__ size_rank, int size_bits, int n_tex,
                              uint64_t* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= load_count(n_host, nd)) return;
    const int m = mat[i];
    const uint64_t tr = (uint64_t)st->rank[mats[m].type & 7];
    uint64_t sub;
    int bits;
    if (policy == WF_POLICY_MATERIAL) {
        sub = (uint64_t)cost_rank[m];
        bits = cost_bits;
    } else {
        const int t = tex[i];
        sub = (uint64_t)size_rank[(t >= 0 && t < n_tex) ? t : n_tex];
        bits = size_bits;
    }
    keys[i] = (tr << bits) | sub;
}

// ---- cost-benefit policy ---------------------------------------------------------------------------
constexpr int kCbBlock = 256;
constexpr int kCbItems = 8;
constexpr int kCbTile = kCbBlock * kCbItems;
constexpr int kCbMaxMaterials = 1024;

__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Per-tile, per-material double sums of luminance(throughput) in a fixed order (wave chunks of 64
// in queue order, peers summed by their lowest lane in lane order), presence flags and hit-point
// bounds. Dynamic LDS: (kCbBlock / 64) * n_mat doubles + n_mat presence words.
__global__ __launch_bounds__(kCbBlock) void k_cb_partials(
    const int32_t* __restrict__ mat, const float* __restrict__ tr, const float* __restrict__ tg,
    const float* __restrict__ tb, const float* __restrict__ px, const float* __restrict__ py,
    const float* __restrict__ pz, int n_host, const int32_t* __restrict__ nd, int n_mat,
    double* __restrict__ partials, uint32_t* __restrict__ present,
    uint32_t* __restrict__ bounds) {
    const int n = load_count(n_host, nd);
    if ((int)blockIdx.x * kCbTile >= n) return;   // the priority sums stop at the last used tile
    extern __shared__ double cb_lds[];
    double* wsum = cb_lds;                                        // [wave][n_mat]
    uint32_t* pres = (uint32_t*)(cb_lds + (kCbBlock / 64) * n_mat);
    __shared__ double vals[kCbBlock];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < (kCbBlock / 64) * n_mat; k += kCbBlock) wsum[k] = 0.0;
    for (int k = threadIdx.x; k < n_mat; k += kCbBlock) pres[k] = 0u;
    __syncthreads();
    const uint64_t lt = lanemask_lt();
    float mn[3] = {3.0e38f, 3.0e38f, 3.0e38f}, mx[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    const int chunk = blockIdx.x * kCbTile + wave * 64 * kCbItems;
    for (int it = 0; it < kCbItems; ++it) {
        const int i = chunk + it * 64 + lane;
        const bool valid = i < n;
        const int m = valid ? mat[i] : 0;
        double x = 0.0;
        if (valid) {
            x = (double)luminance(V3(tr[i], tg[i], tb[i]));
            mn[0] = fminf(mn[0], px[i]); mx[0] = fmaxf(mx[0], px[i]);
            mn[1] = fminf(mn[1], py[i]); mx[1] = fmaxf(mx[1], py[i]);
            mn[2] = fminf(mn[2], pz[i]); mx[2] = fmaxf(mx[2], pz[i]);
        }
        vals[threadIdx.x] = x;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 10; ++b) {
            const bool bit = (m >> b) & 1;
            const uint64_t msk = __ballot(bit);
            peers &= bit ? msk : ~msk;
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && (peers & lt) == 0) {
            double acc = wsum[wave * n_mat + m];
            uint64_t p = peers;
            while (p) {
                const int l = __ffsll((unsigned long long)p) - 1;
                acc += vals[wave * 64 + l];
                p &= p - 1;
            }
            wsum[wave * n_mat + m] = acc;
            pres[m] = 1u;
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (int m = threadIdx.x; m < n_mat; m += kCbBlock) {
        double acc = 0.0;
        for (int w = 0; w < kCbBlock / 64; ++w) acc += wsum[w * n_mat + m];
        partials[(size_t)m * gridDim.x + blockIdx.x] = acc;   // material-major
        if (pres[m]) atomicOr(&present[m], 1u);
    }
    // hit-point bounds: wave reduction, block reduction, one atomic per bound per block
    __shared__ float wb[kCbBlock / 64][6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float a = mn[k], b = mx[k];
        for (int o = 32; o >= 1; o >>= 1) {
            a = fminf(a, __shfl_xor(a, o, 64));
            b = fmaxf(b, __shfl_xor(b, o, 64));
        }
        if (lane == 0) {
            wb[wave][k] = a;
            wb[wave][3 + k] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float r = wb[0][k];
        for (int w = 1; w < kCbBlock / 64; ++w) r = k < 3 ? fminf(r, wb[w][k]) : fmaxf(r, wb[w][k]);
        if (k < 3) atomicMin(&bounds[k], f2o(r));
        else atomicMax(&bounds[k], f2o(r));
    }
}

// Relative cost of texture t in costBenefit's combined cost: 0 without a texture; an id beyond
// the cost table is an untracked texture, which CostTracker::relativeCost rates 1.0 (the
// reference indexes its table with any id >= 0).
__device__ __forceinline__ double texture_cost(const double* __restrict__ tex_cost, int t,
                                               int n_tex) {
    return t < 0 ? 0.0 : (t < n_tex ? tex_cost[t] : 1.0);
}

// priority = benefit / max(relCost(mat) + relCost(tex of mat), 1e-6)   (ShadingQueue.h:458-471)
// One block per material: thread t sums tiles t, t + 256, ... in order, then a fixed pairwise
// tree over the threads (the same association on every run).
__global__ __launch_bounds__(256) void k_cb_priority(const double* __restrict__ partials,
                                                     int tiles, int n_host,
                                                     const int32_t* __restrict__ nd, int n_mat,
                                                     const DevMaterial* __restrict__ mats,
                                                     const double* __restrict__ mat_cost,
                                                     const double* __restrict__ tex_cost,
                                                     int n_tex, double* __restrict__ prio) {
    __shared__ double red[256];
    const int m = blockIdx.x;
    const int used = (load_count(n_host, nd) + kCbTile - 1) / kCbTile;   // tiles with entries
    double acc = 0.0;
    for (int t = threadIdx.x; t < used; t += 256) acc += partials[(size_t)m * tiles + t];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double s = red[0];
    const double sc = mat_cost[m];
    const double c = sc + texture_cost(tex_cost, mats[m].texture_id, n_tex);
    prio[m] = s / (c > 1e-6 ? c : 1e-6);
}

// Competition rank by descending priority among materials present in the queue.
__global__ void k_cb_rank(const double* __restrict__ prio, const uint32_t* __restrict__ present,
                          int n_mat, int32_t* __restrict__ rank) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_mat) return;
    const double p = prio[m];
    int r = 0;
    for (int k = 0; k < n_mat; ++k)
        if (present[k] && prio[k] > p) ++r;
    rank[m] = r;
}

// computeMortonCode (ShadingQueue.h:513-550)
__device__ __forceinline__ uint32_t quantise_hit(float v, float lo, float hi) {
    const float t = (hi > lo) ? (v - lo) / (hi - lo) : 0.0f;
    return (uint32_t)(clamp01(t) * 1023.0f);
}

__global__ void k_cb_keys(const int32_t* __restrict__ mat, const float* __restrict__ px,
                          const float* __restrict__ py, const float* __restrict__ pz, int n_host,
                          const int32_t* __restrict__ nd, const DevMaterial* __restrict__ mats,
                          const int32_t* __restrict__ rank, int rank_bits,
                          const uint32_t* __restrict__ bounds, int spatial,
                          uint64_t* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= load_count(n_host, nd)) return;
    if (!spatial) {   // group key only: a stable sort keeps the input order inside each group
        const int m = mat[i];
        keys[i] = ((uint64_t)(mats[m].type & 7) << rank_bits) | (uint64_t)rank[m];
        return;
    }
    const float lx = o2f(bounds[0]), ly = o2f(bounds[1]), lz = o2f(bounds[2]);
    const float hx = o2f(bounds[3]), hy = o2f(bounds[4]), hz = o2f(bounds[5]);
    const uint32_t morton = spread_bits10(quantise_hit(px[i], lx, hx)) |
                            (spread_bits10(quantise_hit(py[i], ly, hy)) << 1) |
                            (spread_bits10(quantise_hit(pz[i], lz, hz)) << 2);
    const int m = mat[i];
    keys[i] = ((uint64_t)(mats[m].type & 7) << (rank_bits + 30)) | ((uint64_t)rank[m] << 30) |
              (uint64_t)morton;
}

// ---- shading order as per-material ranges (regenerating wavefront) ------------------------------
// Every policy's group key is a function of the hit's material (the engine's costBenefit has no
// spatial term), so the engine does not sort: it counts the hits per material, gives each
// distinct group key one range of the shading queue, the ranges in key order, and appends each
// hit into its key's range (k_split_buckets). Inside a range the order is the append order; the
// queue order of the wavefront changes no result.
constexpr int kPolicyCountBlock = 256;
// copies of the per-material sums: block b adds into copy b % kCountCopies, so that each address
// takes few atomics (device-scope atomics on one address serialise)
constexpr int kCountCopies = 16;

// Hits per material of the traced rays (leaf >= 0) and, with benefit_on, the FP64 sums of their
// throughput luminance (benefit of ShadingQueue.h:458-471); LDS histograms flushed with one atomic
// per material and block into the block's copy. Dynamic LDS: n_mat doubles (benefit_on) then
// n_mat counts.
template <class R>
__global__ __launch_bounds__(kPolicyCountBlock) void k_policy_count(
    R rays, int n_host, const int32_t* __restrict__ nd, const int32_t* __restrict__ leaf,
    const int32_t* __restrict__ leaf_mat, int n_mat, int benefit_on,
    uint32_t* __restrict__ counts, double* __restrict__ benefit) {
    extern __shared__ double pc_lds[];
    double* ben = pc_lds;
    uint32_t* cnt = (uint32_t*)(pc_lds + (benefit_on ? n_mat : 0));
    for (int k = threadIdx.x; k < n_mat; k += kPolicyCountBlock) {
        cnt[k] = 0u;
        if (benefit_on) ben[k] = 0.0;
    }
    __syncthreads();
    const int n = load_count(n_host, nd);
    for (int i = blockIdx.x * kPolicyCountBlock + threadIdx.x; i < n;
         i += gridDim.x * kPolicyCountBlock) {
        const int L = leaf[i];
        if (L < 0) continue;
        const int m = leaf_mat[L];
        atomicAdd(&cnt[m], 1u);
        if (benefit_on) atomicAdd(&ben[m], (double)luminance(ld_ray(rays, i).t));
    }
    __syncthreads();
    const size_t copy = (size_t)(blockIdx.x % kCountCopies) * n_mat;
    for (int k = threadIdx.x; k < n_mat; k += kPolicyCountBlock) {
        if (cnt[k]) atomicAdd(&counts[copy + k], cnt[k]);
        if (benefit_on && ben[k] != 0.0) atomicAdd(&benefit[copy + k], ben[k]);
    }
}

// Group key of every material under the policy -- the sort keys above without their per-entry
// part: typeRank | costRank (material), typeRank | sizeRank(texture of the material) (texture),
// type | priority rank (costBenefit) -- then one range per distinct key among the present
// materials, in key order: bucket[m] = the range of material m, cursor[b] = the first slot of
// range b. One block: thread m gathers material m, then a group of lanes per material runs the
// loops over all materials.
__global__ __launch_bounds__(kPolicyBucketMax) void k_policy_plan(
    int policy, int n_mat, const uint32_t* __restrict__ counts, const double* __restrict__ benefit,
    const DevMaterial* __restrict__ mats, const int32_t* __restrict__ cost_rank, int cost_bits,
    const int32_t* __restrict__ size_rank, int size_bits, int n_tex,
    const uint64_t* __restrict__ tex_mem, const double* __restrict__ mat_cost,
    const double* __restrict__ tex_cost, int rank_bits, int32_t* __restrict__ bucket,
    int32_t* __restrict__ cursor, int32_t* __restrict__ hits_dev) {
    __shared__ uint32_t s_cnt[kPolicyBucketMax];
    __shared__ uint64_t s_key[kPolicyBucketMax];
    __shared__ double s_prio[kPolicyBucketMax];
    __shared__ uint8_t s_first[kPolicyBucketMax];
    __shared__ int32_t s_type[kPolicyBucketMax];
    __shared__ TypeStats st;
    __shared__ uint32_t s_total;
    const int m = threadIdx.x;
    const bool in = m < n_mat;
    const bool cb = policy == WF_POLICY_COST_BENEFIT;
    // every global read up front (independent loads overlap)
    uint32_t cc[kCountCopies];
    double bb[kCountCopies];
#pragma unroll
    for (int k = 0; k < kCountCopies; ++k) {
        cc[k] = in ? counts[(size_t)k * n_mat + m] : 0u;
        bb[k] = in && cb ? benefit[(size_t)k * n_mat + m] : 0.0;
    }
    const DevMaterial* mp = mats + (in ? m : 0);
    const int ty = in ? (mp->type & 7) : 0;
    const int tex = in ? mp->texture_id : -1;
    const bool tex_ok = tex >= 0 && tex < n_tex;
    const int32_t crank = in && policy == WF_POLICY_MATERIAL ? cost_rank[m] : 0;
    const int32_t srank = in && policy == WF_POLICY_TEXTURE ? size_rank[tex_ok ? tex : n_tex] : 0;
    const uint64_t tbytes = in && tex_ok && !cb ? tex_mem[tex] : 0ull;
    const double mcost = in && cb ? mat_cost[m] : 0.0;
    const double tcost = in && cb ? texture_cost(tex_cost, tex, n_tex) : 0.0;
    uint32_t c = 0u;
    double ben = 0.0;
#pragma unroll
    for (int k = 0; k < kCountCopies; ++k) {
        c += cc[k];
        ben += bb[k];
    }
    s_cnt[m] = c;
    if (m < 8) {
        st.count[m] = 0u;
        st.bytes[m] = 0ull;
    }
    if (m == 0) s_total = 0u;
    __syncthreads();
    if (c) atomicAdd(&s_total, c);
    if (cb) {
        if (in) {
            const double cost = mcost + tcost;
            s_prio[m] = ben / (cost > 1e-6 ? cost : 1e-6);
        }
    } else if (c) {
        // per-type hit counts and texture bytes (k_type_stats over the hits, grouped by material)
        atomicAdd(&st.count[ty], c);
        if (tex_ok) atomicAdd(&st.bytes[ty], (unsigned long long)c * tbytes);
    }
    __syncthreads();
    if (!cb && m == 0) type_order(policy, &st);
    __syncthreads();
    if (in) {
        if (policy == WF_POLICY_MATERIAL)
            s_key[m] = ((uint64_t)st.rank[ty] << cost_bits) | (uint64_t)crank;
        else if (policy == WF_POLICY_TEXTURE)
            s_key[m] = ((uint64_t)st.rank[ty] << size_bits) | (uint64_t)srank;
        s_type[m] = ty;
    }
    // The loops over all materials below run on a group of G lanes per material (G a power of
    // two, at most a wave, G * materials <= the block): lane `sub` of group g takes the
    // materials k = sub, sub + G, ... and the group sums its counts with shuffles (integer sums:
    // the same result in any order).
    int gsh = 0;
    while ((64 >> gsh) * n_mat > kPolicyBucketMax) ++gsh;
    const int G = 64 >> gsh;
    const int g = m / G, sub = m % G;
    const bool gin = g < n_mat;
    auto group_sum = [&](uint32_t v) {
        for (int o = 1; o < G; o <<= 1) v += __shfl_xor(v, o, 64);
        return v;
    };
    if (cb) {
        // competition rank by descending priority among present materials
        __syncthreads();
        uint32_t r = 0;
        if (gin)
            for (int k = sub; k < n_mat; k += G)
                if (s_cnt[k] && s_prio[k] > s_prio[g]) ++r;
        r = group_sum(r);
        if (gin && sub == 0) s_key[g] = ((uint64_t)s_type[g] << rank_bits) | (uint64_t)r;
    }
    __syncthreads();
    // the first present material of each key stands for its range
    const uint32_t gc = gin ? s_cnt[g] : 0u;
    const uint64_t key = gin ? s_key[g] : 0ull;
    uint32_t before = 0;   // present materials of the same key before g
    if (gc)
        for (int k = sub; k < g; k += G)
            if (s_cnt[k] && s_key[k] == key) ++before;
    before = group_sum(before);
    const bool first = gc != 0 && before == 0;
    if (gin && sub == 0) s_first[g] = first ? 1 : 0;
    __syncthreads();
    uint32_t b = 0, slot = 0;
    if (gc)
        for (int k = sub; k < n_mat; k += G) {
            if (!s_cnt[k] || s_key[k] >= key) continue;
            b += s_first[k];
            slot += s_cnt[k];
        }
    b = group_sum(b);
    slot = group_sum(slot);
    if (gc && sub == 0) {
        bucket[g] = (int32_t)b;
        if (first) cursor[b] = (int32_t)slot;
    }
    if (m == 0) *hits_dev = (int32_t)s_total;
}

static int bits_for(int max_value) {
    int b = 0;
    while ((1 << b) <= max_value) ++b;
    return b;
}

wf_status shading_keys(wf_context* ctx, const wf_scene* s, const wf_shade_soa* q, Count n,
                       int32_t policy, bool spatial, uint64_t* keys, int32_t* key_bits) {
    const SceneDev& d = s->dev;
    if (policy == WF_POLICY_MATERIAL || policy == WF_POLICY_TEXTURE) {
        void* p = nullptr;
        WF_TRY(workspace(ctx, "type_stats", sizeof(TypeStats), &p));
        TypeStats* st = (TypeStats*)p;
        WF_TRY(wf_fill_u32(ctx, st, 0u, sizeof(TypeStats) / 4));
        {
            Timed tm(ctx, "policy_keys", policy == WF_POLICY_MATERIAL ? 8.0 : 12.0, n);
            int blocks = (int)std::min<int64_t>(grid_for(n.n, 256), 1024);
            k_type_stats<<<blocks, 256, 0, ctx->stream>>>(q->mat, q->tex, n.n, n.dev,
                                                          d.materials, d.tex_mem_bytes,
                                                          d.n_textures, st);
            WF_HIP(hipGetLastError());
            k_type_rank<<<1, 1, 0, ctx->stream>>>(policy, st);
            WF_HIP(hipGetLastError());
            k_policy_keys<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(
                policy, q->mat, q->tex, n.n, n.dev, d.materials, st, d.mat_cost_rank,
                d.cost_rank_bits, d.tex_size_rank, d.size_rank_bits, d.n_textures, keys);
            WF_HIP(hipGetLastError());
        }
        *key_bits = 3 + (policy == WF_POLICY_MATERIAL ? d.cost_rank_bits : d.size_rank_bits);
        return WF_OK;
    }
    if (policy == WF_POLICY_COST_BENEFIT) {
        const int nm = d.n_materials;
        if (nm <= 0) return set_error(WF_ERR_INVALID_ARGUMENT, "costBenefit: scene has no materials");
        if (nm > kCbMaxMaterials)
            return set_error(WF_ERR_UNSUPPORTED,
                             "costBenefit policy supports at most %d materials (scene has %d)",
                             kCbMaxMaterials, nm);
        const int tiles = (n.n + kCbTile - 1) / kCbTile;
        void *pp, *ppres, *pb, *pprio, *prank;
        WF_TRY(workspace(ctx, "cb_partials", sizeof(double) * (size_t)tiles * nm, &pp));
        WF_TRY(workspace(ctx, "cb_present", sizeof(uint32_t) * nm, &ppres));
        WF_TRY(workspace(ctx, "cb_bounds", sizeof(uint32_t) * 8, &pb));
        WF_TRY(workspace(ctx, "cb_prio", sizeof(double) * nm, &pprio));
        WF_TRY(workspace(ctx, "cb_rank", sizeof(int32_t) * nm, &prank));
        WF_TRY(wf_fill_u32(ctx, ppres, 0u, (size_t)nm));
        WF_TRY(wf_fill_u32(ctx, pb, 0xFFFFFFFFu, 3));
        WF_TRY(wf_fill_u32(ctx, (uint32_t*)pb + 3, 0u, 5));
        Timed tm(ctx, "policy_keys", 32.0, n);
        const size_t lds = sizeof(double) * (kCbBlock / 64) * (size_t)nm + sizeof(uint32_t) * nm;
        k_cb_partials<<<tiles, kCbBlock, lds, ctx->stream>>>(q->mat, q->tr, q->tg, q->tb, q->px,
                                                           q->py, q->pz, n.n, n.dev, nm,
                                                           (double*)pp, (uint32_t*)ppres,
                                                           (uint32_t*)pb);
        WF_HIP(hipGetLastError());
        k_cb_priority<<<nm, 256, 0, ctx->stream>>>((double*)pp, tiles, n.n, n.dev, nm,
                                                    d.materials, d.mat_rel_cost, d.tex_rel_cost,
                                                    d.n_textures, (double*)pprio);
        WF_HIP(hipGetLastError());
        k_cb_rank<<<grid_for(nm, 256), 256, 0, ctx->stream>>>((double*)pprio, (uint32_t*)ppres,
                                                              nm, (int32_t*)prank);
        WF_HIP(hipGetLastError());
        const int rb = nm <= 1 ? 0 : bits_for(nm - 1);
        k_cb_keys<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(q->mat, q->px, q->py, q->pz, n.n,
                                                               n.dev, d.materials,
                                                               (int32_t*)prank, rb,
                                                               (uint32_t*)pb, spatial ? 1 : 0,
                                                               keys);
        WF_HIP(hipGetLastError());
        *key_bits = 3 + rb + (spatial ? 30 : 0);
        return WF_OK;
    }
    return set_error(WF_ERR_INVALID_ARGUMENT, "unknown scheduling policy %d", policy);
}

wf_status shading_schedule(wf_context* ctx, const wf_scene* s, const wf_shade_soa* q, Count n,
                           int32_t policy, int32_t* sorted, bool spatial) {
    if (n.n <= 0) return WF_OK;
    if (policy == WF_POLICY_NONE) return iota_count(ctx, sorted, n);
    void *pk, *pk2;
    WF_TRY(workspace(ctx, "sched_keys", sizeof(uint64_t) * n.n, &pk));
    WF_TRY(workspace(ctx, "sched_keys2", sizeof(uint64_t) * n.n, &pk2));
    int32_t bits = 0;
    WF_TRY(shading_keys(ctx, s, q, n, policy, spatial, (uint64_t*)pk, &bits));
    return radix_sort_pairs(ctx, (uint64_t*)pk, nullptr, (uint64_t*)pk2, sorted, n, bits,
                            "shading_sort");
}

template <class R>
static wf_status policy_buckets_any(wf_context* ctx, const wf_scene* s, const R& rays, Count n,
                                    const int32_t* leaf, int32_t policy, int32_t* hits_dev,
                                    PolicyRanges* out, bool* used) {
    const SceneDev& d = s->dev;
    const int nm = d.n_materials;
    *used = false;
    if (policy == WF_POLICY_NONE || nm <= 0 || nm > kPolicyBucketMax || !d.leaf_mat) return WF_OK;
    if (policy < WF_POLICY_NONE || policy > WF_POLICY_COST_BENEFIT)
        return set_error(WF_ERR_INVALID_ARGUMENT, "unknown scheduling policy %d", policy);
    const bool cb = policy == WF_POLICY_COST_BENEFIT;
    void *pc, *pb, *pcur, *pbk;
    WF_TRY(workspace(ctx, "bucket_counts", sizeof(uint32_t) * kCountCopies * nm, &pc));
    WF_TRY(workspace(ctx, "bucket_benefit", sizeof(double) * kCountCopies * nm, &pb));
    WF_TRY(workspace(ctx, "bucket_cursor", sizeof(int32_t) * nm, &pcur));
    WF_TRY(workspace(ctx, "bucket_of", sizeof(int32_t) * nm, &pbk));
    WF_TRY(wf_fill_u32(ctx, pc, 0u, (size_t)kCountCopies * nm));
    if (cb) WF_TRY(wf_fill_u32(ctx, pb, 0u, 2 * (size_t)kCountCopies * nm));
    {
        // per ray: its leaf (+ the throughput column for costBenefit)
        Timed tm(ctx, "policy_count", cb ? 20.0 : 4.0, n);
        const unsigned grid = std::min(grid_for(n.n, kPolicyCountBlock), (unsigned)ctx->cu_count * 16);
        const size_t lds = (cb ? sizeof(double) * nm : 0) + sizeof(uint32_t) * nm;
        k_policy_count<R><<<grid, kPolicyCountBlock, lds, ctx->stream>>>(
            rays, n.n, n.dev, leaf, d.leaf_mat, nm, cb ? 1 : 0, (uint32_t*)pc, (double*)pb);
        WF_HIP(hipGetLastError());
    }
    {
        Timed tm(ctx, "policy_plan", 16.0 * nm);
        k_policy_plan<<<1, kPolicyBucketMax, 0, ctx->stream>>>(
            policy, nm, (const uint32_t*)pc, (const double*)pb, d.materials, d.mat_cost_rank,
            d.cost_rank_bits, d.tex_size_rank, d.size_rank_bits, d.n_textures, d.tex_mem_bytes,
            d.mat_rel_cost, d.tex_rel_cost, nm <= 1 ? 0 : bits_for(nm - 1), (int32_t*)pbk,
            (int32_t*)pcur, hits_dev);
        WF_HIP(hipGetLastError());
    }
    out->bucket = (const int32_t*)pbk;
    out->cursor = (int32_t*)pcur;
    *used = true;
    return WF_OK;
}

wf_status policy_buckets(wf_context* ctx, const wf_scene* s, const RayPacked* rays, Count n,
                         const int32_t* hit_leaf, int32_t policy, int32_t* hits_dev,
                         PolicyRanges* out, bool* used) {
    return policy_buckets_any(ctx, s, *rays, n, hit_leaf, policy, hits_dev, out, used);
}

wf_status policy_buckets(wf_context* ctx, const wf_scene* s, const RayPackedOrdered* rays, Count n,
                         const int32_t* hit_leaf, int32_t policy, int32_t* hits_dev,
                         PolicyRanges* out, bool* used) {
    return policy_buckets_any(ctx, s, *rays, n, hit_leaf, policy, hits_dev, out, used);
}

template <class R>
static wf_status ray_schedule_any(wf_context* ctx, const R& rays, Count n, const float bmin[3],
                                  const float bmax[3], int32_t* sorted, uint64_t* keys_out) {
    if (n.n <= 0) return WF_OK;
    // packed words: the 33-bit key above the ray's index (stable order on equal keys)
    const int idx_bits = index_bits(n.n);
    void *pk, *pk2;
    WF_TRY(workspace(ctx, "ray_keys", sizeof(uint64_t) * n.n, &pk));
    WF_TRY(workspace(ctx, "ray_keys2", sizeof(uint64_t) * n.n, &pk2));
    WF_TRY(ray_keys_any(ctx, rays, n, bmin, bmax, (uint64_t*)pk, idx_bits, keys_out));
    return radix_sort_packed(ctx, (uint64_t*)pk, (uint64_t*)pk2, sorted, n, idx_bits, 33,
                             "ray_sort");
}

wf_status ray_schedule(wf_context* ctx, const wf_ray_soa* rays, Count n, const float bmin[3],
                       const float bmax[3], int32_t* sorted, uint64_t* keys_out) {
    return ray_schedule_any(ctx, *rays, n, bmin, bmax, sorted, keys_out);
}

wf_status ray_schedule(wf_context* ctx, const RayPacked* rays, Count n, const float bmin[3],
                       const float bmax[3], int32_t* sorted, uint64_t* keys_out) {
    return ray_schedule_any(ctx, *rays, n, bmin, bmax, sorted, keys_out);
}

}  // namespace wf

using namespace wf;

extern "C" {

wf_status wf_ray_schedule(wf_context* ctx, const wf_ray_soa* rays, int32_t n,
                          const float bounds_min[3], const float bounds_max[3],
                          int32_t* sorted_indices, uint64_t* keys_out) {
    clear_error();
    WF_REQUIRE(ctx && rays && bounds_min && bounds_max && sorted_indices,
               "wf_ray_schedule: NULL argument");
    WF_REQUIRE(n >= 0, "wf_ray_schedule: negative ray count");
    WF_TRY(ray_schedule(ctx, rays, host_count(n), bounds_min, bounds_max, sorted_indices,
                        keys_out));
    return check_fault(ctx);
}

wf_status wf_ray_schedule_noalloc(wf_context* ctx, const wf_ray_soa* rays, int32_t n,
                                  const float bounds_min[3], const float bounds_max[3],
                                  int32_t* sorted_indices, uint64_t* keys_out) {
    clear_error();
    WF_REQUIRE(ctx && rays && bounds_min && bounds_max && sorted_indices,
               "wf_ray_schedule_noalloc: NULL argument");
    WF_REQUIRE(n >= 0, "wf_ray_schedule_noalloc: negative ray count");
    WF_REQUIRE(n <= WF_NOALLOC_MAX_RAYS,
               "wf_ray_schedule_noalloc: %d rays exceed the fixed scratch of %d rays", n,
               WF_NOALLOC_MAX_RAYS);
    WF_REQUIRE(!ctx->profile, "wf_ray_schedule_noalloc: kernel timing is on (it records events)");
    WF_REQUIRE(ctx->fixed_scratch, "wf_ray_schedule_noalloc: no scratch on device %d",
               ctx->device);
    if (n == 0) return WF_OK;
    unsigned char* base = (unsigned char*)ctx->fixed_scratch;
    uint64_t* words = (uint64_t*)base;
    SortArena arena;
    arena.keys_alt = (uint64_t*)(base + sizeof(uint64_t) * kFixedWords);
    arena.meta = (uint32_t*)(base + 2 * sizeof(uint64_t) * kFixedWords);
    arena.meta_words = kFixedMetaWords;
    const Count cnt = host_count(n);
    const int idx_bits = index_bits(n);
    k_ray_keys<wf_ray_soa><<<grid_for(n, 256), 256, 0, ctx->stream>>>(
        *rays, n, nullptr, ray_key_frame(bounds_min, bounds_max, idx_bits), words, keys_out);
    WF_HIP(hipGetLastError());
    WF_TRY(radix_sort_packed_fixed(ctx, words, sorted_indices, cnt, idx_bits, 33, arena,
                                   "ray_sort"));
    return WF_OK;
}

wf_status wf_materialize_rays(wf_context* ctx, const wf_ray_soa* src, const int32_t* idx,
                              int32_t n, wf_ray_soa* dst) {
    clear_error();
    WF_REQUIRE(ctx && src && idx && dst, "wf_materialize_rays: NULL argument");
    WF_REQUIRE(n >= 0, "wf_materialize_rays: negative count");
    WF_TRY(gather_rays(ctx, src, idx, host_count(n), dst));
    WF_HIP(hipStreamSynchronize(ctx->stream));
    return WF_OK;
}

wf_status wf_shading_schedule(const wf_scene* s, const wf_shade_soa* q, int32_t n, int32_t policy,
                              int32_t* sorted_indices) {
    clear_error();
    WF_REQUIRE(s && q && sorted_indices, "wf_shading_schedule: NULL argument");
    WF_REQUIRE(n >= 0, "wf_shading_schedule: negative count");
    WF_REQUIRE(policy >= WF_POLICY_NONE && policy <= WF_POLICY_COST_BENEFIT,
               "wf_shading_schedule: unknown policy %d", policy);
    WF_TRY(shading_schedule(s->ctx, s, q, host_count(n), policy, sorted_indices));
    return check_fault(s->ctx);
}

wf_status wf_materialize_shading(wf_context* ctx, const wf_shade_soa* src, const int32_t* idx,
                                 int32_t n, wf_shade_soa* dst) {
    clear_error();
    WF_REQUIRE(ctx && src && idx && dst, "wf_materialize_shading: NULL argument");
    WF_REQUIRE(n >= 0, "wf_materialize_shading: negative count");
    WF_TRY(gather_shading(ctx, src, idx, host_count(n), dst));
    WF_HIP(hipStreamSynchronize(ctx->stream));
    return WF_OK;
}

}  // extern "C"
