// Ray traversal replacing EmbreeAccelerator::intersect4/intersect/occluded
// (src/accelerator/Embree.cpp:122-279) and the hit/miss split of WavefrontRenderer::intersectAll
// (src/render/WavefrontRenderer.cpp:459-567).
//
// One lane per ray over the four- or eight-wide BVH. The top of the traversal stack lives in LDS, one
// column per lane (stack[entry][lane] so that a wave's pushes and pops hit 64 different banks),
// deeper entries in a per-lane private array sized from the build's exact stack bound. The triangle test
// is the reference's Moller-Trumbore (include/geometry/Triangle.h:6-42) with the same float
// operations; closest hit = lexicographic minimum of (t, global triangle index) over every
// triangle the test accepts in [tnear, tfar], which makes the result independent of traversal
// order. Box tests are conservative (boxes are padded at build time).
#include "wf_rays.h"
#include "wf_sky.h"
#include "wf_traverse.h"

namespace wf {

// Persistent closest-hit kernel: each lane walks one ray at a time and takes a new ray as soon as
// its walk ends (dynamic refill keeps lanes busy while the slowest rays of a wave finish).
template <class NODE, int STACK, class R>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(
    const NODE* __restrict__ nodes, const float4* __restrict__ tri, int n_tris, R rays,
    int n_host, const int32_t* __restrict__ nd, float tnear, float tfar,
    int32_t* __restrict__ out_leaf, float* __restrict__ out_t, float* __restrict__ out_u,
    float* __restrict__ out_v, uint32_t* __restrict__ out_flag, int32_t* __restrict__ counter,
    int refill_lanes) {
    __shared__ int32_t stack[lds_entries<STACK>() * kTraceBlock];
    __shared__ int4 xchg[kTraceBlock / 64 * kXchgPerWave];   // candidate tests and stashes
    volatile int32_t spill[Walker<false, STACK>::X > 0 ? Walker<false, STACK>::X : 1];
    closest_walks<NODE, STACK>(nodes, tri, n_tris, rays, load_count(n_host, nd), tnear, tfar,
                               out_leaf, out_t, out_u, out_v, out_flag, counter, refill_lanes,
                               stack, xchg, spill);
}

// occluded(): tnear 0.001, tfar = maxDistance - 0.001; a negative tfar reports occluded
// (Embree.cpp:268-278).
template <class NODE, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_trace_occluded(
    const NODE* __restrict__ nodes, const float4* __restrict__ tri, int n_tris,
    const float* __restrict__ ox, const float* __restrict__ oy, const float* __restrict__ oz,
    const float* __restrict__ dx, const float* __restrict__ dy, const float* __restrict__ dz,
    const float* __restrict__ maxd, int n, uint8_t* __restrict__ out) {
    __shared__ int32_t stack[lds_entries<STACK>() * kTraceBlock];
    const int i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const float tfar = maxd[i] - 0.001f;
    bool occ = tfar < 0.0f;
    if (!occ) {
        TraceHit h;
        occ = traverse<true, STACK>(nodes, tri, n_tris, V3(ox[i], oy[i], oz[i]),
                             V3(dx[i], dy[i], dz[i]), 0.001f, tfar, stack, h);
    }
    out[i] = occ ? 1 : 0;
}

// Traversal statistics: BVH nodes visited and triangles tested per ray (lanes of a wave are rays
// 64w..64w+63, so SIMD efficiency follows from the per-ray counts).
template <class NODE, int STACK, bool ANY>
__global__ __launch_bounds__(kTraceBlock) void k_trace_stats(
    const NODE* __restrict__ nodes, const float4* __restrict__ tri, int n_tris,
    wf_ray_soa r, int n, float tnear, float tfar, int32_t* __restrict__ out_nodes,
    int32_t* __restrict__ out_tris) {
    __shared__ int32_t stack[lds_entries<STACK>() * kTraceBlock];
    const int i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    TraceHit h;
    traverse<ANY, STACK, true>(nodes, tri, n_tris, V3(r.ox[i], r.oy[i], r.oz[i]),
                               V3(r.dx[i], r.dy[i], r.dz[i]), tnear, tfar, stack, h);
    out_nodes[i] = h.nodes;
    out_tris[i] = h.tris;
}

// HitRecord mapping (map_hit_vals) of ray i into shading-queue slot j, with the ray's state.
template <class R>
__device__ __forceinline__ void map_hit(const SceneDev& s, const R& in, int i, int L,
                                        float t, float bu, float bv, wf_shade_soa& out, int j) {
    const RayVal ray = ld_ray(in, i);
    const MappedHit h = map_hit_vals(s, ray.o, ray.d, L, t, bu, bv);
    st3(out.px, out.py, out.pz, j, h.P);
    st3(out.nx, out.ny, out.nz, j, h.N);
    out.t[j] = t;
    out.u[j] = h.u;
    out.v[j] = h.v;
    if (out.tri) out.tri[j] = h.g - s.mesh_tri_offset[h.geom];
    out.mat[j] = h.mat;
    out.tex[j] = s.materials[h.mat].texture_id;
    if (out.ox) st3(out.ox, out.oy, out.oz, j, ray.o);
    st3(out.dx, out.dy, out.dz, j, ray.d);
    out.tr[j] = ray.t.x;
    out.tg[j] = ray.t.y;
    out.tb[j] = ray.t.z;
    out.pixel[j] = ray.pixel;
    out.depth[j] = ray.depth;
    out.pdf[j] = ray.pdf;
    if (out.geom) out.geom[j] = h.geom;
    if (out.front) out.front[j] = h.front ? 1 : 0;
}

// Stable split by the exclusive prefix of hit flags (intersectAll's hit and miss queues).
__global__ void k_split(const SceneDev s, wf_ray_soa in, int n_host, const int32_t* __restrict__ nd,
                        const int32_t* __restrict__ leaf, const float* __restrict__ ht,
                        const float* __restrict__ hu, const float* __restrict__ hv,
                        const uint32_t* __restrict__ prefix, wf_shade_soa out, wf_ray_soa miss,
                        const int32_t* __restrict__ tags_in, int32_t* __restrict__ hit_tags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= load_count(n_host, nd)) return;
    const int L = leaf[i];
    const uint32_t pos = prefix[i];
    if (L >= 0) {
        map_hit(s, in, i, L, ht[i], hu[i], hv[i], out, (int)pos);
        if (hit_tags) hit_tags[pos] = tags_in[i];
    } else {
        const int m = i - (int)pos;
        miss.ox[m] = in.ox[i];
        miss.oy[m] = in.oy[i];
        miss.oz[m] = in.oz[i];
        miss.dx[m] = in.dx[i];
        miss.dy[m] = in.dy[i];
        miss.dz[m] = in.dz[i];
        miss.tr[m] = in.tr[i];
        miss.tg[m] = in.tg[i];
        miss.tb[m] = in.tb[i];
        miss.pixel[m] = in.pixel[i];
        miss.depth[m] = in.depth[i];
        miss.pdf[m] = in.pdf[i];
    }
}

// Unordered split: hits and misses appended at counts[0] / counts[1], kAppendItems rays per
// thread (coalesced rounds), one cursor atomic per block and list.
constexpr int kAppendBlock = 1024;
constexpr int kAppendItems = 4;

// The index-mode split: the traced-ray index of each hit appended to hit_list; each miss is
// resolved right here (the sky or environment added to its accumulator, as k_miss does) and its
// accumulator index and job tag appended for path completion (the rays stay where they are and
// the hits are shaded in place). kSplitItems rays per thread (more per thread measured no
// faster: the cursor atomics are not what bounds it).
constexpr int kSplitItems = 4;
__global__ __launch_bounds__(kAppendBlock) void k_split_index(
    const SceneDev s, HitSource src, int n_host, const int32_t* __restrict__ nd, int physics,
    float* __restrict__ accum, int32_t* __restrict__ hit_list, int32_t* __restrict__ miss_pixel,
    int32_t* __restrict__ miss_tag, int32_t* __restrict__ counts) {
    __shared__ uint32_t lds[kSplitItems * kAppendBlock / 64];
    const int n = load_count(n_host, nd);
    const int base = blockIdx.x * kAppendBlock * kSplitItems;
    if (base >= n) return;
    bool hit[kSplitItems], mis[kSplitItems];
    int pos[kSplitItems], m[kSplitItems];
#pragma unroll
    for (int j = 0; j < kSplitItems; ++j) {
        const int i = base + j * kAppendBlock + (int)threadIdx.x;
        const int L = i < n ? src.leaf[i] : -1;
        hit[j] = i < n && L >= 0;
        mis[j] = i < n && L < 0;
    }
    block_append_rounds<kSplitItems, kAppendBlock>(&counts[0], hit, pos, lds);
    block_append_rounds<kSplitItems, kAppendBlock>(&counts[1], mis, m, lds);
#pragma unroll
    for (int j = 0; j < kSplitItems; ++j) {
        const int i = base + j * kAppendBlock + (int)threadIdx.x;
        if (pos[j] >= 0) {
            hit_list[pos[j]] = i;
        } else if (m[j] >= 0) {
            const int slot = hs_slot(src, i);
            const float4 b = src.b[slot], c = src.c[slot];
            const int32_t pixel = reinterpret_cast<const int32_t*>(src.a)[4 * (size_t)slot + 3];
            add_miss(s, V3(b.x, b.y, b.z), V3(c.x, c.y, c.z), pixel, __float_as_int(c.w), b.w,
                     physics, accum);
            miss_pixel[m[j]] = pixel;
            if (miss_tag) miss_tag[m[j]] = src.tags[slot];
        }
    }
}

template <class R>
__global__ __launch_bounds__(kAppendBlock) void k_split_append(
    const SceneDev s, R in, int n_host, const int32_t* __restrict__ nd,
    const int32_t* __restrict__ leaf, const float* __restrict__ ht, const float* __restrict__ hu,
    const float* __restrict__ hv, wf_shade_soa out, wf_ray_soa miss, int32_t* __restrict__ counts,
    const int32_t* __restrict__ tags_in, int32_t* __restrict__ hit_tags,
    int32_t* __restrict__ miss_tags) {
    __shared__ uint32_t lds[kAppendItems * kAppendBlock / 64];
    const int n = load_count(n_host, nd);
    const int base = blockIdx.x * kAppendBlock * kAppendItems;
    if (base >= n) return;
    bool hit[kAppendItems], mis[kAppendItems];
    int L[kAppendItems], pos[kAppendItems], m[kAppendItems];
#pragma unroll
    for (int j = 0; j < kAppendItems; ++j) {
        const int i = base + j * kAppendBlock + (int)threadIdx.x;
        L[j] = i < n ? leaf[i] : -1;
        hit[j] = i < n && L[j] >= 0;
        mis[j] = i < n && L[j] < 0;
    }
    block_append_rounds<kAppendItems, kAppendBlock>(&counts[0], hit, pos, lds);
    block_append_rounds<kAppendItems, kAppendBlock>(&counts[1], mis, m, lds);
#pragma unroll
    for (int j = 0; j < kAppendItems; ++j) {
        const int i = base + j * kAppendBlock + (int)threadIdx.x;
        if (pos[j] >= 0) {
            map_hit(s, in, i, L[j], ht[i], hu[i], hv[i], out, pos[j]);
            if (hit_tags) hit_tags[pos[j]] = tags_in[ray_source(in, i)];
        } else if (m[j] >= 0) {
            st_ray(miss, m[j], ld_ray(in, i));
            if (miss_tags) miss_tags[m[j]] = tags_in[ray_source(in, i)];
        }
    }
}

// Split with the hits appended into the ranges of their group keys laid out by policy_buckets
// (the shading order of a policy without a sort). Each block counts its hits per range in LDS,
// claims one run of slots per range it holds with a single atomic on that range's cursor, and
// lays its hits out in LDS in range order; thread k then maps the k-th hit of that order, so
// that consecutive lanes write consecutive slots of one run (coalesced stores) and only the
// reads of a hit's ray and triangle follow the permutation. Misses are appended as in
// k_split_append. INDEX: each hit's traced-ray index goes into its range of hit_list, or, with
// records (rec.a set), its packed ray, closest hit and tag are copied into its range of the
// records (read in ray order, written in runs); the misses are resolved as in k_split_index
// (index mode: no shading-queue record is mapped). Dynamic LDS:
// 3 n_materials words (there are at most n_materials ranges).
constexpr int kBucketItems = 8;
constexpr int kBucketTile = kAppendBlock * kBucketItems;
template <class R, bool INDEX>
__global__ __launch_bounds__(kAppendBlock) void k_split_buckets(
    const SceneDev s, R in, int n_host, const int32_t* __restrict__ nd,
    const int32_t* __restrict__ leaf, const float* __restrict__ ht, const float* __restrict__ hu,
    const float* __restrict__ hv, wf_shade_soa out, wf_ray_soa miss, int32_t* __restrict__ counts,
    const int32_t* __restrict__ bucket, int32_t* __restrict__ cursor,
    const int32_t* __restrict__ tags_in, int32_t* __restrict__ hit_tags,
    int32_t* __restrict__ miss_tags, int32_t* __restrict__ hit_list, HitRecords rec,
    HitSource src, int physics, float* __restrict__ accum) {
    extern __shared__ uint32_t sb_lds[];
    __shared__ uint32_t lds[kBucketItems * kAppendBlock / 64];
    __shared__ int32_t order[kBucketTile];   // ray of the k-th hit of the block, in range order
    __shared__ uint16_t order_range[kBucketTile];   // and its range
    __shared__ uint32_t s_hits;
    const int n = load_count(n_host, nd);
    const int base = blockIdx.x * kBucketTile;
    if (base >= n) return;
    const int nm = s.n_materials;
    uint32_t* cnt = sb_lds;              // hits per range in this block
    uint32_t* local = sb_lds + nm;       // first position of each range in `order`
    uint32_t* run = sb_lds + 2 * nm;     // first slot of this block's run of each range
    for (int k = threadIdx.x; k < nm; k += kAppendBlock) cnt[k] = 0u;
    __syncthreads();
    bool mis[kBucketItems];
    int L[kBucketItems], mt[kBucketItems], m[kBucketItems];
    uint32_t rk[kBucketItems];
#pragma unroll
    for (int j = 0; j < kBucketItems; ++j) {
        const int i = base + j * kAppendBlock + (int)threadIdx.x;
        L[j] = i < n ? leaf[i] : -1;
        mis[j] = i < n && L[j] < 0;
        mt[j] = L[j] >= 0 ? bucket[s.leaf_mat[L[j]]] : 0;
        rk[j] = L[j] >= 0 ? atomicAdd(&cnt[mt[j]], 1u) : 0u;
    }
    __syncthreads();
    // exclusive scan of the block's counts in range order (wave 0), and the runs
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        uint32_t carry = 0;
        for (int c0 = 0; c0 < nm; c0 += 64) {
            const int k = c0 + lane;
            const uint32_t v = k < nm ? cnt[k] : 0u;
            uint32_t inc = v;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(inc, o, 64);
                if (lane >= o) inc += y;
            }
            if (k < nm) local[k] = carry + inc - v;
            carry += __shfl(inc, 63, 64);
        }
        if (lane == 0) s_hits = carry;
    }
    for (int k = threadIdx.x; k < nm; k += kAppendBlock)
        if (cnt[k]) run[k] = (uint32_t)atomicAdd(&cursor[k], (int32_t)cnt[k]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBucketItems; ++j)
        if (L[j] >= 0) {
            const uint32_t k = local[mt[j]] + rk[j];
            order[k] = base + j * kAppendBlock + (int)threadIdx.x;
            order_range[k] = (uint16_t)mt[j];
        }
    // (its barriers also publish `order`)
    block_append_rounds<kBucketItems, kAppendBlock>(&counts[1], mis, m, lds);
#pragma unroll
    for (int j = 0; j < kBucketItems; ++j) {
        if (m[j] < 0) continue;
        const int i = base + j * kAppendBlock + (int)threadIdx.x;
        if (INDEX) {
            // resolved here, as in k_split_index
            const int slot = hs_slot(src, i);
            const float4 b = src.b[slot], c = src.c[slot];
            const int32_t pixel = reinterpret_cast<const int32_t*>(src.a)[4 * (size_t)slot + 3];
            add_miss(s, V3(b.x, b.y, b.z), V3(c.x, c.y, c.z), pixel, __float_as_int(c.w), b.w,
                     physics, accum);
            miss.pixel[m[j]] = pixel;
            if (miss_tags) miss_tags[m[j]] = src.tags[slot];
            continue;
        }
        st_ray(miss, m[j], ld_ray(in, i));
        if (miss_tags) miss_tags[m[j]] = tags_in[ray_source(in, i)];
    }
    const int hits = (int)s_hits;
    for (int k = threadIdx.x; k < hits; k += kAppendBlock) {
        const int i = order[k];
        const int mi = order_range[k];
        const int pos = (int)(run[mi] + ((uint32_t)k - local[mi]));
        if (INDEX && !rec.a) {
            hit_list[pos] = i;
            continue;
        }
        if (INDEX) {
            const int slot = hs_slot(src, i);
            rec.a[pos] = src.a[slot];
            rec.b[pos] = src.b[slot];
            rec.c[pos] = src.c[slot];
            rec.leaf[pos] = src.leaf[i];
            rec.t[pos] = src.t[i];
            rec.u[pos] = src.u[i];
            rec.v[pos] = src.v[i];
            if (rec.tag) rec.tag[pos] = src.tags[slot];
            continue;
        }
        map_hit(s, in, i, leaf[i], ht[i], hu[i], hv[i], out, pos);
        if (hit_tags) hit_tags[pos] = tags_in[ray_source(in, i)];
    }
}

// One HitRecord per input ray (slot i), hit flag per ray.
__global__ void k_records(const SceneDev s, wf_ray_soa in, int n, const int32_t* __restrict__ leaf,
                          const float* __restrict__ ht, const float* __restrict__ hu,
                          const float* __restrict__ hv, wf_shade_soa out,
                          uint8_t* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int L = leaf[i];
    flags[i] = L >= 0 ? 1 : 0;
    if (L >= 0) map_hit(s, in, i, L, ht[i], hu[i], hv[i], out, i);
}

__global__ void k_hits_to_soa(const SceneDev s, int n, const int32_t* __restrict__ leaf,
                              const float* __restrict__ ht, const float* __restrict__ hu,
                              const float* __restrict__ hv, wf_hit_soa out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int L = leaf[i];
    if (L >= 0) {
        const int32_t g = __float_as_int(s.tri[3 * (size_t)L].w);
        const int32_t geom = __float_as_int(s.tri[3 * (size_t)L + 1].w);
        out.geom[i] = geom;
        out.prim[i] = g - s.mesh_tri_offset[geom];
    } else {
        out.geom[i] = -1;
        out.prim[i] = -1;
    }
    out.t[i] = ht[i];
    out.bu[i] = hu[i];
    out.bv[i] = hv[i];
}

wf_status scene_ready(const wf_scene* s) {
    if (!s) return set_error(WF_ERR_INVALID_ARGUMENT, "NULL scene");
    if (!s->dev.built) return set_error(WF_ERR_NOT_BUILT, "accelerator not built (call wf_accel_build)");
    if (s->dev.stack_bound > kStackMax)
        return set_error(WF_ERR_UNSUPPORTED,
                         "BVH needs a traversal stack of %d entries (at most %d supported)",
                         s->dev.stack_bound, kStackMax);
    return WF_OK;
}

template <class R>
static wf_status trace_closest_any(wf_context* ctx, const wf_scene* s, const R& r, Count n,
                                   float tnear, float tfar, int32_t* hit_leaf, float* hit_t,
                                   float* hit_bu, float* hit_bv, uint32_t* flags) {
    WF_TRY(scene_ready(s));
    if (n.n <= 0) return WF_OK;
    void* pc = nullptr;
    WF_TRY(workspace(ctx, "trace_counter", sizeof(int32_t) * kWorkInts, &pc));
    WF_TRY(wf_fill_u32(ctx, pc, 0u, kWorkInts));
    Timed tm(ctx, "traverse_closest", 44.0, n);
    with_bvh(s->dev, [&](auto nodes, auto S) {
        k_trace_closest<std::remove_cv_t<std::remove_pointer_t<decltype(nodes)>>,
                        decltype(S)::value, R>
            <<<persistent_grid(n.n, ctx), kTraceBlock, 0, ctx->stream>>>(
                nodes, s->dev.tri, s->dev.n_tris, r, n.n, n.dev, tnear, tfar, hit_leaf, hit_t,
                hit_bu, hit_bv, flags, (int32_t*)pc, ctx->tuning.refill_lanes);
    });
    WF_HIP(hipGetLastError());
    return WF_OK;
}

wf_status trace_closest(wf_context* ctx, const wf_scene* s, const wf_ray_soa* r, Count n,
                        float tnear, float tfar, int32_t* hit_leaf, float* hit_t, float* hit_bu,
                        float* hit_bv, uint32_t* flags) {
    return trace_closest_any(ctx, s, *r, n, tnear, tfar, hit_leaf, hit_t, hit_bu, hit_bv, flags);
}

wf_status trace_closest(wf_context* ctx, const wf_scene* s, const RayPacked* r, Count n,
                        float tnear, float tfar, int32_t* hit_leaf, float* hit_t, float* hit_bu,
                        float* hit_bv, uint32_t* flags) {
    return trace_closest_any(ctx, s, *r, n, tnear, tfar, hit_leaf, hit_t, hit_bu, hit_bv, flags);
}

wf_status trace_closest(wf_context* ctx, const wf_scene* s, const RayPackedOrdered* r, Count n,
                        float tnear, float tfar, int32_t* hit_leaf, float* hit_t, float* hit_bu,
                        float* hit_bv, uint32_t* flags) {
    return trace_closest_any(ctx, s, *r, n, tnear, tfar, hit_leaf, hit_t, hit_bu, hit_bv, flags);
}

wf_status trace_occluded(const wf_scene* s, const float* ox, const float* oy, const float* oz,
                         const float* dx, const float* dy, const float* dz, const float* maxd,
                         int32_t n, uint8_t* out) {
    WF_TRY(scene_ready(s));
    if (n <= 0) return WF_OK;
    wf_context* ctx = s->ctx;
    Timed tm(ctx, "traverse_any", 29.0 * n);
    with_bvh(s->dev, [&](auto nodes, auto S) {
        k_trace_occluded<std::remove_cv_t<std::remove_pointer_t<decltype(nodes)>>,
                         decltype(S)::value>
            <<<grid_for(n, kTraceBlock), kTraceBlock, 0, ctx->stream>>>(
                nodes, s->dev.tri, s->dev.n_tris, ox, oy, oz, dx, dy, dz, maxd, n, out);
    });
    WF_HIP(hipGetLastError());
    return WF_OK;
}

wf_status hit_buffers(wf_context* ctx, Count n, int32_t** leaf, float** t, float** u, float** v) {
    const size_t nn = (size_t)(n.n > 0 ? n.n : 1);
    void *pl, *pt, *pu, *pv;
    WF_TRY(workspace(ctx, "hit_leaf", sizeof(int32_t) * nn, &pl));
    WF_TRY(workspace(ctx, "hit_t", sizeof(float) * nn, &pt));
    WF_TRY(workspace(ctx, "hit_u", sizeof(float) * nn, &pu));
    WF_TRY(workspace(ctx, "hit_v", sizeof(float) * nn, &pv));
    *leaf = (int32_t*)pl;
    *t = (float*)pt;
    *u = (float*)pu;
    *v = (float*)pv;
    return WF_OK;
}

// Hits and misses of traced rays appended at device cursors (queue order free). With ranges
// (from policy_buckets, which has also written the hit count to counts_dev[0]) the hits go into
// the ranges of their group keys instead.
template <class R>
static wf_status split_append(wf_context* ctx, const wf_scene* s, const R& rays, Count n,
                              const int32_t* hl, const float* ht, const float* hu,
                              const float* hv, wf_shade_soa* hits_out, wf_ray_soa* misses_out,
                              int32_t* counts_dev, const int32_t* tags_in, int32_t* hit_tags,
                              int32_t* miss_tags, const PolicyRanges* ranges = nullptr) {
    if (ranges) {
        WF_TRY(wf_fill_u32(ctx, counts_dev + 1, 0u, 1));
        if (n.n <= 0) return WF_OK;
        Timed tm(ctx, "split", 97.0, n);
        const size_t lds = sizeof(uint32_t) * 3 * (size_t)s->dev.n_materials;
        k_split_buckets<R, false><<<grid_for(n.n, kBucketTile), kAppendBlock, lds,
                                    ctx->stream>>>(
            s->dev, rays, n.n, n.dev, hl, ht, hu, hv, *hits_out, *misses_out, counts_dev,
            ranges->bucket, ranges->cursor, tags_in, hit_tags, miss_tags, nullptr, HitRecords{},
            HitSource{}, 0, nullptr);
        WF_HIP(hipGetLastError());
        return WF_OK;
    }
    WF_TRY(wf_fill_u32(ctx, counts_dev, 0u, 2));
    if (n.n > 0) {
        Timed tm(ctx, "split", 97.0, n);
        k_split_append<R><<<grid_for(n.n, kAppendBlock * kAppendItems), kAppendBlock, 0,
                            ctx->stream>>>(
            s->dev, rays, n.n, n.dev, hl, ht, hu, hv, *hits_out, *misses_out, counts_dev,
            tags_in, hit_tags, miss_tags);
        WF_HIP(hipGetLastError());
    }
    return WF_OK;
}

wf_status split_traced(wf_context* ctx, const wf_scene* s, const RayPacked* rays, Count n,
                       const int32_t* hit_leaf, const float* hit_t, const float* hit_bu,
                       const float* hit_bv, wf_shade_soa* hits_out, wf_ray_soa* misses_out,
                       int32_t* counts_dev, const int32_t* tags_in, int32_t* hit_tags,
                       int32_t* miss_tags, const PolicyRanges* ranges) {
    return split_append(ctx, s, *rays, n, hit_leaf, hit_t, hit_bu, hit_bv, hits_out, misses_out,
                        counts_dev, tags_in, hit_tags, miss_tags, ranges);
}

wf_status split_traced(wf_context* ctx, const wf_scene* s, const RayPackedOrdered* rays, Count n,
                       const int32_t* hit_leaf, const float* hit_t, const float* hit_bu,
                       const float* hit_bv, wf_shade_soa* hits_out, wf_ray_soa* misses_out,
                       int32_t* counts_dev, const int32_t* tags_in, int32_t* hit_tags,
                       int32_t* miss_tags, const PolicyRanges* ranges) {
    return split_append(ctx, s, *rays, n, hit_leaf, hit_t, hit_bu, hit_bv, hits_out, misses_out,
                        counts_dev, tags_in, hit_tags, miss_tags, ranges);
}

wf_status split_index(wf_context* ctx, const wf_scene* s, const HitSource& src, Count n,
                      int32_t* hit_list, int32_t physics, float* accum, int32_t* miss_pixel,
                      int32_t* miss_tag, int32_t* counts_dev, const PolicyRanges* ranges,
                      const HitRecords* records) {
    // per ray: leaf 4 read, hit index 4 written; per miss also direction/pdf and
    // throughput/depth 32, accumulator index 4, tag 4 read, accumulator 12 read and written,
    // completion columns 8 written (the misses are about half of the rays)
    const double bytes = 8.0 + 0.5 * 72.0;
    if (ranges) {
        WF_TRY(wf_fill_u32(ctx, counts_dev + 1, 0u, 1));
        if (n.n <= 0) return WF_OK;
        // + the leaf's material, and with records per hit (about half) the 68-byte record read
        // and written
        Timed tm(ctx, "split", bytes + 4.0 + (records ? 0.5 * 2.0 * 68.0 : 0.0), n);
        const size_t lds = sizeof(uint32_t) * 3 * (size_t)s->dev.n_materials;
        wf_ray_soa miss{};
        miss.pixel = miss_pixel;
        k_split_buckets<RayPacked, true><<<grid_for(n.n, kBucketTile), kAppendBlock, lds,
                                           ctx->stream>>>(
            s->dev, RayPacked{}, n.n, n.dev, src.leaf, nullptr, nullptr, nullptr, wf_shade_soa{},
            miss, counts_dev, ranges->bucket, ranges->cursor, nullptr, nullptr, miss_tag,
            hit_list, records ? *records : HitRecords{}, src, physics, accum);
        WF_HIP(hipGetLastError());
        return WF_OK;
    }
    WF_TRY(wf_fill_u32(ctx, counts_dev, 0u, 2));
    if (n.n > 0) {
        Timed tm(ctx, "split", bytes, n);
        k_split_index<<<grid_for(n.n, kAppendBlock * kSplitItems), kAppendBlock, 0,
                        ctx->stream>>>(s->dev, src, n.n, n.dev, physics, accum, hit_list,
                                       miss_pixel, miss_tag, counts_dev);
        WF_HIP(hipGetLastError());
    }
    return WF_OK;
}

// HitRecord mapping of listed hits into a shading queue (the engine's diagnostic).
__global__ void k_gather_hits(const SceneDev s, HitSource h, int n_host,
                              const int32_t* __restrict__ nd, wf_shade_soa out,
                              int32_t* __restrict__ tags_out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= load_count(n_host, nd)) return;
    const int i = hs_traced(h, j);
    const int slot = hs_slot(h, i);
    const RayPacked r{const_cast<float4*>(h.a), const_cast<float4*>(h.b), const_cast<float4*>(h.c)};
    map_hit(s, r, slot, h.leaf[i], h.t[i], h.u[i], h.v[i], out, j);
    if (tags_out && h.tags) tags_out[j] = h.tags[slot];
}

wf_status gather_hits(wf_context* ctx, const wf_scene* s, const HitSource& h, Count n,
                      wf_shade_soa* dst, int32_t* tags_out) {
    if (n.n <= 0) return WF_OK;
    k_gather_hits<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(s->dev, h, n.n, n.dev, *dst,
                                                                tags_out);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

// Closest hit, then hits and misses appended at device cursors (queue order free).
template <class R>
static wf_status split_unordered(wf_context* ctx, const wf_scene* s, const R& rays, Count n,
                                 wf_shade_soa* hits_out, wf_ray_soa* misses_out,
                                 int32_t* counts_dev, const int32_t* tags_in, int32_t* hit_tags,
                                 int32_t* miss_tags) {
    int32_t* hl;
    float *ht, *hu, *hv;
    WF_TRY(hit_buffers(ctx, n, &hl, &ht, &hu, &hv));
    WF_TRY(trace_closest_any(ctx, s, rays, n, 0.001f, 1e9f, hl, ht, hu, hv, nullptr));
    return split_append(ctx, s, rays, n, hl, ht, hu, hv, hits_out, misses_out, counts_dev,
                        tags_in, hit_tags, miss_tags);
}

// Closest hit + stable split; counts_dev[0] = hits, counts_dev[1] = misses (device int32).
wf_status intersect_split(wf_context* ctx, const wf_scene* s, const wf_ray_soa* rays, Count n,
                          wf_shade_soa* hits_out, wf_ray_soa* misses_out, int32_t* counts_dev,
                          const int32_t* tags_in, int32_t* hit_tags, int32_t* miss_tags,
                          bool unordered) {
    void *pl, *pt, *pu, *pv, *pf;
    size_t nn = (size_t)(n.n > 0 ? n.n : 1);
    WF_TRY(workspace(ctx, "hit_leaf", sizeof(int32_t) * nn, &pl));
    WF_TRY(workspace(ctx, "hit_t", sizeof(float) * nn, &pt));
    WF_TRY(workspace(ctx, "hit_u", sizeof(float) * nn, &pu));
    WF_TRY(workspace(ctx, "hit_v", sizeof(float) * nn, &pv));
    WF_TRY(workspace(ctx, "hit_flag", sizeof(uint32_t) * nn, &pf));
    if (unordered)
        return split_unordered(ctx, s, *rays, n, hits_out, misses_out, counts_dev, tags_in,
                               hit_tags, miss_tags);
    WF_TRY(trace_closest(ctx, s, rays, n, 0.001f, 1e9f, (int32_t*)pl, (float*)pt, (float*)pu,
                         (float*)pv, (uint32_t*)pf));
    WF_TRY(exclusive_scan_u32(ctx, (uint32_t*)pf, (uint32_t*)pf, n, counts_dev, "split"));
    if (n.n > 0) {
        Timed tm(ctx, "split", 97.0, n);
        k_split<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(
            s->dev, *rays, n.n, n.dev, (int32_t*)pl, (float*)pt, (float*)pu, (float*)pv,
            (uint32_t*)pf, *hits_out, *misses_out, tags_in, hit_tags);
        WF_HIP(hipGetLastError());
    }
    if (n.dev) {
        WF_TRY(sub_count(ctx, counts_dev + 1, n.dev, counts_dev));
    } else {
        // misses = n - hits, on the device so that callers read both counts from one place
        void* pn = nullptr;
        WF_TRY(workspace(ctx, "split_n", sizeof(int32_t), &pn));
        WF_TRY(wf_fill_u32(ctx, pn, (uint32_t)n.n, 1));
        WF_TRY(sub_count(ctx, counts_dev + 1, (const int32_t*)pn, counts_dev));
    }
    return WF_OK;
}

}  // namespace wf

using namespace wf;

extern "C" {

wf_status wf_intersect(const wf_scene* s, const wf_ray_soa* rays, int32_t n, float tnear,
                       float tfar, wf_hit_soa* hits) {
    clear_error();
    WF_REQUIRE(s && rays && hits, "wf_intersect: NULL argument");
    WF_REQUIRE(n >= 0, "wf_intersect: negative ray count");
    WF_TRY(scene_ready(s));
    if (n == 0) return WF_OK;
    wf_context* ctx = s->ctx;
    void *pl, *pt, *pu, *pv;
    WF_TRY(workspace(ctx, "hit_leaf", sizeof(int32_t) * n, &pl));
    WF_TRY(workspace(ctx, "hit_t", sizeof(float) * n, &pt));
    WF_TRY(workspace(ctx, "hit_u", sizeof(float) * n, &pu));
    WF_TRY(workspace(ctx, "hit_v", sizeof(float) * n, &pv));
    WF_TRY(trace_closest(ctx, s, rays, host_count(n), tnear, tfar, (int32_t*)pl, (float*)pt,
                         (float*)pu, (float*)pv, nullptr));
    k_hits_to_soa<<<grid_for(n, 256), 256, 0, ctx->stream>>>(s->dev, n, (int32_t*)pl, (float*)pt,
                                                             (float*)pu, (float*)pv, *hits);
    WF_HIP(hipGetLastError());
    WF_HIP(hipStreamSynchronize(ctx->stream));
    return WF_OK;
}

wf_status wf_intersect_records(const wf_scene* s, const wf_ray_soa* rays, int32_t n, float tnear,
                               float tfar, wf_shade_soa* records, uint8_t* hit_flags) {
    clear_error();
    WF_REQUIRE(s && rays && records && hit_flags, "wf_intersect_records: NULL argument");
    WF_REQUIRE(n >= 0, "wf_intersect_records: negative ray count");
    WF_TRY(scene_ready(s));
    if (n == 0) return WF_OK;
    wf_context* ctx = s->ctx;
    void *pl, *pt, *pu, *pv;
    WF_TRY(workspace(ctx, "hit_leaf", sizeof(int32_t) * n, &pl));
    WF_TRY(workspace(ctx, "hit_t", sizeof(float) * n, &pt));
    WF_TRY(workspace(ctx, "hit_u", sizeof(float) * n, &pu));
    WF_TRY(workspace(ctx, "hit_v", sizeof(float) * n, &pv));
    WF_TRY(trace_closest(ctx, s, rays, host_count(n), tnear, tfar, (int32_t*)pl, (float*)pt,
                         (float*)pu, (float*)pv, nullptr));
    k_records<<<grid_for(n, 256), 256, 0, ctx->stream>>>(s->dev, *rays, n, (int32_t*)pl,
                                                         (float*)pt, (float*)pu, (float*)pv,
                                                         *records, hit_flags);
    WF_HIP(hipGetLastError());
    WF_HIP(hipStreamSynchronize(ctx->stream));
    return WF_OK;
}

wf_status wf_occluded(const wf_scene* s, const float* ox, const float* oy, const float* oz,
                      const float* dx, const float* dy, const float* dz, const float* max_dist,
                      int32_t n, uint8_t* occluded) {
    clear_error();
    WF_REQUIRE(s && ox && oy && oz && dx && dy && dz && max_dist && occluded,
               "wf_occluded: NULL argument");
    WF_REQUIRE(n >= 0, "wf_occluded: negative ray count");
    WF_TRY(trace_occluded(s, ox, oy, oz, dx, dy, dz, max_dist, n, occluded));
    WF_HIP(hipStreamSynchronize(s->ctx->stream));
    return WF_OK;
}

wf_status wf_trace_stats(const wf_scene* s, const wf_ray_soa* rays, int32_t n, float tnear,
                         float tfar, int32_t any_hit, int32_t* nodes_visited,
                         int32_t* triangles_tested) {
    clear_error();
    WF_REQUIRE(s && rays && nodes_visited && triangles_tested, "wf_trace_stats: NULL argument");
    WF_REQUIRE(n >= 0, "wf_trace_stats: negative ray count");
    WF_TRY(scene_ready(s));
    if (n == 0) return WF_OK;
    wf_context* ctx = s->ctx;
    with_bvh(s->dev, [&](auto nodes, auto S) {
        using N = std::remove_cv_t<std::remove_pointer_t<decltype(nodes)>>;
        if (any_hit)
            k_trace_stats<N, decltype(S)::value, true>
                <<<grid_for(n, kTraceBlock), kTraceBlock, 0, ctx->stream>>>(
                    nodes, s->dev.tri, s->dev.n_tris, *rays, n, tnear, tfar, nodes_visited,
                    triangles_tested);
        else
            k_trace_stats<N, decltype(S)::value, false>
                <<<grid_for(n, kTraceBlock), kTraceBlock, 0, ctx->stream>>>(
                    nodes, s->dev.tri, s->dev.n_tris, *rays, n, tnear, tfar, nodes_visited,
                    triangles_tested);
    });
    WF_HIP(hipGetLastError());
    WF_HIP(hipStreamSynchronize(ctx->stream));
    return WF_OK;
}

wf_status wf_intersect_split(const wf_scene* s, const wf_ray_soa* rays, int32_t n,
                             wf_shade_soa* hits_out, int32_t* n_hits, wf_ray_soa* misses_out,
                             int32_t* n_misses) {
    clear_error();
    WF_REQUIRE(s && rays && hits_out && misses_out && n_hits && n_misses,
               "wf_intersect_split: NULL argument");
    WF_REQUIRE(n >= 0, "wf_intersect_split: negative ray count");
    WF_TRY(scene_ready(s));
    void* pc = nullptr;
    WF_TRY(workspace(s->ctx, "split_count", sizeof(int32_t) * 2, &pc));
    WF_TRY(intersect_split(s->ctx, s, rays, host_count(n), hits_out, misses_out, (int32_t*)pc));
    int32_t h = 0;
    WF_TRY(readback(s->ctx, &h, pc, sizeof(int32_t)));
    *n_hits = h;
    *n_misses = n - h;
    return WF_OK;
}

}  // extern "C"
