// Host-side internals shared by the HIP translation units of libwavefront_hip.so: the context
// (stream, grow-only workspace pool, event timing), the scene's device layout, and the internal
// launch entry points each .hip file provides.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/wavefront_hip.h"

namespace wf {

// ---- errors ----------------------------------------------------------------------------------
wf_status set_error(wf_status code, const char* fmt, ...);
void clear_error();

#define WF_HIP(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ::wf::set_error(WF_ERR_HIP, "%s failed: %s (%s:%d)", #expr,          \
                                   hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

#define WF_TRY(expr)                      \
    do {                                  \
        wf_status s_ = (expr);            \
        if (s_ != WF_OK) return s_;       \
    } while (0)

#define WF_REQUIRE(cond, ...)                                                 \
    do {                                                                      \
        if (!(cond)) return ::wf::set_error(WF_ERR_INVALID_ARGUMENT, __VA_ARGS__); \
    } while (0)

// ---- context ---------------------------------------------------------------------------------
struct ProfEntry {
    double ms = 0.0;
    int64_t launches = 0;
    double bytes = 0.0;
};

struct PendingTiming {
    std::string name;
    hipEvent_t start, stop;
    double bytes;              // algorithmic bytes, or bytes per unit when count_slot >= 0
    int count_slot;            // pinned slot receiving the launch's device-side element count
    double bytes2 = 0.0;       // a second kind of unit: fixed bytes, or per unit of count_slot2
    int count_slot2 = -1;
};

}  // namespace wf

struct wf_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // grow-only named scratch buffers (freed with the context)
    struct Buf {
        void* ptr = nullptr;
        size_t bytes = 0;
    };
    std::map<std::string, Buf> workspace;
    // pinned host staging for small device->host readbacks
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // timing
    bool profile = false;
    std::string profile_only;          // non-empty: time only this kernel name
    std::map<std::string, wf::ProfEntry> prof;
    std::vector<wf::PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    int32_t* count_slots = nullptr;    // pinned, kCountSlots entries
    wf_tuning tuning{48, 12, WF_BVH_PLOC, 16, 4, 8, 0, 0, 33};
    int next_count_slot = 0;
    int cu_count = 256;
    // Device word that kernels set when a bounded device-wide wait (the radix sort's look-back)
    // runs out; checked at the next synchronisation (check_fault). Owned by the context, or a
    // slot of an engine's status block (copied back with every iteration's status).
    int32_t* fault = nullptr;
    bool fault_owned = false;
    // This device's copy of the allocation-free ray schedule's scratch array (its address as
    // hipGetSymbolAddress returns it on this device).
    void* fixed_scratch = nullptr;
};

namespace wf {

// Named scratch buffer of at least `bytes` on the context's device (grow-only).
wf_status workspace(wf_context* ctx, const char* name, size_t bytes, void** out);
// Pinned host scratch of at least `bytes`.
wf_status pinned(wf_context* ctx, size_t bytes, void** out);
// Copies `bytes` from device to host through pinned memory and synchronises the stream.
wf_status readback(wf_context* ctx, void* host, const void* dev, size_t bytes);
// Synchronises the stream and reads the context's fault word: WF_ERR_HIP (and the word cleared)
// when a kernel reported an exhausted wait since the last check.
wf_status check_fault(wf_context* ctx);
// The error of a set fault word.
wf_status fault_error();

struct Count;
constexpr int kCountSlots = 16384;

// Brackets a kernel launch with events when profiling is on. With a device-side element count
// the algorithmic bytes are bytes_per_unit times the count, copied back after the launch.
struct Timed {
    wf_context* ctx;
    const char* name;
    double bytes;
    const int32_t* count_dev = nullptr;
    double bytes2 = 0.0;
    const int32_t* count_dev2 = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(wf_context* c, const char* n, double algorithmic_bytes);
    Timed(wf_context* c, const char* n, double bytes_per_unit, const Count& count);
    // a launch over two kinds of units (per-unit bytes and count of each)
    Timed(wf_context* c, const char* n, double bytes_per_unit, const Count& count,
          double bytes_per_unit2, const Count& count2);
    ~Timed();
};
wf_status harvest_timings(wf_context* ctx);

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// ---- scene layout on the device ---------------------------------------------------------------
struct DevMaterial {   // wf_material packed for the device
    int32_t type;
    int32_t model;
    int32_t texture_id;
    int32_t use_spatial_checker;
    float albedo[3];
    float roughness;
    float emission;
    float spot_outer_deg;
    float spot_falloff_deg;
    float ior;
    float checker_cell_size;
    float light_pdf_area;      // unbiased track: selection pdf per unit area of this emitter
    float pad[2];
};

struct DevTexture {
    int32_t width, height, format, pad;
    int64_t offset;            // texel offset (RGBA8: u32 units; float RGB: float4 units)
};

// Four-wide node (128 B): child boxes per axis in SoA form, child index >= 0 = wide node,
// < 0 and != kEmptyChild = ~leaf triangle.
struct Bvh4Node {   // an empty child slot has lo = +inf, hi = -inf (never entered)
    float4 lox, loy, loz;
    float4 hix, hiy, hiz;
    int4 child;
    int4 pad;
};
constexpr int32_t kEmptyChild = (int32_t)0x80000000;

// Eight-wide node (128 B) with child boxes quantised to 8 bits per bound on a per-node grid:
// bound = origin + q * 2^(e - 127) per axis (origin in `origin.xyz`, the three biased exponents
// in the bytes of `origin.w`). Quantisation rounds outward, checked with the decode's own float
// operations, so every decoded box contains its child.
struct Bvh8Node {
    float4 origin;
    int4 child_lo;     // children 0..3 (>= 0 node, < 0 ~leaf, kEmptyChild empty)
    int4 child_hi;     // children 4..7
    uint4 qa;          // lo.x bytes of children 0..7, then lo.y
    uint4 qb;          // lo.z, then hi.x
    uint4 qc;          // hi.y, then hi.z
    int4 pad[2];
};

struct SceneDev {
    int32_t n_tris = 0, n_meshes = 0, n_materials = 0, n_textures = 0, n_vertices = 0;
    // triangles in BVH leaf order: v0 (w = global index bits), e1 (w = mesh index bits), e2
    float4* tri = nullptr;
    // their vertex UVs in leaf order, two records per leaf: (u0, v0, u1, v1), (u2, v2, 0, 0), so
    // that a hit reads its shading data next to its triangle
    float4* tri_uv_leaf = nullptr;
    int32_t* leaf_mat = nullptr;       // material of each leaf triangle (shading-order buckets)
    // per global triangle
    int32_t* tri_geom = nullptr;
    float* tri_uv = nullptr;           // 6 per triangle (zeros without UVs)
    float4* tri_v0 = nullptr;          // unsorted v0/e1/e2 for light sampling (global index)
    float4* tri_e1 = nullptr;
    float4* tri_e2 = nullptr;
    // per mesh
    int32_t* mesh_tri_offset = nullptr;
    int32_t* mesh_material = nullptr;
    int32_t* mesh_has_uv = nullptr;
    // materials / textures
    DevMaterial* materials = nullptr;
    DevTexture* textures = nullptr;
    uint32_t* texels8 = nullptr;
    float4* texels32 = nullptr;
    float* srgb_lut = nullptr;         // 256 entries
    uint64_t* tex_mem_bytes = nullptr; // Texture::memoryBytes per texture (w*h*12)
    // policy tables
    int32_t* mat_cost_rank = nullptr;  // dense rank of relativeCost (descending)
    int32_t cost_rank_bits = 0;
    double* mat_rel_cost = nullptr;
    double* tex_rel_cost = nullptr;
    int32_t* tex_size_rank = nullptr;  // dense rank of memoryBytes (descending), n_textures+1 entries
    int32_t size_rank_bits = 0;
    // lights
    int32_t n_emissive_meshes = 0;
    int32_t* emissive_meshes = nullptr;
    int32_t* mesh_tri_count = nullptr;
    float total_emissive_area = 0.0f;
    int32_t n_light_tris = 0;
    int32_t* light_tris = nullptr;         // global triangle id per emitter slot
    float4* light_v0 = nullptr;            // original vertices per emitter slot
    float4* light_v1 = nullptr;
    float4* light_v2 = nullptr;
    int32_t* light_mat = nullptr;          // material per emitter slot
    int32_t* emissive_light_base = nullptr;  // first emitter slot of each emissive mesh
    float* light_prob = nullptr;
    int32_t* light_alias = nullptr;
    wf_directional_light directional{};
    int32_t env_w = 0, env_h = 0;
    float4* env = nullptr;             // raw RGB (w unused)
    float* env_prob = nullptr;
    int32_t* env_alias = nullptr;
    float* env_pdf = nullptr;
    float env_intensity = 1.0f;
    float env_tint[3] = {1.0f, 1.0f, 1.0f};
    float env_choice_prob = 0.0f;
    // BVH (two-child LBVH collapsed to four-wide nodes)
    Bvh4Node* nodes4 = nullptr;        // four-wide nodes (width 4)
    Bvh8Node* nodes8 = nullptr;        // eight-wide quantised nodes (width 8)
    int32_t width = 4;
    int32_t n_nodes = 0;               // two-child internal nodes of the build
    int32_t n_nodes4 = 0;
    int32_t max_depth = 0;             // two-child depth (deepest leaf's ancestor count)
    int32_t stack_bound = 0;           // max entries the four-wide traversal stack can hold
    bool built = false;
    float bounds_min[3] = {0, 0, 0}, bounds_max[3] = {0, 0, 0};
};

}  // namespace wf

// CostTracker (include/render/CostTracker.h:8-63), restated on the host.
struct wf_cost_tracker {
    std::vector<double> average;       // averageCostNanoseconds
    std::vector<uint8_t> initialized;
    std::vector<int32_t> samples;      // sampleCount
    double total = 0.0;                // cachedTotalCost
    int32_t tracked = 0;               // trackedItemCount
};

namespace wf {
// Device sums of timed shading (wall-clock ticks) and their counts, per material and texture.
struct CostProbe {
    unsigned long long* mat_sum = nullptr;
    uint32_t* mat_cnt = nullptr;
    unsigned long long* tex_sum = nullptr;
    uint32_t* tex_cnt = nullptr;
};
}  // namespace wf

struct wf_scene {
    wf_context* ctx = nullptr;
    int device = 0;            // kept so that destroy does not depend on the context
    wf::SceneDev dev;
    std::vector<void*> allocations;
    size_t device_bytes = 0;
    // host copies needed for table rebuilds
    std::vector<wf_material> materials;
    std::vector<uint64_t> tex_mem;
    // engine reused by wf_render's regenerating schedule
    wf_engine* render_engine = nullptr;
    // live cost tracking (wf_scene_track_costs)
    wf_cost_tracker* track_mat = nullptr;
    wf_cost_tracker* track_tex = nullptr;
    wf::CostProbe probe;
    double ns_per_tick = 10.0;
};

namespace wf {

wf_status scene_alloc(wf_scene* s, size_t bytes, void** out);
template <typename T>
wf_status scene_upload(wf_scene* s, const T* host, size_t count, T** out) {
    if (count == 0) {
        *out = nullptr;
        return WF_OK;
    }
    void* p = nullptr;
    WF_TRY(scene_alloc(s, count * sizeof(T), &p));
    WF_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s->ctx->stream));
    *out = (T*)p;
    return WF_OK;
}

// Element count of a stage. With `dev` null, `n` is exact. Otherwise kernels read the count from
// device memory (*dev <= n, written by an earlier kernel on the same stream) and `n` only sizes
// the launch grid: no host round trip, so the stage can run inside a captured graph.
struct Count {
    int32_t n = 0;
    const int32_t* dev = nullptr;
};
inline Count host_count(int32_t n) { return Count{n, nullptr}; }

// Per-ray job tags of an engine wavefront: tag = job slot << kTagJobShift | sample index.
constexpr int kTagJobShift = 20;
constexpr int32_t kTagSampleMask = (1 << kTagJobShift) - 1;
constexpr int kMaxJobs = 1 << (31 - kTagJobShift);

// Per-job constants of an engine wavefront (device table indexed by job slot).
struct JobDev {
    int32_t pixel_base;        // first framebuffer pixel of the job
    int32_t width, height;
    int32_t spp;               // sample budget per pixel
    int32_t max_depth;
    uint32_t seed;
    int32_t tag;               // job id in the counter-based RNG (wf_shade_params::job)
    int32_t adaptive;
    float adaptive_threshold;
    int32_t adaptive_min_samples;
    int32_t adaptive_interval;
    int32_t lanes;             // samples in flight per pixel (power of two); sample s runs in
                               // lane s & (lanes - 1), whose accumulator is pixel + lane * stride
    float cam[12];             // eye, top-left, horizontal, vertical (Camera.h:13-33)
    int32_t lane_stride;       // accumulator index distance between the lanes of a pixel
    int32_t first_sample;      // RNG index of sample 0
};

// Framebuffer pixel of an accumulator index a of job J running sample s.
__host__ __device__ __forceinline__ int32_t lane_pixel(const JobDev& J, int32_t a, int32_t s) {
    return a - (s & (J.lanes - 1)) * J.lane_stride;
}

// Camera basis (Camera.h:13-33): v = eye, top-left corner, horizontal, vertical (12 floats).
struct CamBasis {
    float v[12];
};
CamBasis camera_basis(const wf_camera& c, int width, int height);

// ---- primitives (wf_sort.hip) ------------------------------------------------------------------
// Stable LSD radix sort of (key, value) pairs on key bits [0, key_bits). When vals_in is NULL
// the values start as 0..n-1. The sorted pairs end in keys_out/vals_out (which must not alias
// the inputs); keys_in is not modified. A host count sorts only the bit range where the keys
// differ (one readback); a device count sorts all key_bits.
wf_status radix_sort_pairs(wf_context* ctx, const uint64_t* keys_in, const int32_t* vals_in,
                           uint64_t* keys_out, int32_t* vals_out, Count n, int32_t key_bits,
                           const char* tag);
// Stable LSD radix sort of packed words: word i holds its key in bits [idx_bits, idx_bits +
// key_bits) and its own position i in bits [0, idx_bits). Sorting on the key bits alone is the
// same permutation as sorting (key, i) pairs, with 8 bytes moved per element instead of 12;
// order receives the low idx_bits bits of the sorted words. scratch (n words) is overwritten;
// words is not. A host count sorts only the key bits where the words differ. key_lo > 0 leaves the
// low key_lo key bits out of the order (a stable sort by the key's top key_bits - key_lo bits).
wf_status radix_sort_packed(wf_context* ctx, const uint64_t* words, uint64_t* scratch,
                            int32_t* order, Count n, int32_t idx_bits, int32_t key_bits,
                            const char* tag, int32_t key_lo = 0);
// Coherence metrics of one shading batch (wf_coherence.hip): adds the batch's material/texture
// run lengths (and, with lines, cache-line homogeneities) times its size into acc[0..3], its
// size into acc[4], and stores the batch's own values in acc[5..8] (9 device doubles).
wf_status coherence_accumulate(wf_context* ctx, const int32_t* mat, const int32_t* tex, Count n,
                               bool lines, double* acc);
// Reads acc back and stores the size-weighted averages in st (left 0 without batches).
wf_status coherence_read(wf_context* ctx, const double* acc, wf_render_stats* st);
// Exclusive prefix sum of n u32 flags/counts into out (may alias in); total (optional) is
// written to a device int32.
wf_status exclusive_scan_u32(wf_context* ctx, const uint32_t* in, uint32_t* out, Count n,
                             int32_t* total_dev, const char* tag);
// dst[i] = i
wf_status iota_i32(wf_context* ctx, int32_t* dst, int32_t n);
wf_status iota_count(wf_context* ctx, int32_t* dst, Count n);
// *dst = *a - *b (device int32s), one thread.
wf_status sub_count(wf_context* ctx, int32_t* dst, const int32_t* a, const int32_t* b);

// ---- stage entry points (used by the C-ABI, the render loop and the engine) -------------------
// Every stage runs on ctx's stream with ctx's workspaces; the scene may belong to another
// context on the same device.
wf_status scene_ready(const wf_scene* s);
wf_status trace_closest(wf_context* ctx, const wf_scene* s, const wf_ray_soa* rays, Count n,
                        float tnear, float tfar, int32_t* hit_leaf, float* hit_t, float* hit_bu,
                        float* hit_bv, uint32_t* flags);
wf_status trace_occluded(const wf_scene* s, const float* ox, const float* oy, const float* oz,
                         const float* dx, const float* dy, const float* dz, const float* maxd,
                         int32_t n, uint8_t* out);
// Closest hit + stable split; counts_dev[0] = hits, counts_dev[1] = misses (device int32).
// tags_in/hit_tags (optional) carry job tags from the rays to the hit queue.
// unordered: hits and misses are appended in any order (one kernel, no scan).
wf_status intersect_split(wf_context* ctx, const wf_scene* s, const wf_ray_soa* rays, Count n,
                          wf_shade_soa* hits_out, wf_ray_soa* misses_out, int32_t* counts_dev,
                          const int32_t* tags_in = nullptr, int32_t* hit_tags = nullptr,
                          int32_t* miss_tags = nullptr, bool unordered = false);
wf_status ray_schedule(wf_context* ctx, const wf_ray_soa* rays, Count n, const float bmin[3],
                       const float bmax[3], int32_t* sorted, uint64_t* keys_out);
// ShadingQueue::schedule. spatial = false (engine use): the cost-benefit key keeps only the
// (material type, priority rank) group and the stable sort keeps the input order inside each
// group (the rays were Morton-sorted before tracing, so hits arrive in spatial order); the
// C-ABI keeps the reference's full key.
wf_status shading_schedule(wf_context* ctx, const wf_scene* s, const wf_shade_soa* q, Count n,
                           int32_t policy, int32_t* sorted, bool spatial = true);

// Job context of a tagged wavefront for the shading kernel: per-entry tags plus the job table.
// Engine-internal ray queue: three 16-byte columns per ray, a = (origin, pixel bits),
// b = (direction, pdf), c = (throughput, depth bits).
// Frame of RayQueue::sortKey (include/scheduling/RayQueue.h:135-162): the scene bounds minimum,
// the extents max(max - min, 1e-6) per axis, and the number of index bits below the 33-bit key
// in a packed sort word (0: the bare key).
struct RayKeyFrame {
    float lo[3];
    float ext[3];
    int32_t idx_bits;
};
RayKeyFrame ray_key_frame(const float bmin[3], const float bmax[3], int32_t idx_bits);
// Smallest b >= 1 with 2^b >= n: the index bits of packed sort words over n elements.
int32_t index_bits(int64_t n);

// Fixed buffers for a radix sort that may not allocate: the second buffer of packed words and a
// meta block (tile counters, digit totals, look-back states) of meta_words 32-bit words.
struct SortArena {
    uint64_t* keys_alt = nullptr;
    uint32_t* meta = nullptr;
    size_t meta_words = 0;
};
// Stable sort of packed words (key_bits above idx_bits) by all key bits with 8-bit digits into
// the arena: no allocation and no readback. words is overwritten (it is the other ping-pong
// buffer); order receives the low idx_bits of each sorted word. A look-back wait that runs out
// sets the context's fault word, which the next synchronising call reports.
wf_status radix_sort_packed_fixed(wf_context* ctx, uint64_t* words, int32_t* order, Count n,
                                  int32_t idx_bits, int32_t key_bits, const SortArena& arena,
                                  const char* tag);
// Looks up this device's copy of the fixed scratch array (at context creation; the context's
// device is current).
wf_status fixed_scratch_init(wf_context* ctx);

struct RayPacked {
    float4* a = nullptr;
    float4* b = nullptr;
    float4* c = nullptr;
};
// Packed rays visited in a sorted order without moving them: ray i is (a, b, c)[order[i]], and
// its job tag is read at order[i] too.
struct RayPackedOrdered {
    const float4* a = nullptr;
    const float4* b = nullptr;
    const float4* c = nullptr;
    const int32_t* order = nullptr;
};
wf_status trace_closest(wf_context* ctx, const wf_scene* s, const RayPacked* rays, Count n,
                        float tnear, float tfar, int32_t* hit_leaf, float* hit_t, float* hit_bu,
                        float* hit_bv, uint32_t* flags);
struct DeferredShadow;
// Closest hits of packed rays into hit buffers (leaf, t, u, v per ray), with the shadow walks of
// a deferred bounce traced by the same persistent launch.
wf_status trace_closest_deferred(wf_context* ctx, const wf_scene* s, const RayPacked* rays,
                                 Count n, int32_t* hit_leaf, float* hit_t, float* hit_bu,
                                 float* hit_bv, const DeferredShadow& d);
wf_status trace_closest_deferred(wf_context* ctx, const wf_scene* s, const RayPackedOrdered* rays,
                                 Count n, int32_t* hit_leaf, float* hit_t, float* hit_bu,
                                 float* hit_bv, const DeferredShadow& d);
wf_status trace_closest(wf_context* ctx, const wf_scene* s, const RayPackedOrdered* rays, Count n,
                        float tnear, float tfar, int32_t* hit_leaf, float* hit_t, float* hit_bu,
                        float* hit_bv, uint32_t* flags);
// The context's closest-hit buffers for n rays.
wf_status hit_buffers(wf_context* ctx, Count n, int32_t** leaf, float** t, float** u, float** v);
// The hit/miss split of traced packed rays (hits and misses appended at device cursors).
// Ranges of the shading queue per group key of a shading policy (policy_buckets below).
constexpr int kPolicyBucketMax = 1024;
struct PolicyRanges {
    const int32_t* bucket = nullptr;   // range of each material
    int32_t* cursor = nullptr;         // next free slot of each range (advanced by the split)
};
wf_status split_traced(wf_context* ctx, const wf_scene* s, const RayPacked* rays, Count n,
                       const int32_t* hit_leaf, const float* hit_t, const float* hit_bu,
                       const float* hit_bv, wf_shade_soa* hits_out, wf_ray_soa* misses_out,
                       int32_t* counts_dev, const int32_t* tags_in, int32_t* hit_tags,
                       int32_t* miss_tags,
                       const PolicyRanges* ranges = nullptr);
// The same over rays in a sorted order (tags_in in the unsorted order, read through it).
wf_status split_traced(wf_context* ctx, const wf_scene* s, const RayPackedOrdered* rays, Count n,
                       const int32_t* hit_leaf, const float* hit_t, const float* hit_bu,
                       const float* hit_bv, wf_shade_soa* hits_out, wf_ray_soa* misses_out,
                       int32_t* counts_dev, const int32_t* tags_in, int32_t* hit_tags,
                       int32_t* miss_tags,
                       const PolicyRanges* ranges = nullptr);
wf_status ray_schedule(wf_context* ctx, const RayPacked* rays, Count n, const float bmin[3],
                       const float bmax[3], int32_t* sorted, uint64_t* keys_out);

// The columns of a shaded entry that the kernels after shading read (next-ray append, shadow
// walks, path completion): hit point, accumulator index, depth and job tag, by shading position.
struct HitCols {
    const float* px = nullptr;
    const float* py = nullptr;
    const float* pz = nullptr;
    const int32_t* pixel = nullptr;
    const int32_t* depth = nullptr;
    const int32_t* tag = nullptr;    // NULL without jobs
};

// Hits shaded in place (the engine's index mode): entry j of the shading order is traced ray
// list[j] of the packed rays (a, b, c), visited through `order` when it is set, with its closest
// hit in the hit buffers (leaf, t, u, v by traced-ray index). The same form lists misses.
struct HitSource {
    const float4* a = nullptr;
    const float4* b = nullptr;
    const float4* c = nullptr;
    const int32_t* order = nullptr;  // sorted view: traced ray i is slot order[i]; NULL = i
    const int32_t* list = nullptr;   // traced-ray index of each entry
    const int32_t* leaf = nullptr;
    const float* t = nullptr;
    const float* u = nullptr;
    const float* v = nullptr;
    const int32_t* tags = nullptr;   // job tag of each ray slot
};

// Hits copied into shading order (a policy's ranges) by the split: packed ray, closest hit and
// job tag of each entry, so that shading reads them sequentially (HitSource with no list).
struct HitRecords {
    float4* a = nullptr;
    float4* b = nullptr;
    float4* c = nullptr;
    int32_t* leaf = nullptr;
    float* t = nullptr;
    float* u = nullptr;
    float* v = nullptr;
    int32_t* tag = nullptr;
};

// Per-entry outputs of k_shade consumed by the shadow and compaction kernels.
struct ShadeOut {
    float* ldx; float* ldy; float* ldz; float* lmax;   // area/environment shadow ray
    float* lr; float* lg; float* lb;                   // its contribution
    float* dr; float* dg; float* db;                   // directional-light contribution
    uint32_t* shadow_flag;                             // bit0 area/env ray, bit1 directional
    float* ndx; float* ndy; float* ndz;                // scattered direction
    float* ntr; float* ntg; float* ntb;                // next throughput
    float* npdf;
    uint32_t* alive;
    // unordered append mode (queue order free): shadow entries and next rays are appended with
    // wave-aggregated atomics instead of a scan + compaction
    int32_t append;
    int32_t* shadow_list;
    int32_t* shadow_entries;                           // device counter
    int32_t* shadow_rays;                              // device counter
    int32_t* next_count;                               // device counter
    wf_ray_soa next;
    RayPacked next_packed;                             // used instead of `next` when set
    int32_t* next_tags;
    uint64_t* next_keys;                               // sort word of each next ray, when set
    RayKeyFrame key_frame;
    CostProbe probe;                                   // live cost tracking, when set
    // index-mode shading: the hit columns of every entry (HitCols), for the kernels after it
    float* hpx; float* hpy; float* hpz;
    int32_t* hpixel; int32_t* hdepth; int32_t* htag;
};

// Shadow entries of a shaded bounce whose occlusion walks were handed to the next closest-hit
// launch instead of a launch of their own: what that launch needs of the bounce.
struct DeferredShadow {
    bool armed = false;
    HitCols hc{};                 // the bounce's hit points and accumulators
    ShadeOut o{};                 // its shading outputs (shadow rays and contributions)
    const int32_t* list = nullptr;   // shadow entries (shading-queue indices)
    Count n{};                    // their count (device)
    float* accum = nullptr;
};

struct ShadeJobs {
    const int32_t* tags = nullptr;     // per shading-queue entry
    const JobDev* jobs = nullptr;
    int32_t* next_tags = nullptr;      // tags of the compacted next rays
    RayPacked next_packed;             // unordered mode: next rays in the packed layout
    uint64_t* next_keys = nullptr;     // unordered mode: sort words of the next rays
    RayKeyFrame key_frame{};           // ... in this frame
    bool unordered = false;            // next rays / shadow entries may come in any order
    DeferredShadow* defer = nullptr;   // unordered mode: record the shadow walks here instead
                                       // of running them (trace_closest_deferred runs them)
};
// One bounce of shadeAll: emission into accum, shadow rays resolved, next rays compacted into
// `next`; counts_dev[0] = next rays, counts_dev[1] = shadow rays traced, counts_dev[2] = shadow
// entries (device int32). With jobs.tags, RNG sample/job/seed and the depth limit come from each
// entry's job; *alive_out (optional) receives the per-entry "path continues" flags.
wf_status shade_bounce(wf_context* ctx, const wf_scene* s, const wf_shade_soa* q, Count n,
                       const wf_shade_params* p, float* accum, wf_ray_soa* next,
                       int32_t* counts_dev, const ShadeJobs& jobs = ShadeJobs{},
                       const uint32_t** alive_out = nullptr, const CostProbe* probe = nullptr);
// Relative-cost tables of the scene (dense descending material ranks, relative costs); NULL =
// neutral.
wf_status scene_upload_costs(wf_scene* s, const double* mat_cost, const double* tex_cost);
// Live cost tracking around one bounce of the pass loop (no-ops without attached trackers):
// refresh the tables from the trackers, clear the probe, record the probe into the trackers.
wf_status costs_refresh(wf_scene* s);
wf_status costs_clear(wf_scene* s);
wf_status costs_record(wf_scene* s);
wf_status shade_misses(wf_context* ctx, const wf_scene* s, const wf_ray_soa* q, Count n,
                       int32_t physics, float* accum);
// Index mode (queue order free, shadow walks deferred): the hits listed by h (h.list) shaded in
// place; *hc_out receives the hit columns the kernels after shading read.
wf_status shade_hits(wf_context* ctx, const wf_scene* s, const HitSource& h, Count n,
                     const wf_shade_params* p, float* accum, int32_t* counts_dev,
                     const ShadeJobs& jobs, const uint32_t** alive_out, HitCols* hc_out);
// Index-mode split of traced packed rays (src: the rays and their hit buffers; its list is not
// used): the traced-ray index of every hit appended to hit_list (with ranges, into the range of
// its group key), or, with ranges and records, the hit's ray, closest hit and tag copied into
// its range of `records` (an order of many sparse ranges scatters each range's hits over the
// whole ray buffer: copied once in ray order, they are read sequentially by the shading);
// every miss resolved in place (sky or environment into accum)
// with its accumulator index and job tag appended to miss_pixel / miss_tag for path
// completion; counts_dev[0] = hits, counts_dev[1] = misses.
wf_status split_index(wf_context* ctx, const wf_scene* s, const HitSource& src, Count n,
                      int32_t* hit_list, int32_t physics, float* accum, int32_t* miss_pixel,
                      int32_t* miss_tag, int32_t* counts_dev,
                      const PolicyRanges* ranges = nullptr, const HitRecords* records = nullptr);
// The shading queue of listed hits (HitRecord mapping of each), with their job tags: the
// engine's diagnostic of an index-mode iteration.
wf_status gather_hits(wf_context* ctx, const wf_scene* s, const HitSource& h, Count n,
                      wf_shade_soa* dst, int32_t* tags_out);
wf_status render_regenerate(wf_scene* s, const wf_render_params* rp, float* image_rgb,
                            int32_t* sample_counts, wf_render_stats* stats);
// Binary hierarchy by PLOC over n >= 2 Morton-ordered leaf boxes (Karras-layout outputs).
wf_status ploc_build(wf_context* ctx, int n, int radius, const float4* slo, const float4* shi,
                     int32_t* left, int32_t* right, int32_t* iparent, int32_t* lparent,
                     float4* nlo, float4* nhi);
// Surface-area-optimal collapse of a binary hierarchy (root 0, parent links iparent/lparent)
// into four-wide nodes (out4) or eight-wide quantised nodes (out8), whichever is given.
wf_status collapse_wide(wf_context* ctx, int n, const int32_t* left, const int32_t* right,
                        int32_t* iparent, const int32_t* lparent, const float4* slo,
                        const float4* shi, const float4* nlo, const float4* nhi, Bvh4Node* out4,
                        Bvh8Node* out8, int32_t* n_nodes, int32_t* stack_bound);
wf_status ray_keys(wf_context* ctx, const wf_ray_soa* rays, Count n, const float bmin[3],
                   const float bmax[3], uint64_t* keys);
wf_status shading_keys(wf_context* ctx, const wf_scene* s, const wf_shade_soa* q, Count n,
                       int32_t policy, bool spatial, uint64_t* keys, int32_t* key_bits);
wf_status gather_rays(wf_context* ctx, const wf_ray_soa* src, const int32_t* idx, Count n,
                      wf_ray_soa* dst, const int32_t* tags_in = nullptr,
                      int32_t* tags_out = nullptr);
wf_status gather_shading(wf_context* ctx, const wf_shade_soa* src, const int32_t* idx, Count n,
                         wf_shade_soa* dst, const int32_t* tags_in = nullptr,
                         int32_t* tags_out = nullptr);

// Shading order of the regenerating wavefront without a sort: the engine's queue order changes
// no result, so a policy's order is produced by appending every hit into the range of its group
// key, the ranges laid out in key order. Counts the materials of the traced rays' hits (and, for
// costBenefit, their throughput luminance), computes each material's group key and leaves in
// `out` the range of every material (bucket, n_materials entries) and the first slot of every
// range (cursor); hits_dev receives the hit count. False in *used when the scene has no
// materials or more than kPolicyBucketMax of them (then the caller sorts and gathers instead).
wf_status policy_buckets(wf_context* ctx, const wf_scene* s, const RayPacked* rays, Count n,
                         const int32_t* hit_leaf, int32_t policy, int32_t* hits_dev,
                         PolicyRanges* out, bool* used);
wf_status policy_buckets(wf_context* ctx, const wf_scene* s, const RayPackedOrdered* rays, Count n,
                         const int32_t* hit_leaf, int32_t policy, int32_t* hits_dev,
                         PolicyRanges* out, bool* used);

}  // namespace wf
