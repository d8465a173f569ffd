/*
 * wavefront_hip.h -- C-ABI of the MI355X (gfx950) wavefront path-tracing hot path.
 *
 * The reference renderer (fhidargue/wavefront-renderer) has no FFI: its seam is the set of C++
 * class surfaces that WavefrontRenderer::renderScene drives once per bounce. Every entry point
 * below replaces one of those surfaces; the comment on each names the reference interface
 * (file:line in the reference tree) it stands in for.
 *
 * Conventions
 *  - Every entry point returns wf_status (WF_OK == 0). On failure wf_last_error() returns a
 *    thread-local message describing the failing call.
 *  - Queue buffers (wf_ray_soa, wf_shade_soa, index arrays, key arrays) are caller-owned DEVICE
 *    pointers. The library never frees them. Scene/context handles are library-owned and have an
 *    explicit destroy, like the reference's move-only Embree handles (include/accelerator/Embree.h:16-22).
 *  - Calls are stream-ordered on the context's HIP stream unless stated otherwise; functions
 *    that return a count to the host synchronise that stream.
 *  - Scene data is copied on wf_scene_create, so there is no back-pointer hazard
 *    (the reference keeps a raw Scene* inside the accelerator, Embree.h:29).
 *  - Calls that share one context must not overlap (one in-flight frame per context), as the
 *    reference renderer is not re-entrant (include/render/WavefrontRenderer.h:46-48).
 */
#ifndef WAVEFRONT_HIP_H
#define WAVEFRONT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t wf_status;
#define WF_OK 0
#define WF_ERR_INVALID_ARGUMENT 1
#define WF_ERR_HIP 2
#define WF_ERR_OUT_OF_MEMORY 3
#define WF_ERR_CAPACITY 4
#define WF_ERR_NOT_BUILT 5
#define WF_ERR_UNSUPPORTED 6
#define WF_ERR_ADMISSION 7

/* ---- material / policy enums (values are the reference's enum order) -------------------- */

/* MaterialType, include/shading/Material.h:9-17. The CostBenefit policy orders by this value. */
#define WF_MAT_DIFFUSE 0
#define WF_MAT_METAL 1
#define WF_MAT_EMISSIVE 2
#define WF_MAT_SPOTLIGHT 3
#define WF_MAT_GLASS 4
#define WF_MAT_PLASTIC 5

/* BSDF model of a material: the reference's own BSDFs (Material.cpp:165-266) or the microfacet
 * models (GGX conductor for METAL, exact-Fresnel smooth dielectric for GLASS). */
#define WF_MODEL_REFERENCE 0
#define WF_MODEL_MICROFACET 1

/* SchedulingPolicy, include/scheduling/ShadingQueue.h:18-24. */
#define WF_POLICY_NONE 0
#define WF_POLICY_MATERIAL 1
#define WF_POLICY_TEXTURE 2
#define WF_POLICY_COST_BENEFIT 3

/* Light transport track.
 * REFERENCE: balance-heuristic MIS, emitter mesh then triangle picked uniformly with pdf 1/A_total
 *            (src/core/Scene.cpp:192-237, include/shading/MIS.h:14-31), nearest-texel environment.
 * UNBIASED:  power-heuristic MIS, area*power alias table over emissive triangles, environment map
 *            importance-sampled through a luminance alias table with environment MIS, NEE for
 *            every non-delta BSDF. */
#define WF_PHYSICS_REFERENCE 0
#define WF_PHYSICS_UNBIASED 1

/* Texture texel formats. RGBA8_SRGB texels are decoded with the reference's sRGB curve
 * (src/shading/Texture.cpp:20-28) through a 256-entry table; RGB_F32 texels are used as is. */
#define WF_TEX_RGBA8_SRGB 0
#define WF_TEX_RGB_F32 1

/* ---- scene description (mirrors Scene::addMaterial/addTexture/addMesh, Scene.cpp:5-56) ---- */

typedef struct wf_material {
    int32_t type;              /* WF_MAT_* */
    float albedo[3];
    float roughness;
    float emission;
    int32_t texture_id;        /* -1 = none */
    float spot_outer_deg;
    float spot_falloff_deg;
    float ior;
    int32_t use_spatial_checker;
    float checker_cell_size;
    int32_t model;             /* WF_MODEL_* */
} wf_material;

typedef struct wf_texture_desc {
    int32_t width;
    int32_t height;
    int32_t format;            /* WF_TEX_* */
    const void* data;          /* host pointer: width*height*4 bytes (RGBA8) or *3 floats */
} wf_texture_desc;

typedef struct wf_mesh_desc {
    const float* positions;    /* host, xyz interleaved, vertex_count*3 */
    int32_t vertex_count;
    const int32_t* indices;    /* host, 3 per triangle, mesh-local vertex indices */
    int32_t triangle_count;
    const float* uv_u;         /* host, face-varying, 3 per triangle, or NULL (Mesh.h:8-20) */
    const float* uv_v;
    int32_t material_id;
} wf_mesh_desc;

typedef struct wf_directional_light {   /* DirectionalLight, include/core/Scene.h:57-62 */
    float direction[3];
    float color[3];
    float intensity;           /* 0 = absent */
} wf_directional_light;

typedef struct wf_environment_desc {    /* EnvironmentMap, include/core/EnvironmentMap.h */
    int32_t width;             /* 0 = no map: the reference's sky gradient is used */
    int32_t height;
    const float* rgb;          /* host, width*height*3 floats, equirectangular */
    float intensity;
    float tint[3];
} wf_environment_desc;

typedef struct wf_scene_desc {
    const wf_mesh_desc* meshes;           /* geomID == array position (attach order, Embree.cpp:77-108) */
    int32_t mesh_count;
    const wf_material* materials;
    int32_t material_count;
    const wf_texture_desc* textures;
    int32_t texture_count;
    wf_directional_light directional;
    wf_environment_desc environment;
} wf_scene_desc;

/* ---- queues (structure of arrays, field order of RayQueue.h:66-90 and ShadingQueue.h:29-53) - */

typedef struct wf_ray_soa {       /* RayQueue */
    float* ox; float* oy; float* oz;
    float* dx; float* dy; float* dz;
    float* tr; float* tg; float* tb;   /* throughput */
    int32_t* pixel;
    int32_t* depth;
    float* pdf;
} wf_ray_soa;

typedef struct wf_shade_soa {     /* ShadingQueue, plus the geomID and front-face bit it drops */
    float* px; float* py; float* pz;   /* hit point */
    float* nx; float* ny; float* nz;   /* normal facing the ray */
    float* t;
    float* u; float* v;                /* interpolated UV, or barycentrics without UVs */
    int32_t* tri;                      /* primID within its mesh */
    int32_t* mat;
    int32_t* tex;
    float* ox; float* oy; float* oz;
    float* dx; float* dy; float* dz;
    float* tr; float* tg; float* tb;
    int32_t* pixel;
    int32_t* depth;
    float* pdf;
    int32_t* geom;                     /* geomID (mesh index) */
    uint8_t* front;                    /* HitRecord::frontFace (HitRecord.h:15) */
} wf_shade_soa;

/* Per-ray closest-hit result (Embree RTCHit in the reference). prim < 0 means miss. */
typedef struct wf_hit_soa {
    float* t;
    int32_t* geom;                     /* geomID */
    int32_t* prim;                     /* primID within the mesh */
    float* bu; float* bv;              /* barycentrics of vertex 1 and 2 */
} wf_hit_soa;

/* ---- context and device memory ------------------------------------------------------------ */

const char* wf_last_error(void);
const char* wf_version(void);

typedef struct wf_context wf_context;
typedef struct wf_scene wf_scene;

wf_status wf_context_create(int32_t device, wf_context** out);
wf_status wf_context_destroy(wf_context* ctx);
/* The context's HIP stream (hipStream_t) every stream-ordered call below is issued on. */
wf_status wf_context_stream(wf_context* ctx, void** out_stream);
wf_status wf_device_alloc(wf_context* ctx, size_t bytes, void** out);
wf_status wf_device_free(wf_context* ctx, void* ptr);
wf_status wf_memcpy_htod(wf_context* ctx, void* dst, const void* src, size_t bytes);
wf_status wf_memcpy_dtoh(wf_context* ctx, void* dst, const void* src, size_t bytes);
wf_status wf_memcpy_dtod(wf_context* ctx, void* dst, const void* src, size_t bytes);
wf_status wf_fill_u32(wf_context* ctx, void* dst, uint32_t value, size_t count);
/* Streaming device copy (16-byte loads and stores, grid-stride) of `bytes` (a multiple of 16).
 * No reference counterpart: it measures the achievable HBM bandwidth beside the hot path and
 * calibrates hardware byte counters against a kernel of known traffic. */
wf_status wf_stream_copy(wf_context* ctx, void* dst, const void* src, size_t bytes);
wf_status wf_synchronize(wf_context* ctx);

/* ---- scene and accelerator ------------------------------------------------------------------- */

typedef struct wf_build_stats {   /* EmbreeAccelerator m_meshCount/m_totalTriangles/... (Embree.h:31-35) */
    int32_t mesh_count;
    int32_t triangle_count;
    int32_t vertex_count;
    int32_t node_count;           /* wide BVH nodes (of bvh_width children each) */
    int32_t max_depth;            /* depth of the binary hierarchy the wide BVH is collapsed from */
    int32_t stack_bound;          /* most entries the traversal stack can hold (<= 64 supported) */
    int32_t bvh_width;            /* 4 or 8: the node width the build used (a tuning of 8 falls
                                     back to 4 for one triangle or a stack bound above 64) */
    double build_ms;              /* device time of the BVH build kernels */
    float bounds_min[3];          /* Scene::computeBounds over all vertices (Scene.cpp:59-77) */
    float bounds_max[3];
} wf_build_stats;

/* Scene::addMesh/addMaterial/addTexture + upload. Copies everything to the device. */
wf_status wf_scene_create(wf_context* ctx, const wf_scene_desc* desc, wf_scene** out);
wf_status wf_scene_destroy(wf_scene* scene);
/* Scene::buildAccelerator + EmbreeAccelerator::build (Scene.cpp:79-113, Embree.cpp:66-119):
 * GPU LBVH over the flattened triangle soup collapsed to four-wide nodes, emitter tables and
 * bounds. */
wf_status wf_accel_build(wf_scene* scene, wf_build_stats* stats);
/* Device bytes the scene and its accelerator hold (admission accounting). */
wf_status wf_scene_device_bytes(const wf_scene* scene, size_t* out);

/* EmbreeAccelerator::intersect4 / intersect over a whole batch (Embree.cpp:122-256): closest hit
 * for rays [0,n) with tnear/tfar; writes the raw hit per ray (prim < 0 = miss). */
wf_status wf_intersect(const wf_scene* scene, const wf_ray_soa* rays, int32_t n, float tnear,
                       float tfar, wf_hit_soa* hits);
/* EmbreeAccelerator::intersect4 / intersect with the HitRecord mapping (Embree.cpp:164-174,
 * 240-252) for rays [0,n): records slot i receives ray i's HitRecord (plus the ray state) when
 * hit_flags[i] == 1. Any tnear/tfar (Scene::hit passes its own range, Scene.cpp:120-160). */
wf_status wf_intersect_records(const wf_scene* scene, const wf_ray_soa* rays, int32_t n,
                               float tnear, float tfar, wf_shade_soa* records,
                               uint8_t* hit_flags);
/* Traversal statistics for rays [0,n) (closest hit, or any hit when any_hit != 0): BVH nodes
 * visited and triangles tested per ray into device int32 arrays. Lanes of a 64-wide wave are
 * consecutive rays, so lane divergence follows from the counts. No reference counterpart. */
wf_status wf_trace_stats(const wf_scene* scene, const wf_ray_soa* rays, int32_t n, float tnear,
                         float tfar, int32_t any_hit, int32_t* nodes_visited,
                         int32_t* triangles_tested);
/* EmbreeAccelerator::occluded (Embree.cpp:258-279) for n shadow rays: occluded[i] = 1 when any
 * triangle lies in [0.001, max_dist[i]-0.001] (and, as in the reference, whenever that upper
 * bound is negative). Inputs are device arrays. */
wf_status wf_occluded(const wf_scene* scene, const float* ox, const float* oy, const float* oz,
                      const float* dx, const float* dy, const float* dz, const float* max_dist,
                      int32_t n, uint8_t* occluded);
/* WavefrontRenderer::intersectAll (src/render/WavefrontRenderer.cpp:459-567): closest hit, then a
 * stable split into a ShadingQueue (HitRecord mapping of Embree.cpp:240-252) and a miss RayQueue,
 * both in ascending input order. Counts are returned to the host. */
wf_status wf_intersect_split(const wf_scene* scene, const wf_ray_soa* rays, int32_t n,
                             wf_shade_soa* hits_out, int32_t* n_hits, wf_ray_soa* misses_out,
                             int32_t* n_misses);

/* ---- scheduling ---------------------------------------------------------------------------- */

/* RayQueue::sortKey + schedule (include/scheduling/RayQueue.h:135-179): 33-bit octant+Morton
 * key per ray against the given bounds, then a stable sort; sorted_indices[i] is the input index
 * at sorted position i. keys_out (optional, may be NULL) receives the unsorted keys. */
wf_status wf_ray_schedule(wf_context* ctx, const wf_ray_soa* rays, int32_t n,
                          const float bounds_min[3], const float bounds_max[3],
                          int32_t* sorted_indices, uint64_t* keys_out);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): wf_ray_schedule with all of
 * its scratch (sort words, the sort's ping-pong buffer, tile counters, digit totals, look-back
 * states) in one fixed array per device of WF_NOALLOC_MAX_RAYS rays (about 36 MB, held by each
 * device the library runs on until the process ends). It calls no hipMalloc, hipFree, pinned
 * host allocation or event creation, and rejects more than WF_NOALLOC_MAX_RAYS rays and contexts
 * with kernel timing on. It only enqueues work on the context's stream: the permutation (the same
 * as wf_ray_schedule's, bit for bit) is ready, and an exhausted look-back wait is reported, at the
 * next synchronising call (wf_synchronize). The array keeps the last call's contents; every call
 * zero-fills what must start at zero. Calls that overlap on one device (another context's stream,
 * another thread) share the array and race: serialise them. */
#define WF_NOALLOC_MAX_RAYS (1 << 21)
wf_status wf_ray_schedule_noalloc(wf_context* ctx, const wf_ray_soa* rays, int32_t n,
                                  const float bounds_min[3], const float bounds_max[3],
                                  int32_t* sorted_indices, uint64_t* keys_out);
/* RayQueue::materialize (RayQueue.h:182-215), out of place: dst[i] = src[idx[i]]. */
wf_status wf_materialize_rays(wf_context* ctx, const wf_ray_soa* src, const int32_t* idx,
                              int32_t n, wf_ray_soa* dst);

/* Cost tables for the Material/CostBenefit policies and cost-aware RR: relativeCost per material
 * and per texture (CostTracker::relativeCost, include/render/CostTracker.h:44-54). NULL restores
 * the neutral tables (every cost 1.0), which is what a CostTracker with no samples returns. */
wf_status wf_scene_set_costs(wf_scene* scene, const double* material_rel_cost,
                             const double* texture_rel_cost);

/* CostTracker (include/render/CostTracker.h:8-63): per item (material or texture) an
 * exponential moving average (weight 0.05) of its shading cost in nanoseconds, the first sample
 * taken as is; relativeCost = average / mean of the tracked items' averages once the item has 50
 * samples, else 1. Host objects; no device work. */
typedef struct wf_cost_tracker wf_cost_tracker;
wf_status wf_cost_tracker_create(int32_t items, wf_cost_tracker** out);
wf_status wf_cost_tracker_destroy(wf_cost_tracker* tracker);
/* CostTracker::record */
wf_status wf_cost_tracker_record(wf_cost_tracker* tracker, int32_t id, double cost_ns);
/* averageCostNanoseconds, sampleCount and relativeCost of every item (NULL arrays skipped). */
wf_status wf_cost_tracker_query(const wf_cost_tracker* tracker, double* average_ns,
                                int32_t* samples, double* relative);
/* Live cost tracking (shadeAll's cost sampling, cpp:612-627, 746-758, 800-844): while trackers
 * are attached, every bounce of wf_render's pass loop times a quarter of its shaded hits on the
 * device (a per-hit coin, the reference's shouldTimeThisRay): the surface-colour fetch of a
 * textured material per texture and the BSDF scatter per material, both as the wave's wall-clock
 * time. Each bounce's per-item mean is recorded into the trackers, and before every bounce's
 * shading-queue schedule the scene's relative-cost tables (the policies' and cost-aware RR's
 * input) are refreshed from them. The reference starts fresh trackers per render; attach new
 * ones to do the same. NULL, NULL detaches: the tables then stay as last set. Either tracker
 * must have as many items as the scene has materials / textures. */
wf_status wf_scene_track_costs(wf_scene* scene, wf_cost_tracker* materials,
                               wf_cost_tracker* textures);

/* ShadingQueue::schedule (ShadingQueue.h:90-141, 277-329, 437-550) for the four policies.
 * sorted_indices[i] = queue index shaded at position i. Ties the reference leaves to an unstable
 * sort are broken by ascending queue index. */
wf_status wf_shading_schedule(const wf_scene* scene, const wf_shade_soa* queue, int32_t n,
                              int32_t policy, int32_t* sorted_indices);
/* ShadingQueue::materialize (ShadingQueue.h:233-275), out of place. */
wf_status wf_materialize_shading(wf_context* ctx, const wf_shade_soa* src, const int32_t* idx,
                                 int32_t n, wf_shade_soa* dst);
/* The coherence metrics of one shading queue in its shading order (renderScene's batch
 * statistics, cpp:67-92 computeAverageRunLength and cpp:201-246 cacheLineHomogeneity):
 * out[0] material run length and out[1] texture run length over the entries with a texture
 * (texture_ids >= 0), out[2] / out[3] the fraction of 16-entry lines of material / texture IDs
 * holding one ID (0 when cache_lines is 0). Device arrays of n IDs; out is host memory. */
wf_status wf_coherence_stats(wf_context* ctx, const int32_t* material_ids,
                             const int32_t* texture_ids, int32_t n, int32_t cache_lines,
                             double out[4]);

/* ---- shading --------------------------------------------------------------------------------- */

typedef struct wf_shade_params {
    int32_t physics;           /* WF_PHYSICS_* */
    int32_t rr_min_depth;      /* reference: always 3 (constructor argument ignored) */
    int32_t cost_aware_rr;     /* enableCostAwareRR */
    float firefly_threshold;   /* reference default 10 */
    uint32_t seed;
    int32_t sample_index;      /* RNG counter: sample pass */
    int32_t job;               /* RNG counter: job tag (0 in the plain form) */
} wf_shade_params;

/* WavefrontRenderer::shadeAll (cpp:569-877) for one bounce: emission+MIS and NEE into the
 * per-pixel accumulator (float RGB, pixel-indexed), shadow rays resolved with any-hit, scatter and
 * Russian roulette, and the next RayQueue appended in shading order. n_next returned. */
wf_status wf_shade(const wf_scene* scene, const wf_shade_soa* queue, int32_t n,
                   const wf_shade_params* params, float* accum_rgb, wf_ray_soa* next,
                   int32_t* n_next);
/* The miss loop (cpp:256-265) with getSkyColor (cpp:879-889): accum[pixel] += T * sky(dir)
 * (WF_PHYSICS_UNBIASED adds environment MIS for rays leaving non-delta bounces). */
wf_status wf_shade_misses(const wf_scene* scene, const wf_ray_soa* misses, int32_t n,
                          int32_t physics, float* accum_rgb);

/* ---- whole-frame render (WavefrontRenderer::renderScene, cpp:94-399) --------------------- */

typedef struct wf_camera {        /* Camera(lookFrom, lookAt, up, vfov, w, h), include/core/Camera.h:13-33 */
    float eye[3];
    float target[3];
    float up[3];
    float vfov_deg;
} wf_camera;

typedef struct wf_render_params {
    int32_t width;
    int32_t height;
    int32_t spp;               /* samplesPerPixel (a budget when adaptive) */
    int32_t max_depth;
    int32_t policy;            /* WF_POLICY_* */
    int32_t ray_sort;          /* enableRaySort */
    int32_t adaptive;          /* enableAdaptiveSampling */
    float adaptive_threshold;  /* AdaptiveSampler relative threshold (reference default 0.05) */
    int32_t adaptive_min_samples;   /* reference: 32 */
    int32_t adaptive_interval;      /* reference: 16 */
    int32_t schedule;          /* WF_SCHEDULE_*: same image, different order of work */
    int32_t sample_lanes;      /* REGENERATE, non-adaptive: samples of a pixel in flight at once
                                  (a power of two <= 64; 0 = 64); see
                                  wf_engine_config::sample_lanes */
    int32_t wavefront;         /* REGENERATE: ray slots (paths in flight); 0 = automatic */
    int32_t coherence_stats;   /* WF_COHERENCE_*: the reference's coherence metrics of every
                                  shading batch (cpp:187-246), returned in wf_render_stats */
    int32_t first_sample;      /* RNG index of the first sample: the render runs samples
                                  first_sample .. first_sample + spp - 1, so consecutive calls
                                  continue one sequence (progressive previews); 0 normally */
    wf_shade_params shade;
    wf_camera camera;
} wf_render_params;

/* wf_render_params::coherence_stats. RUN_LENGTH: average material and texture run lengths of
 * every shading batch in shading order, over hits with a texture (computeAverageRunLength,
 * cpp:67-92), weighted by batch size (cpp:187-199). CACHE_LINES adds the fraction of 16-int
 * (64-byte) lines of the batch's material and texture IDs holding one ID (enableMemoryCoherenceStats,
 * cpp:201-246). A batch is a (pass, depth) shading queue of the pass loop, or one iteration's
 * shading queue of the regenerating schedule. */
#define WF_COHERENCE_OFF 0
#define WF_COHERENCE_RUN_LENGTH 1
#define WF_COHERENCE_CACHE_LINES 2

typedef struct wf_render_stats {  /* the reference's stats blocks (cpp:321-370) plus ray counts */
    int64_t rays_traced;       /* sum over bounces of rays entering the bounce */
    int64_t shadow_rays;
    int64_t shaded_hits;       /* "Total shaded hits" */
    int64_t misses;
    int32_t samples_run;       /* "Samples actually run" */
    int32_t active_pixels_end; /* "Pixels still active at end" */
    double frame_ms;           /* device time raygen .. final resolve */
    double sort_ms;            /* ray sort + materialize */
    double intersect_ms;       /* traversal + hit/miss split (regenerating schedule: with the
                                * previous bounce's shadow rays, traced in the same launch) */
    double schedule_ms;        /* shading-queue policy sort + materialize */
    double shade_ms;           /* shade + shadow resolve (regenerating schedule: shade only,
                                * its shadow rays count under intersect_ms) */
    double miss_ms;
    /* with coherence_stats (else 0): "Average material run length", "Average texture run
     * length", "Material ID cache line homogeneity", "Texture  ID cache line homogeneity" */
    double material_run_length;
    double texture_run_length;
    double material_homogeneity;
    double texture_homogeneity;
} wf_render_stats;

/* wf_render_params::schedule. REGENERATE keeps a full wavefront: a pixel starts its next sample
 * as soon as a previous path of it ends. With sample_lanes = 1 a pixel has one path in flight,
 * so its accumulator receives the same additions in the same order as the pass loop and the
 * image is bit-identical to PASSES; with L lanes it has L, each summing its own samples, and the
 * image differs from the pass loop only by the order of the float additions. PASSES is the
 * reference's loop: all pixels sample pass by pass, bounce by bounce (cpp:141-312). */
#define WF_SCHEDULE_REGENERATE 0
#define WF_SCHEDULE_PASSES 1

/* Phase times in wf_render_stats (sort/intersect/schedule/shade/miss) are measured while kernel
 * timing is enabled (wf_profile_enable); otherwise they are 0 and the regenerating schedule
 * replays each iteration from a captured HIP graph. */

/* Renders one image. image_rgb: caller-owned device float buffer width*height*3 receiving
 * accumulator / per-pixel sample count (writeAveragedImage, cpp:127-137). sample_counts
 * (optional device int32[width*height]) receives the per-pixel sample counts. */
wf_status wf_render(wf_scene* scene, const wf_render_params* params, float* image_rgb,
                    int32_t* sample_counts, wf_render_stats* stats);

/* ---- batched and served forms: the job-tagged wavefront engine ---------------------------
 * No reference counterpart (the reference renders one image per process, main.cpp:149-215).
 * An engine keeps one scene's wavefront resident: a ray queue of `capacity` slots shared by the
 * paths of every admitted job, and framebuffers for up to `pixel_budget` pixels. Jobs are
 * admitted while a job slot and a contiguous framebuffer range are free; each job's image is
 * identical to wf_render of the same job with the same seed, job tag and sample lanes. */
typedef struct wf_engine wf_engine;

typedef struct wf_engine_config {
    int32_t capacity;          /* ray slots (paths in flight) */
    int32_t max_jobs;          /* job slots (<= 2048) */
    int64_t pixel_budget;      /* framebuffer pixels across admitted jobs */
    int32_t policy;            /* WF_POLICY_* shading-queue order */
    int32_t ray_sort;          /* RayQueue Morton sort of every iteration's rays */
    int32_t use_graph;         /* replay each iteration from a captured HIP graph */
    int32_t sample_lanes;      /* samples of a pixel in flight at once (a power of two <= 64;
                                  0 = 64).
                                  Lane l of a non-adaptive job runs samples l, l+L, l+2L, ... into
                                  an accumulator of its own; the lanes are summed in lane order
                                  when the job resolves, so an image depends on L but not on the
                                  schedule. A job uses the largest power of two <= min(L, spp)
                                  lanes; adaptive jobs use one (the estimate is sequential). */
    wf_shade_params shade;     /* physics, RR and firefly settings (seed/sample/job unused) */
} wf_engine_config;

typedef struct wf_job_desc {
    int32_t width;
    int32_t height;
    int32_t spp;               /* sample budget per pixel */
    int32_t max_depth;
    uint32_t seed;
    int32_t job_tag;           /* RNG job id: equal (seed, job_tag) give equal images */
    int32_t adaptive;
    float adaptive_threshold;
    int32_t adaptive_min_samples;
    int32_t adaptive_interval;
    int32_t first_sample;      /* RNG index of the job's first sample (see wf_render_params) */
    wf_camera camera;
} wf_job_desc;

typedef struct wf_job_result {
    int32_t width;
    int32_t height;
    int32_t samples_run;       /* most samples any pixel ran */
    int32_t active_pixels_end; /* adaptive: pixels that used the whole budget unconverged */
} wf_job_result;

typedef struct wf_engine_stats {
    int64_t rays_traced;
    int64_t shadow_rays;
    int64_t shaded_hits;
    int64_t misses;
    int64_t paths;             /* samples started */
    int64_t iterations;
    int64_t queued_pixels;     /* pixels waiting to start a sample */
    int32_t live_rays;
    int32_t running_jobs;
    int32_t free_pixels;
    int32_t reserved;
    double phase_ms[5];        /* sort, intersect, schedule, shade, miss (with phase timing;
                                * intersect holds the previous bounce's shadow rays) */
} wf_engine_stats;

/* Device bytes an engine of this capacity, pixel budget and sample lanes (0 = 64) needs
 * (admission planning). */
wf_status wf_engine_bytes(int32_t capacity, int64_t pixel_budget, int32_t sample_lanes,
                          size_t* out);
/* The engine runs on a stream of its own on ctx's device; it waits for work queued on the
 * scene's context before each run. */
wf_status wf_engine_create(wf_context* ctx, wf_scene* scene, const wf_engine_config* config,
                           wf_engine** out);
wf_status wf_engine_destroy(wf_engine* engine);
wf_status wf_engine_device_bytes(const wf_engine* engine, size_t* out);
/* Admits a job; WF_ERR_ADMISSION when no job slot or framebuffer range is free. */
wf_status wf_engine_admit(wf_engine* engine, const wf_job_desc* job, int32_t* job_id);
/* Runs iterations until at least one job finishes, max_iterations (<= 0: no limit) have run or
 * nothing is left to do; the finished job ids (up to finished_cap) go to `finished`. */
wf_status wf_engine_run(wf_engine* engine, int32_t max_iterations, int32_t* finished,
                        int32_t finished_cap, int32_t* n_finished);
/* Non-blocking form for driving several engines (one per resident scene, each on its own stream)
 * from one host thread: launch enqueues `iterations` iterations and returns at once; wait blocks
 * until they ran and reports the jobs finished by then. wf_engine_run may not be called between
 * them. */
wf_status wf_engine_launch(wf_engine* engine, int32_t iterations);
wf_status wf_engine_wait(wf_engine* engine, int32_t* finished, int32_t finished_cap,
                         int32_t* n_finished);
/* Resolves a finished job into caller-owned device buffers (image_rgb: width*height*3 floats;
 * sample_counts optional) and frees its slot and framebuffer range. */
wf_status wf_engine_result(wf_engine* engine, int32_t job_id, float* image_rgb,
                           int32_t* sample_counts, wf_job_result* result);
wf_status wf_engine_query(wf_engine* engine, wf_engine_stats* out);
/* Per-phase event timing of every iteration (disables graph replay while on). */
wf_status wf_engine_phase_timing(wf_engine* engine, int32_t enable);
/* The shading queue of the engine's last iteration, in the order it was shaded: the order a
 * shading policy produced (ShadingQueue::schedule + materialize, include/scheduling/
 * ShadingQueue.h:233-275, applied inside the engine), and the queue the reference's coherence
 * statistics are computed from (src/render/WavefrontRenderer.cpp:187-246). Copies
 * min(capacity, length) entries into the non-NULL device columns of dst (columns the engine does
 * not keep -- ray origin, tri, geom -- are left alone); *n receives the length. Waits for the
 * engine's stream. */
wf_status wf_engine_shading_queue(wf_engine* engine, wf_shade_soa* dst, int32_t capacity,
                                  int32_t* n);

/* ---- launch tuning (the autotuner's knobs; results never depend on them) ------------------ */
#define WF_BVH_PLOC 0   /* locally-ordered clustering of the Morton-sorted triangles */
#define WF_BVH_LBVH 1   /* Karras radix tree over the Morton codes */
typedef struct wf_tuning {
    int32_t refill_lanes;        /* persistent traversal: idle lanes of a wave before it claims new
                                  * rays (1..64; 64 = only when the whole wave is idle) */
    int32_t trace_blocks_per_cu; /* persistent traversal grid: 128-lane blocks per CU (1..64) */
    int32_t bvh_builder;         /* WF_BVH_* hierarchy of the next wf_accel_build */
    int32_t ploc_radius;         /* PLOC neighbour search radius (1..32) */
    int32_t bvh_width;           /* node width of the next wf_accel_build: 4 (default), float
                                  * boxes; 8, boxes quantised outward to 8 bits per node grid
                                  * (fewer, smaller nodes; more box tests per node) */
    int32_t radix_bits;          /* digit width of the radix sort passes: 6, 8 (default) or 11 */
    int32_t shadow_blocks_per_cu;   /* the shadow-ray kernel's grid (0 = trace_blocks_per_cu): it
                                     * traces few rays per lane, so its shape is tuned apart */
    int32_t shadow_refill_lanes;    /* the shadow-ray kernel's refill threshold (0 = refill_lanes) */
    int32_t ray_key_bits;        /* regenerating wavefront: how many top bits of RayQueue::sortKey's
                                  * 33-bit key (octant, then Morton from the top) order the rays of
                                  * each bounce (1..33); the queue order there is free, so a
                                  * shorter key trades coherence for fewer sort passes.
                                  * wf_ray_schedule always sorts by the full key */
} wf_tuning;
/* Engines created on the context take its tuning at every run. */
wf_status wf_context_set_tuning(wf_context* ctx, const wf_tuning* tuning);
wf_status wf_context_get_tuning(wf_context* ctx, wf_tuning* tuning);

/* ---- per-kernel device timing (HIP events on the context stream) ------------------------ */

/* Enables/disables event timing around every library kernel launch and clears the totals. */
wf_status wf_profile_enable(wf_context* ctx, int32_t enable);
/* Totals for a kernel name ("ray_keys", "ray_sort", "traverse_closest", "split", ...):
 * summed device milliseconds, launch count and algorithmic HBM bytes accounted to it. */
wf_status wf_profile_query(wf_context* ctx, const char* kernel, double* total_ms,
                           int64_t* launches, double* algorithmic_bytes);
/* Newline-separated list of profiled kernel names into buf (truncated to cap). */
wf_status wf_profile_names(wf_context* ctx, char* buf, int32_t cap);
/* Restricts timing to one kernel name (NULL or "" = every kernel), so that a timed region can
 * carry events around its dominant kernel only. Clears the totals. */
wf_status wf_profile_filter(wf_context* ctx, const char* kernel);

#ifdef __cplusplus
}
#endif

#endif /* WAVEFRONT_HIP_H */
