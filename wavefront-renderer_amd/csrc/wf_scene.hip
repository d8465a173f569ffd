// Scene upload (Scene::addMesh/addMaterial/addTexture, src/core/Scene.cpp:5-56) and the
// accelerator build that replaces EmbreeAccelerator::build (src/accelerator/Embree.cpp:66-119):
// a GPU linear BVH (Karras, "Maximizing parallelism in the construction of BVHs, octrees and
// k-d trees", HPG 2012) over the flattened triangle soup.
//
// Build: padded triangle boxes + centroid bounds -> 48-bit Morton codes of the centroids ->
// stable radix sort -> Karras hierarchy (one triangle per leaf, ties of equal codes split by
// index) -> bottom-up box refit (second thread to arrive at a node merges its children) ->
// collapse into four-wide nodes (every even-depth node adopts its grandchildren; 128 B, child
// boxes in SoA form) -> the exact bound on the traversal stack. Triangles are stored in leaf
// order as v0/e1/e2 (48 B, v0.w = the global triangle index), exactly the edges
// Moller-Trumbore recomputes.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

#include "wf_device.h"
#include "wf_internal.h"
#include "wf_traverse.h"

namespace wf {

wf_status scene_alloc(wf_scene* s, size_t bytes, void** out) {
    *out = nullptr;
    if (bytes == 0) return WF_OK;
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        *out = nullptr;
        return set_error(WF_ERR_OUT_OF_MEMORY, "scene allocation of %zu bytes failed: %s", bytes,
                         hipGetErrorString(e));
    }
    s->allocations.push_back(*out);
    s->device_bytes += bytes;
    return WF_OK;
}

// ---- host-side float math in the reference's order (for scene-level constants) ---------------
namespace {
struct H3 {
    float x, y, z;
    H3 operator-(const H3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    H3 cross(const H3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
    float length() const { return std::sqrt(x * x + y * y + z * z); }
};
inline float hlum(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

void build_alias(const std::vector<double>& w, std::vector<float>& prob,
                 std::vector<int32_t>& alias) {
    // Walker/Vose alias table: prob[i] keeps outcome i, otherwise alias[i].
    const size_t n = w.size();
    prob.assign(n, 1.0f);
    alias.resize(n);
    std::iota(alias.begin(), alias.end(), 0);
    double total = 0.0;
    for (double x : w) total += x;
    if (n == 0 || !(total > 0.0)) return;
    std::vector<double> q(n);
    std::vector<size_t> small, large;
    for (size_t i = 0; i < n; ++i) {
        q[i] = w[i] * (double)n / total;
        (q[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        size_t si = small.back();
        small.pop_back();
        size_t l = large.back();
        prob[si] = (float)q[si];
        alias[si] = (int32_t)l;
        q[l] = (q[l] + q[si]) - 1.0;
        if (q[l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (size_t i : large) prob[i] = 1.0f;
    for (size_t i : small) prob[i] = 1.0f;
}

int bits_for(int max_value) {
    int b = 0;
    while ((1 << b) <= max_value) ++b;
    return b;
}
}  // namespace

// Dense rank of per-material relative costs, descending (equal costs share a rank).
wf_status scene_upload_costs(wf_scene* s, const double* mat_cost, const double* tex_cost) {
    auto& d = s->dev;
    const int nm = d.n_materials, nt = d.n_textures;
    std::vector<double> mc(nm, 1.0), tc(nt, 1.0);
    if (mat_cost) std::copy(mat_cost, mat_cost + nm, mc.begin());
    if (tex_cost) std::copy(tex_cost, tex_cost + nt, tc.begin());
    std::vector<double> uniq = mc;
    std::sort(uniq.begin(), uniq.end(), [](double a, double b) { return a > b; });
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    std::vector<int32_t> rank(nm);
    for (int m = 0; m < nm; ++m)
        rank[m] = (int32_t)(std::find(uniq.begin(), uniq.end(), mc[m]) - uniq.begin());
    d.cost_rank_bits = uniq.size() <= 1 ? 0 : bits_for((int)uniq.size() - 1);
    if (nm) {
        WF_HIP(hipMemcpyAsync(d.mat_cost_rank, rank.data(), sizeof(int32_t) * nm,
                              hipMemcpyHostToDevice, s->ctx->stream));
        WF_HIP(hipMemcpyAsync(d.mat_rel_cost, mc.data(), sizeof(double) * nm,
                              hipMemcpyHostToDevice, s->ctx->stream));
    }
    if (nt)
        WF_HIP(hipMemcpyAsync(d.tex_rel_cost, tc.data(), sizeof(double) * nt,
                              hipMemcpyHostToDevice, s->ctx->stream));
    WF_HIP(hipStreamSynchronize(s->ctx->stream));
    return WF_OK;
}

// ---- BVH build kernels -------------------------------------------------------------------------
__device__ __forceinline__ uint32_t float_to_ordered(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(uint32_t u) {
    uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __uint_as_float(b);
}

__global__ void k_bvh_prep(const float4* __restrict__ v0, const float4* __restrict__ e1,
                           const float4* __restrict__ e2, int n, float4* __restrict__ lo,
                           float4* __restrict__ hi, uint32_t* __restrict__ cbounds) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    V3 a = xyz(v0[g]);
    V3 b = a + xyz(e1[g]);
    V3 c = a + xyz(e2[g]);
    float l[3], h[3], cen[3];
    const float pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z}, pc[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float mn = fminf(pa[k], fminf(pb[k], pc[k]));
        float mx = fmaxf(pa[k], fmaxf(pb[k], pc[k]));
        // padding: the box must contain every point Moller-Trumbore can accept
        float pad = (fabsf(mn) + fabsf(mx) + (mx - mn)) * 4e-6f + 1e-30f;
        l[k] = mn - pad;
        h[k] = mx + pad;
        cen[k] = 0.5f * (mn + mx);
    }
    lo[g] = make_float4(l[0], l[1], l[2], 0.0f);
    hi[g] = make_float4(h[0], h[1], h[2], 0.0f);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        atomicMin(&cbounds[k], float_to_ordered(cen[k]));
        atomicMax(&cbounds[3 + k], float_to_ordered(cen[k]));
    }
}

__device__ __forceinline__ uint64_t spread16(uint64_t x) {
    x &= 0xFFFFull;
    x = (x | (x << 16)) & 0x0000FF0000FFull;
    x = (x | (x << 8)) & 0x00F00F00F00Full;
    x = (x | (x << 4)) & 0x0C30C30C30C3ull;
    x = (x | (x << 2)) & 0x249249249249ull;
    return x;
}

__global__ void k_bvh_morton(const float4* __restrict__ lo, const float4* __restrict__ hi, int n,
                             const uint32_t* __restrict__ cbounds, uint64_t* __restrict__ keys) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    float4 l = lo[g], h = hi[g];
    float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
    uint64_t key = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float mn = ordered_to_float(cbounds[k]), mx = ordered_to_float(cbounds[3 + k]);
        float ext = mx - mn;
        float t = ext > 0.0f ? (c[k] - mn) / ext : 0.0f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        uint32_t q = (uint32_t)(t * 65535.0f);
        key |= spread16(q) << k;
    }
    keys[g] = key;
}

__global__ void k_bvh_gather_tris(const int32_t* __restrict__ order, int n,
                                  const float4* __restrict__ v0, const float4* __restrict__ e1,
                                  const float4* __restrict__ e2, const float4* __restrict__ lo,
                                  const float4* __restrict__ hi,
                                  const int32_t* __restrict__ tri_geom,
                                  const float* __restrict__ tri_uv, float4* __restrict__ tri,
                                  float4* __restrict__ uv_leaf, float4* __restrict__ slo,
                                  float4* __restrict__ shi,
                                  const int32_t* __restrict__ mesh_material,
                                  int32_t* __restrict__ leaf_mat) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int g = order[j];
    float4 a = v0[g];
    a.w = __int_as_float(g);
    float4 b = e1[g];
    const int32_t geom = tri_geom[g];
    b.w = __int_as_float(geom);
    leaf_mat[j] = mesh_material[geom];
    tri[3 * (size_t)j] = a;
    tri[3 * (size_t)j + 1] = b;
    tri[3 * (size_t)j + 2] = e2[g];
    const float* uv = tri_uv + 6 * (size_t)g;
    uv_leaf[2 * (size_t)j] = make_float4(uv[0], uv[1], uv[2], uv[3]);
    uv_leaf[2 * (size_t)j + 1] = make_float4(uv[4], uv[5], 0.0f, 0.0f);
    slo[j] = lo[g];
    shi[j] = hi[g];
}

__device__ __forceinline__ int karras_delta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

__global__ void k_bvh_karras(const uint64_t* __restrict__ keys, int n, int32_t* __restrict__ left,
                             int32_t* __restrict__ right, int32_t* __restrict__ iparent,
                             int32_t* __restrict__ lparent) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (karras_delta(keys, n, i, i + 1) - karras_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = karras_delta(keys, n, i, i - d);
    int lmax = 2;
    while (karras_delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (karras_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = karras_delta(keys, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        int t = (l + div - 1) / div;
        if (karras_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    int gamma = i + s * d + (d < 0 ? -1 : 0);
    int lo_i = i < j ? i : j, hi_i = i < j ? j : i;
    int lc = (lo_i == gamma) ? ~gamma : gamma;
    int rc = (hi_i == gamma + 1) ? ~(gamma + 1) : (gamma + 1);
    left[i] = lc;
    right[i] = rc;
    if (lc < 0) lparent[~lc] = i; else iparent[lc] = i;
    if (rc < 0) lparent[~rc] = i; else iparent[rc] = i;
    if (i == 0) iparent[0] = -1;
}

__device__ __forceinline__ float4 coherent_load(const float4* p) {
    const uint32_t* q = (const uint32_t*)p;
    float4 r;
    r.x = __uint_as_float(__hip_atomic_load(q + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    r.y = __uint_as_float(__hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    r.z = __uint_as_float(__hip_atomic_load(q + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    r.w = 0.0f;
    return r;
}

__device__ __forceinline__ void coherent_store(float4* p, float4 v) {
    uint32_t* q = (uint32_t*)p;
    __hip_atomic_store(q + 0, __float_as_uint(v.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, __float_as_uint(v.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 2, __float_as_uint(v.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bottom-up refit: each leaf climbs; the second arrival at a node merges both children.
__global__ void k_bvh_refit(int n, const int32_t* __restrict__ left,
                            const int32_t* __restrict__ right, const int32_t* __restrict__ iparent,
                            const int32_t* __restrict__ lparent, const float4* __restrict__ slo,
                            const float4* __restrict__ shi, float4* nlo, float4* nhi,
                            uint32_t* visits) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int p = lparent[j];
    while (p >= 0) {
        __threadfence();
        uint32_t old = atomicAdd(&visits[p], 1u);
        if (old == 0) return;  // first arrival: the sibling subtree is not finished yet
        __threadfence();
        int c[2] = {left[p], right[p]};
        float4 l2[2], h2[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (c[k] < 0) {
                l2[k] = slo[~c[k]];
                h2[k] = shi[~c[k]];
            } else {
                l2[k] = coherent_load(&nlo[c[k]]);
                h2[k] = coherent_load(&nhi[c[k]]);
            }
        }
        float4 L = make_float4(fminf(l2[0].x, l2[1].x), fminf(l2[0].y, l2[1].y),
                               fminf(l2[0].z, l2[1].z), 0.0f);
        float4 H = make_float4(fmaxf(h2[0].x, h2[1].x), fmaxf(h2[0].y, h2[1].y),
                               fmaxf(h2[0].z, h2[1].z), 0.0f);
        coherent_store(&nlo[p], L);
        coherent_store(&nhi[p], H);
        p = iparent[p];
    }
}

// Depth of each leaf = number of internal ancestors; max over leaves.
__global__ void k_bvh_depth(int n, const int32_t* __restrict__ iparent,
                            const int32_t* __restrict__ lparent, int32_t* max_depth) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    // a valid hierarchy has fewer than n ancestors per leaf: the bound turns a malformed tree
    // into a reported error instead of an endless walk
    int d = 0;
    int p = lparent[j];
    while (p >= 0 && d < n) {
        ++d;
        p = (p < n - 1) ? iparent[p] : -1;
    }
    if (p >= 0 || lparent[j] >= n - 1) d = n;
    atomicMax(max_depth, d);
}

__global__ void k_bvh4_single(const float4* __restrict__ slo, const float4* __restrict__ shi,
                              Bvh4Node* __restrict__ out) {
    float4 l = slo[0], h = shi[0];
    const float inf = __int_as_float(0x7F800000);   // empty slots: the inverted infinite box
    Bvh4Node nd;
    nd.lox = make_float4(l.x, inf, inf, inf);
    nd.loy = make_float4(l.y, inf, inf, inf);
    nd.loz = make_float4(l.z, inf, inf, inf);
    nd.hix = make_float4(h.x, -inf, -inf, -inf);
    nd.hiy = make_float4(h.y, -inf, -inf, -inf);
    nd.hiz = make_float4(h.z, -inf, -inf, -inf);
    nd.child = make_int4(~0, kEmptyChild, kEmptyChild, kEmptyChild);
    nd.pad = make_int4(0, 0, 0, 0);
    out[0] = nd;
}

}  // namespace wf

using namespace wf;

extern "C" {

wf_status wf_scene_create(wf_context* ctx, const wf_scene_desc* desc, wf_scene** out) {
    clear_error();
    WF_REQUIRE(ctx && desc && out, "wf_scene_create: NULL argument");
    WF_REQUIRE(desc->mesh_count >= 0 && desc->material_count >= 0 && desc->texture_count >= 0,
               "wf_scene_create: negative count");
    WF_REQUIRE(desc->mesh_count == 0 || desc->meshes, "wf_scene_create: meshes is NULL");
    WF_REQUIRE(desc->material_count == 0 || desc->materials, "wf_scene_create: materials is NULL");
    WF_REQUIRE(desc->texture_count == 0 || desc->textures, "wf_scene_create: textures is NULL");
    for (int m = 0; m < desc->mesh_count; ++m) {
        const wf_mesh_desc& md = desc->meshes[m];
        WF_REQUIRE(md.triangle_count >= 0 && md.vertex_count >= 0, "mesh %d: negative count", m);
        WF_REQUIRE(md.triangle_count == 0 || (md.positions && md.indices),
                   "mesh %d: NULL positions/indices", m);
        WF_REQUIRE(md.material_id >= 0 && md.material_id < desc->material_count,
                   "mesh %d: material_id %d out of range [0,%d)", m, md.material_id,
                   desc->material_count);
        for (int k = 0; k < md.triangle_count * 3; ++k)
            WF_REQUIRE(md.indices[k] >= 0 && md.indices[k] < md.vertex_count,
                       "mesh %d: index %d out of range", m, md.indices[k]);
    }
    for (int t = 0; t < desc->texture_count; ++t) {
        const wf_texture_desc& td = desc->textures[t];
        WF_REQUIRE(td.width > 0 && td.height > 0 && td.data, "texture %d: empty", t);
        WF_REQUIRE(td.format == WF_TEX_RGBA8_SRGB || td.format == WF_TEX_RGB_F32,
                   "texture %d: unknown format %d", t, td.format);
    }
    WF_HIP(hipSetDevice(ctx->device));
    auto* s = new wf_scene();
    s->ctx = ctx;
    s->device = ctx->device;
    auto fail = [&](wf_status st) {
        for (void* p : s->allocations) (void)hipFree(p);
        delete s;
        return st;
    };
    SceneDev& d = s->dev;
    // flatten (geomID = mesh order; global triangle index = mesh offset + primID)
    std::vector<float4> v0, e1, e2;
    std::vector<int32_t> tri_geom, mesh_off, mesh_cnt, mesh_mat, mesh_uv;
    std::vector<float> tri_uv;
    float bmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    int nverts = 0;
    for (int m = 0; m < desc->mesh_count; ++m) {
        const wf_mesh_desc& md = desc->meshes[m];
        mesh_off.push_back((int32_t)v0.size());
        mesh_cnt.push_back(md.triangle_count);
        mesh_mat.push_back(md.material_id);
        bool has_uv = md.uv_u && md.uv_v;
        mesh_uv.push_back(has_uv ? 1 : 0);
        nverts += md.vertex_count;
        for (int v = 0; v < md.vertex_count; ++v)
            for (int k = 0; k < 3; ++k) {
                bmin[k] = std::min(bmin[k], md.positions[v * 3 + k]);
                bmax[k] = std::max(bmax[k], md.positions[v * 3 + k]);
            }
        for (int k = 0; k < md.triangle_count; ++k) {
            const float* p = md.positions;
            int i0 = md.indices[k * 3], i1 = md.indices[k * 3 + 1], i2 = md.indices[k * 3 + 2];
            H3 a{p[i0 * 3], p[i0 * 3 + 1], p[i0 * 3 + 2]};
            H3 b{p[i1 * 3], p[i1 * 3 + 1], p[i1 * 3 + 2]};
            H3 c{p[i2 * 3], p[i2 * 3 + 1], p[i2 * 3 + 2]};
            H3 x = b - a, y = c - a;
            v0.push_back(make_float4(a.x, a.y, a.z, 0.0f));
            e1.push_back(make_float4(x.x, x.y, x.z, 0.0f));
            e2.push_back(make_float4(y.x, y.y, y.z, 0.0f));
            tri_geom.push_back(m);
            for (int cidx = 0; cidx < 3; ++cidx) {
                tri_uv.push_back(has_uv ? md.uv_u[k * 3 + cidx] : 0.0f);
                tri_uv.push_back(has_uv ? md.uv_v[k * 3 + cidx] : 0.0f);
            }
        }
    }
    d.n_tris = (int32_t)v0.size();
    d.n_meshes = desc->mesh_count;
    d.n_materials = desc->material_count;
    d.n_textures = desc->texture_count;
    d.n_vertices = nverts;
    for (int k = 0; k < 3; ++k) {
        d.bounds_min[k] = bmin[k];
        d.bounds_max[k] = bmax[k];
    }
    wf_status st = WF_OK;
#define UP(host, count, dst)                                                   \
    do {                                                                       \
        st = scene_upload(s, (host), (count), &(dst));                         \
        if (st != WF_OK) return fail(st);                                      \
    } while (0)
    UP(v0.data(), v0.size(), d.tri_v0);
    UP(e1.data(), e1.size(), d.tri_e1);
    UP(e2.data(), e2.size(), d.tri_e2);
    UP(tri_geom.data(), tri_geom.size(), d.tri_geom);
    UP(tri_uv.data(), tri_uv.size(), d.tri_uv);
    UP(mesh_off.data(), mesh_off.size(), d.mesh_tri_offset);
    UP(mesh_cnt.data(), mesh_cnt.size(), d.mesh_tri_count);
    UP(mesh_mat.data(), mesh_mat.size(), d.mesh_material);
    UP(mesh_uv.data(), mesh_uv.size(), d.mesh_has_uv);

    // materials
    s->materials.assign(desc->materials, desc->materials + desc->material_count);
    // emitters: meshes whose material is Emissive/SpotLight (Scene.cpp:84-90); total area is a
    // float sum in mesh then triangle order (Scene.cpp:162-190)
    std::vector<int32_t> emissive, emissive_base, light_mat;
    std::vector<float4> lv0, lv1, lv2;
    float total_area = 0.0f;
    std::vector<int32_t> light_tris;
    std::vector<double> light_w;
    double light_total = 0.0;
    for (int m = 0; m < desc->mesh_count; ++m) {
        const wf_material& mt = s->materials[mesh_mat[m]];
        if (mt.type != WF_MAT_EMISSIVE && mt.type != WF_MAT_SPOTLIGHT) continue;
        emissive.push_back(m);
        emissive_base.push_back((int32_t)light_tris.size());
        double lum = hlum(mt.albedo[0] * mt.emission, mt.albedo[1] * mt.emission,
                          mt.albedo[2] * mt.emission);
        for (int k = 0; k < mesh_cnt[m]; ++k) {
            int g = mesh_off[m] + k;
            H3 x{e1[g].x, e1[g].y, e1[g].z}, y{e2[g].x, e2[g].y, e2[g].z};
            float area = x.cross(y).length() * 0.5f;
            total_area += area;
            light_tris.push_back(g);
            light_mat.push_back(mesh_mat[m]);
            {
                const wf_mesh_desc& md = desc->meshes[m];
                const float* p = md.positions;
                int i0 = md.indices[k * 3], i1 = md.indices[k * 3 + 1], i2 = md.indices[k * 3 + 2];
                lv0.push_back(make_float4(p[i0 * 3], p[i0 * 3 + 1], p[i0 * 3 + 2], 0.0f));
                lv1.push_back(make_float4(p[i1 * 3], p[i1 * 3 + 1], p[i1 * 3 + 2], 0.0f));
                lv2.push_back(make_float4(p[i2 * 3], p[i2 * 3 + 1], p[i2 * 3 + 2], 0.0f));
            }
            light_w.push_back((double)area * lum);
            light_total += (double)area * lum;
        }
    }
    d.total_emissive_area = total_area;
    d.n_emissive_meshes = (int32_t)emissive.size();
    UP(emissive.data(), emissive.size(), d.emissive_meshes);
    std::vector<float> lprob;
    std::vector<int32_t> lalias;
    build_alias(light_w, lprob, lalias);
    d.n_light_tris = (int32_t)light_tris.size();
    UP(light_tris.data(), light_tris.size(), d.light_tris);
    UP(lv0.data(), lv0.size(), d.light_v0);
    UP(lv1.data(), lv1.size(), d.light_v1);
    UP(lv2.data(), lv2.size(), d.light_v2);
    UP(light_mat.data(), light_mat.size(), d.light_mat);
    UP(emissive_base.data(), emissive_base.size(), d.emissive_light_base);
    UP(lprob.data(), lprob.size(), d.light_prob);
    UP(lalias.data(), lalias.size(), d.light_alias);

    // environment
    d.directional = desc->directional;
    const wf_environment_desc& env = desc->environment;
    if (env.width > 0 && env.height > 0 && env.rgb) {
        d.env_w = env.width;
        d.env_h = env.height;
        d.env_intensity = env.intensity;
        for (int k = 0; k < 3; ++k) d.env_tint[k] = env.tint[k];
        size_t np = (size_t)env.width * env.height;
        std::vector<float4> px(np);
        std::vector<double> ew(np);
        double etot = 0.0;
        for (int y = 0; y < env.height; ++y) {
            double c = std::cos((0.5 - (y + 0.5) / env.height) * M_PI);
            for (int x = 0; x < env.width; ++x) {
                size_t i = (size_t)y * env.width + x;
                px[i] = make_float4(env.rgb[i * 3], env.rgb[i * 3 + 1], env.rgb[i * 3 + 2], 0.0f);
                ew[i] = (double)hlum(env.rgb[i * 3], env.rgb[i * 3 + 1], env.rgb[i * 3 + 2]) * c;
                etot += ew[i];
            }
        }
        std::vector<float> eprob, epdf(np);
        std::vector<int32_t> ealias;
        build_alias(ew, eprob, ealias);
        for (size_t i = 0; i < np; ++i) epdf[i] = etot > 0.0 ? (float)(ew[i] / etot) : 0.0f;
        d.env_choice_prob = etot > 0.0 ? (light_tris.empty() ? 1.0f : 0.5f) : 0.0f;
        UP(px.data(), px.size(), d.env);
        UP(eprob.data(), eprob.size(), d.env_prob);
        UP(ealias.data(), ealias.size(), d.env_alias);
        UP(epdf.data(), epdf.size(), d.env_pdf);
    }

    std::vector<DevMaterial> dm(desc->material_count);
    for (int m = 0; m < desc->material_count; ++m) {
        const wf_material& w = s->materials[m];
        DevMaterial& x = dm[m];
        std::memset(&x, 0, sizeof(x));
        x.type = w.type;
        x.model = w.model;
        x.texture_id = w.texture_id;
        x.use_spatial_checker = w.use_spatial_checker;
        for (int k = 0; k < 3; ++k) x.albedo[k] = w.albedo[k];
        x.roughness = w.roughness;
        x.emission = w.emission;
        x.spot_outer_deg = w.spot_outer_deg;
        x.spot_falloff_deg = w.spot_falloff_deg;
        x.ior = w.ior;
        x.checker_cell_size = w.checker_cell_size;
        if ((w.type == WF_MAT_EMISSIVE || w.type == WF_MAT_SPOTLIGHT) && light_total > 0.0) {
            double lum = hlum(w.albedo[0] * w.emission, w.albedo[1] * w.emission,
                              w.albedo[2] * w.emission);
            x.light_pdf_area = (float)(lum / light_total);
        }
    }
    UP(dm.data(), dm.size(), d.materials);

    // textures: RGBA8 sRGB texels stay 8-bit and decode through a 256-entry table computed with
    // the reference's srgbToLinear (Texture.cpp:20-28); float textures become float4
    std::vector<DevTexture> dt(desc->texture_count);
    std::vector<uint32_t> t8;
    std::vector<float4> t32;
    s->tex_mem.resize(desc->texture_count);
    for (int t = 0; t < desc->texture_count; ++t) {
        const wf_texture_desc& td = desc->textures[t];
        size_t np = (size_t)td.width * td.height;
        dt[t].width = td.width;
        dt[t].height = td.height;
        dt[t].format = td.format;
        s->tex_mem[t] = (uint64_t)np * 12u;  // Texture::memoryBytes (vector<Color>)
        if (td.format == WF_TEX_RGBA8_SRGB) {
            dt[t].offset = (int64_t)t8.size();
            const uint32_t* src = (const uint32_t*)td.data;
            t8.insert(t8.end(), src, src + np);
        } else {
            dt[t].offset = (int64_t)t32.size();
            const float* f = (const float*)td.data;
            for (size_t i = 0; i < np; ++i) t32.push_back(make_float4(f[i * 3], f[i * 3 + 1], f[i * 3 + 2], 0.0f));
        }
    }
    UP(dt.data(), dt.size(), d.textures);
    UP(t8.data(), t8.size(), d.texels8);
    UP(t32.data(), t32.size(), d.texels32);
    std::vector<float> lut(256);
    for (int i = 0; i < 256; ++i) {
        float c = (float)i / 255.0f;
        lut[i] = c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
    }
    UP(lut.data(), lut.size(), d.srgb_lut);
    UP(s->tex_mem.data(), s->tex_mem.size(), d.tex_mem_bytes);
    // texture size rank (descending memoryBytes; an invalid id counts as 0 bytes)
    {
        std::vector<uint64_t> sizes(s->tex_mem);
        sizes.push_back(0);
        std::vector<uint64_t> uniq = sizes;
        std::sort(uniq.begin(), uniq.end(), [](uint64_t a, uint64_t b) { return a > b; });
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        std::vector<int32_t> rank(sizes.size());
        for (size_t i = 0; i < sizes.size(); ++i)
            rank[i] = (int32_t)(std::find(uniq.begin(), uniq.end(), sizes[i]) - uniq.begin());
        d.size_rank_bits = uniq.size() <= 1 ? 0 : bits_for((int)uniq.size() - 1);
        UP(rank.data(), rank.size(), d.tex_size_rank);
    }
#undef UP
    void* p = nullptr;
    if ((st = scene_alloc(s, sizeof(int32_t) * std::max(1, d.n_materials), &p)) != WF_OK) return fail(st);
    d.mat_cost_rank = (int32_t*)p;
    if ((st = scene_alloc(s, sizeof(double) * std::max(1, d.n_materials), &p)) != WF_OK) return fail(st);
    d.mat_rel_cost = (double*)p;
    if ((st = scene_alloc(s, sizeof(double) * std::max(1, d.n_textures), &p)) != WF_OK) return fail(st);
    d.tex_rel_cost = (double*)p;
    if ((st = scene_upload_costs(s, nullptr, nullptr)) != WF_OK) return fail(st);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(set_error(WF_ERR_HIP, "scene upload: %s", hipGetErrorString(e)));
    if ((st = check_fault(ctx)) != WF_OK) return fail(st);
    *out = s;
    return WF_OK;
}

wf_status wf_scene_destroy(wf_scene* s) {
    if (!s) return WF_OK;
    // the owning context may already be gone: synchronise the device, not the context stream
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    if (s->render_engine) wf_engine_destroy(s->render_engine);
    for (void* p : s->allocations) (void)hipFree(p);
    delete s;
    return WF_OK;
}

wf_status wf_scene_device_bytes(const wf_scene* s, size_t* out) {
    WF_REQUIRE(s && out, "wf_scene_device_bytes: NULL argument");
    *out = s->device_bytes;
    return WF_OK;
}

wf_status wf_scene_set_costs(wf_scene* s, const double* mat_cost, const double* tex_cost) {
    clear_error();
    WF_REQUIRE(s, "wf_scene_set_costs: NULL scene");
    return scene_upload_costs(s, mat_cost, tex_cost);
}

wf_status wf_accel_build(wf_scene* s, wf_build_stats* stats) {
    clear_error();
    WF_REQUIRE(s, "wf_accel_build: NULL scene");
    wf_context* ctx = s->ctx;
    SceneDev& d = s->dev;
    const int n = d.n_tris;
    hipEvent_t e0, e1;
    WF_HIP(hipEventCreate(&e0));
    WF_HIP(hipEventCreate(&e1));
    WF_HIP(hipEventRecord(e0, ctx->stream));
    if (n > 0) {
        void *plo, *phi, *pkeys, *pkeys2, *porder, *pcb, *pslo, *pshi;
        WF_TRY(workspace(ctx, "bvh_lo", sizeof(float4) * n, &plo));
        WF_TRY(workspace(ctx, "bvh_hi", sizeof(float4) * n, &phi));
        WF_TRY(workspace(ctx, "bvh_keys", sizeof(uint64_t) * n, &pkeys));
        WF_TRY(workspace(ctx, "bvh_keys2", sizeof(uint64_t) * n, &pkeys2));
        WF_TRY(workspace(ctx, "bvh_order", sizeof(int32_t) * n, &porder));
        WF_TRY(workspace(ctx, "bvh_cbounds", sizeof(uint32_t) * 8, &pcb));
        WF_TRY(workspace(ctx, "bvh_slo", sizeof(float4) * n, &pslo));
        WF_TRY(workspace(ctx, "bvh_shi", sizeof(float4) * n, &pshi));
        uint32_t cb_init[8] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u, 0u, 0u};
        WF_HIP(hipMemcpyAsync(pcb, cb_init, sizeof(cb_init), hipMemcpyHostToDevice, ctx->stream));
        k_bvh_prep<<<grid_for(n, 256), 256, 0, ctx->stream>>>(d.tri_v0, d.tri_e1, d.tri_e2, n,
                                                               (float4*)plo, (float4*)phi,
                                                               (uint32_t*)pcb);
        WF_HIP(hipGetLastError());
        k_bvh_morton<<<grid_for(n, 256), 256, 0, ctx->stream>>>((float4*)plo, (float4*)phi, n,
                                                                 (uint32_t*)pcb, (uint64_t*)pkeys);
        WF_HIP(hipGetLastError());
        WF_TRY(radix_sort_pairs(ctx, (uint64_t*)pkeys, nullptr, (uint64_t*)pkeys2,
                                (int32_t*)porder, host_count(n), 48, "bvh_sort"));
        if (!d.tri) {
            void* pt = nullptr;
            WF_TRY(scene_alloc(s, sizeof(float4) * 3 * (size_t)n, &pt));
            d.tri = (float4*)pt;
            WF_TRY(scene_alloc(s, sizeof(float4) * 2 * (size_t)n, &pt));
            d.tri_uv_leaf = (float4*)pt;
            WF_TRY(scene_alloc(s, sizeof(int32_t) * (size_t)n, &pt));
            d.leaf_mat = (int32_t*)pt;
        }
        k_bvh_gather_tris<<<grid_for(n, 256), 256, 0, ctx->stream>>>(
            (int32_t*)porder, n, d.tri_v0, d.tri_e1, d.tri_e2, (float4*)plo, (float4*)phi,
            d.tri_geom, d.tri_uv, d.tri, d.tri_uv_leaf, (float4*)pslo, (float4*)pshi,
            d.mesh_material, d.leaf_mat);
        WF_HIP(hipGetLastError());
        int32_t depth = 0;
        // a single triangle always gets the four-wide root; four-wide nodes are addressed by
        // 32-bit byte offsets, so a tree of more than 2^32 / 128 of them is built eight-wide
        const bool four_fits = (size_t)std::max(1, n - 1) * sizeof(Bvh4Node) <= ((size_t)1 << 32);
        d.width = (n > 1 && (ctx->tuning.bvh_width == 8 || !four_fits)) ? 8 : 4;
        if (d.width == 4 && !d.nodes4) {
            void* pn4 = nullptr;
            WF_TRY(scene_alloc(s, sizeof(Bvh4Node) * (size_t)std::max(1, n - 1), &pn4));
            d.nodes4 = (Bvh4Node*)pn4;
        }
        if (d.width == 8 && !d.nodes8) {
            void* pn8 = nullptr;
            WF_TRY(scene_alloc(s, sizeof(Bvh8Node) * (size_t)std::max(1, n - 1), &pn8));
            d.nodes8 = (Bvh8Node*)pn8;
        }
        if (n == 1) {
            k_bvh4_single<<<1, 1, 0, ctx->stream>>>((float4*)pslo, (float4*)pshi, d.nodes4);
            WF_HIP(hipGetLastError());
            d.n_nodes = 1;
            d.n_nodes4 = 1;
            d.stack_bound = 0;
            depth = 1;
        } else {
            void *pl, *pr, *pip, *plp, *pnlo, *pnhi, *pvis, *pdep;
            WF_TRY(workspace(ctx, "bvh_left", sizeof(int32_t) * n, &pl));
            WF_TRY(workspace(ctx, "bvh_right", sizeof(int32_t) * n, &pr));
            WF_TRY(workspace(ctx, "bvh_iparent", sizeof(int32_t) * n, &pip));
            WF_TRY(workspace(ctx, "bvh_lparent", sizeof(int32_t) * n, &plp));
            WF_TRY(workspace(ctx, "bvh_nlo", sizeof(float4) * n, &pnlo));
            WF_TRY(workspace(ctx, "bvh_nhi", sizeof(float4) * n, &pnhi));
            WF_TRY(workspace(ctx, "bvh_visits", sizeof(uint32_t) * n, &pvis));
            WF_TRY(workspace(ctx, "bvh_depth", sizeof(int32_t), &pdep));
            WF_TRY(wf_fill_u32(ctx, pvis, 0u, (size_t)n));
            WF_TRY(wf_fill_u32(ctx, pdep, 0u, 1));
            // binary hierarchy: PLOC clustering (default) or the Karras radix tree
            bool built = false;
            if (ctx->tuning.bvh_builder == WF_BVH_PLOC) {
                wf_status pst = ploc_build(ctx, n, ctx->tuning.ploc_radius, (float4*)pslo,
                                           (float4*)pshi, (int32_t*)pl, (int32_t*)pr,
                                           (int32_t*)pip, (int32_t*)plp, (float4*)pnlo,
                                           (float4*)pnhi);
                if (pst == WF_OK) built = true;
                else if (pst != WF_ERR_UNSUPPORTED) return pst;
                else clear_error();
            }
            if (!built) {
                k_bvh_karras<<<grid_for(n - 1, 256), 256, 0, ctx->stream>>>(
                    (uint64_t*)pkeys2, n, (int32_t*)pl, (int32_t*)pr, (int32_t*)pip, (int32_t*)plp);
                WF_HIP(hipGetLastError());
                k_bvh_refit<<<grid_for(n, 256), 256, 0, ctx->stream>>>(
                    n, (int32_t*)pl, (int32_t*)pr, (int32_t*)pip, (int32_t*)plp, (float4*)pslo,
                    (float4*)pshi, (float4*)pnlo, (float4*)pnhi, (uint32_t*)pvis);
                WF_HIP(hipGetLastError());
            }
            k_bvh_depth<<<grid_for(n, 256), 256, 0, ctx->stream>>>(n, (int32_t*)pip,
                                                                   (int32_t*)plp, (int32_t*)pdep);
            WF_HIP(hipGetLastError());
            WF_TRY(readback(ctx, &depth, pdep, sizeof(int32_t)));
            if (depth >= n) return set_error(WF_ERR_HIP, "BVH build produced a malformed hierarchy");
            d.n_nodes = n - 1;
            // surface-area-optimal collapse to four- or eight-wide nodes
            WF_TRY(collapse_wide(ctx, n, (int32_t*)pl, (int32_t*)pr, (int32_t*)pip,
                                 (int32_t*)plp, (float4*)pslo, (float4*)pshi, (float4*)pnlo,
                                 (float4*)pnhi, d.width == 4 ? d.nodes4 : nullptr,
                                 d.width == 8 ? d.nodes8 : nullptr, &d.n_nodes4,
                                 &d.stack_bound));
            if (d.width == 8 && d.stack_bound > kStackMax) {
                // an eight-wide node pushes up to seven entries: fall back to four-wide nodes
                // when the eight-wide stack bound exceeds the largest traversal instantiation
                if (!four_fits)
                    return set_error(WF_ERR_UNSUPPORTED,
                                     "BVH of %d triangles: eight-wide stack bound %d exceeds %d "
                                     "and four-wide nodes would exceed 4 GiB", n, d.stack_bound,
                                     kStackMax);
                d.width = 4;
                if (!d.nodes4) {
                    void* pn4 = nullptr;
                    WF_TRY(scene_alloc(s, sizeof(Bvh4Node) * (size_t)(n - 1), &pn4));
                    d.nodes4 = (Bvh4Node*)pn4;
                }
                WF_TRY(collapse_wide(ctx, n, (int32_t*)pl, (int32_t*)pr, (int32_t*)pip,
                                     (int32_t*)plp, (float4*)pslo, (float4*)pshi, (float4*)pnlo,
                                     (float4*)pnhi, d.nodes4, nullptr, &d.n_nodes4,
                                     &d.stack_bound));
            }
        }
        d.max_depth = depth;
    } else {
        d.n_nodes = 0;
        d.n_nodes4 = 0;
        d.stack_bound = 0;
        d.max_depth = 0;
    }
    WF_HIP(hipEventRecord(e1, ctx->stream));
    WF_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    WF_TRY(check_fault(ctx));
    d.built = true;
    if (stats) {
        stats->mesh_count = d.n_meshes;
        stats->triangle_count = d.n_tris;
        stats->vertex_count = d.n_vertices;
        stats->node_count = d.n_nodes4;
        stats->max_depth = d.max_depth;
        stats->stack_bound = d.stack_bound;
        stats->bvh_width = d.width;
        stats->build_ms = ms;
        for (int k = 0; k < 3; ++k) {
            stats->bounds_min[k] = d.bounds_min[k];
            stats->bounds_max[k] = d.bounds_max[k];
        }
    }
    return WF_OK;
}

}  // extern "C"
