// Ray queue access for kernels that run on either layout: the C-ABI's structure of twelve
// columns (wf_ray_soa) or the engine's three 16-byte columns (RayPacked). A gather or a scattered
// append of a packed ray moves three aligned 16-byte records instead of twelve 4-byte fields
// in twelve different cache lines.
#pragma once

#include "wf_device.h"
#include "wf_internal.h"

namespace wf {

struct RayVal {
    V3 o, d, t;
    int32_t pixel, depth;
    float pdf;
};

__device__ __forceinline__ void ld_od(const wf_ray_soa& r, int i, V3& o, V3& d) {
    o = V3(r.ox[i], r.oy[i], r.oz[i]);
    d = V3(r.dx[i], r.dy[i], r.dz[i]);
}
__device__ __forceinline__ void ld_od(const RayPacked& r, int i, V3& o, V3& d) {
    const float4 a = r.a[i], b = r.b[i];
    o = V3(a.x, a.y, a.z);
    d = V3(b.x, b.y, b.z);
}

__device__ __forceinline__ RayVal ld_ray(const wf_ray_soa& r, int i) {
    RayVal v;
    v.o = V3(r.ox[i], r.oy[i], r.oz[i]);
    v.d = V3(r.dx[i], r.dy[i], r.dz[i]);
    v.t = V3(r.tr[i], r.tg[i], r.tb[i]);
    v.pixel = r.pixel[i];
    v.depth = r.depth[i];
    v.pdf = r.pdf[i];
    return v;
}
__device__ __forceinline__ RayVal ld_ray(const RayPacked& r, int i) {
    const float4 a = r.a[i], b = r.b[i], c = r.c[i];
    RayVal v;
    v.o = V3(a.x, a.y, a.z);
    v.pixel = __float_as_int(a.w);
    v.d = V3(b.x, b.y, b.z);
    v.pdf = b.w;
    v.t = V3(c.x, c.y, c.z);
    v.depth = __float_as_int(c.w);
    return v;
}

__device__ __forceinline__ void ld_od(const RayPackedOrdered& r, int i, V3& o, V3& d) {
    const int j = r.order[i];
    const float4 a = r.a[j], b = r.b[j];
    o = V3(a.x, a.y, a.z);
    d = V3(b.x, b.y, b.z);
}
__device__ __forceinline__ RayVal ld_ray(const RayPackedOrdered& r, int i) {
    const int j = r.order[i];
    const float4 a = r.a[j], b = r.b[j], c = r.c[j];
    RayVal v;
    v.o = V3(a.x, a.y, a.z);
    v.pixel = __float_as_int(a.w);
    v.d = V3(b.x, b.y, b.z);
    v.pdf = b.w;
    v.t = V3(c.x, c.y, c.z);
    v.depth = __float_as_int(c.w);
    return v;
}

// Slot of ray i's per-ray side data (job tags): i, or its source slot in an ordered view.
__device__ __forceinline__ int ray_source(const wf_ray_soa&, int i) { return i; }
__device__ __forceinline__ int ray_source(const RayPacked&, int i) { return i; }
__device__ __forceinline__ int ray_source(const RayPackedOrdered& r, int i) { return r.order[i]; }

// A column queue without origin columns (ox NULL) stores the rest of the ray.
__device__ __forceinline__ void st_ray(const wf_ray_soa& r, int i, const RayVal& v) {
    if (r.ox) {
        r.ox[i] = v.o.x; r.oy[i] = v.o.y; r.oz[i] = v.o.z;
    }
    r.dx[i] = v.d.x; r.dy[i] = v.d.y; r.dz[i] = v.d.z;
    r.tr[i] = v.t.x; r.tg[i] = v.t.y; r.tb[i] = v.t.z;
    r.pixel[i] = v.pixel;
    r.depth[i] = v.depth;
    r.pdf[i] = v.pdf;
}
__device__ __forceinline__ void st_ray(const RayPacked& r, int i, const RayVal& v) {
    r.a[i] = make_float4(v.o.x, v.o.y, v.o.z, __int_as_float(v.pixel));
    r.b[i] = make_float4(v.d.x, v.d.y, v.d.z, v.pdf);
    r.c[i] = make_float4(v.t.x, v.t.y, v.t.z, __int_as_float(v.depth));
}

// HitRecord mapping of EmbreeAccelerator::intersect4 (Embree.cpp:240-252) for a ray (o, d) whose
// closest hit is leaf triangle L at (t, bu, bv): hit point, the triangle's outward normal turned
// against the ray (frontFace), UV (Mesh::interpolateUV, src/geometry/Mesh.cpp:46-63, from the
// UVs stored beside the leaf; the barycentrics without UVs), global triangle, mesh and material.
struct MappedHit {
    V3 P, N;
    float u, v;
    int32_t g, geom, mat;
    bool front;
};
__device__ __forceinline__ MappedHit map_hit_vals(const SceneDev& s, const V3& o, const V3& d,
                                                  int L, float t, float bu, float bv) {
    const float4 A = s.tri[3 * (size_t)L];
    const float4 B = s.tri[3 * (size_t)L + 1];
    const float4 C = s.tri[3 * (size_t)L + 2];
    MappedHit h;
    h.g = __float_as_int(A.w);
    h.P = o + d * t;
    h.geom = __float_as_int(B.w);
    h.mat = s.mesh_material[h.geom];
    h.u = bu;
    h.v = bv;
    if (s.mesh_has_uv[h.geom]) {
        const float4 ua = s.tri_uv_leaf[2 * (size_t)L], ub = s.tri_uv_leaf[2 * (size_t)L + 1];
        const float w0 = 1.0f - bu - bv;
        h.u = w0 * ua.x + bu * ua.z + bv * ub.x;
        h.v = w0 * ua.y + bu * ua.w + bv * ub.y;
    }
    const V3 outward = xyz(B).cross(xyz(C)).normalized();
    h.front = d.dot(outward) < 0.0f;
    h.N = h.front ? outward : outward * -1.0f;
    return h;
}

// Ray `e` of a hit source (its list entry; e itself for records already in shading order) and
// the slot of its side data.
__device__ __forceinline__ int hs_traced(const HitSource& h, int e) {
    return h.list ? h.list[e] : e;
}
__device__ __forceinline__ int hs_slot(const HitSource& h, int i) {
    return h.order ? h.order[i] : i;
}
__device__ __forceinline__ RayVal hs_ray(const HitSource& h, int slot) {
    const float4 a = h.a[slot], b = h.b[slot], c = h.c[slot];
    RayVal v;
    v.o = V3(a.x, a.y, a.z);
    v.pixel = __float_as_int(a.w);
    v.d = V3(b.x, b.y, b.z);
    v.pdf = b.w;
    v.t = V3(c.x, c.y, c.z);
    v.depth = __float_as_int(c.w);
    return v;
}

// spreadBits10 (RayQueue.h:45-55)
__device__ __forceinline__ uint32_t spread_bits10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x30000FFu;
    v = (v | (v << 8)) & 0x300F00Fu;
    v = (v | (v << 4)) & 0x30C30C3u;
    v = (v | (v << 2)) & 0x9249249u;
    return v;
}

// std::clamp(v, 0, 1) for non-NaN v
__device__ __forceinline__ float clamp01(float v) { return v < 0.0f ? 0.0f : (1.0f < v ? 1.0f : v); }

// RayQueue::sortKey (RayQueue.h:135-162): octant << 30 | Morton code of the origin quantised to
// 10 bits per axis over the scene bounds (33 bits).
__device__ __forceinline__ uint64_t ray_sort_key(const V3& o, const V3& d, const RayKeyFrame& f) {
    const uint64_t octant = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) |
                            (d.z < 0.0f ? 4u : 0u);
    const uint32_t gx = (uint32_t)(clamp01((o.x - f.lo[0]) / f.ext[0]) * 1023.0f);
    const uint32_t gy = (uint32_t)(clamp01((o.y - f.lo[1]) / f.ext[1]) * 1023.0f);
    const uint32_t gz = (uint32_t)(clamp01((o.z - f.lo[2]) / f.ext[2]) * 1023.0f);
    const uint32_t morton = spread_bits10(gx) | (spread_bits10(gy) << 1) | (spread_bits10(gz) << 2);
    return (octant << 30) | (uint64_t)morton;
}

// The sort word of the ray in queue slot `slot`: the key above the slot index, or the bare key
// when the frame has no index bits.
__device__ __forceinline__ uint64_t ray_sort_word(const V3& o, const V3& d, const RayKeyFrame& f,
                                                  int slot) {
    const uint64_t key = ray_sort_key(o, d, f);
    return f.idx_bits ? (key << f.idx_bits) | (uint64_t)(uint32_t)slot : key;
}

}  // namespace wf
