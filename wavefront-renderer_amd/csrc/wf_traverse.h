// Ray traversal shared by the closest-hit, occlusion and shadow kernels: the reference's
// Moller-Trumbore test (include/geometry/Triangle.h:6-42) with the same float operations, a
// conservative slab test, and a stack walk over the two-child BVH with the stack in LDS
// (stack[depth * kTraceBlock + lane]).
#pragma once

#include "wf_device.h"
#include "wf_internal.h"

namespace wf {

constexpr int kTraceBlock = 128;
constexpr int kStack = 64;

__device__ __forceinline__ bool mt_test(const V3& o, const V3& d, const V3& v0, const V3& e1,
                                        const V3& e2, float tmin, float tmax, float& t, float& u,
                                        float& v) {
    const float EPSILON = 0.0000001f;
    V3 pvec = d.cross(e2);
    float det = e1.dot(pvec);
    if (det > -EPSILON && det < EPSILON) return false;
    float inv = 1.0f / det;
    V3 tvec = o - v0;
    float uu = inv * tvec.dot(pvec);
    if (uu < 0.0f || uu > 1.0f) return false;
    V3 qvec = tvec.cross(e1);
    float vv = inv * d.dot(qvec);
    if (vv < 0.0f || uu + vv > 1.0f) return false;
    float tt = inv * e2.dot(qvec);
    if (tt < tmin || tt > tmax) return false;
    t = tt;
    u = uu;
    v = vv;
    return true;
}

__device__ __forceinline__ float safe_rcp(float x) {
    float a = fabsf(x) < 1e-20f ? copysignf(1e-20f, x) : x;
    return 1.0f / a;
}

__device__ __forceinline__ bool slab(const float4& lo, const float4& hi, const V3& o,
                                     const V3& inv, float tmax, float& tenter) {
    float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
    float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
    float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
    float t0 = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    float t1 = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    tenter = t0;
    return t0 <= t1 * 1.00000024f;
}

struct TraceHit {
    float t;
    int32_t leaf;   // leaf (sorted) triangle index, -1 = none
    int32_t g;      // global triangle index
    float u, v;
};

template <bool ANY>
__device__ __forceinline__ bool traverse(const BvhNode* __restrict__ nodes,
                                         const float4* __restrict__ tri, int n_tris,
                                         const V3& o, const V3& d, float tnear, float tfar,
                                         int32_t* __restrict__ stack, TraceHit& h) {
    h.t = tfar;
    h.leaf = -1;
    h.g = 0x7FFFFFFF;
    h.u = h.v = 0.0f;
    if (n_tris <= 0) return false;
    const V3 inv(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
    const int lane = threadIdx.x;
    int sp = 0;
    int node = 0;
    auto test = [&](int leaf) -> bool {
        const float4 a = tri[3 * (size_t)leaf];
        const float4 b = tri[3 * (size_t)leaf + 1];
        const float4 c = tri[3 * (size_t)leaf + 2];
        float t, u, v;
        if (!mt_test(o, d, xyz(a), xyz(b), xyz(c), tnear, tfar, t, u, v)) return false;
        const int32_t g = __float_as_int(a.w);
        if (ANY) {
            h.t = t;
            h.leaf = leaf;
            h.g = g;
            return true;
        }
        if (h.leaf < 0 || t < h.t || (t == h.t && g < h.g)) {
            h.t = t;
            h.leaf = leaf;
            h.g = g;
            h.u = u;
            h.v = v;
        }
        return false;
    };
    for (;;) {
        const BvhNode nd = nodes[node];
        const int c0 = __float_as_int(nd.a.w);
        const int c1 = __float_as_int(nd.b.w);
        float t0, t1;
        const float tcur = ANY ? tfar : h.t;
        bool h0 = slab(nd.a, nd.b, o, inv, tcur, t0);
        bool h1 = slab(nd.c, nd.d, o, inv, tcur, t1);
        if (h0 && c0 < 0) {
            if (test(~c0)) return true;
            h0 = false;
        }
        if (h1 && c1 < 0) {
            if (test(~c1)) return true;
            h1 = false;
        }
        if (h0 && h1) {
            const bool first0 = t0 <= t1;
            stack[sp * kTraceBlock + lane] = first0 ? c1 : c0;
            ++sp;
            node = first0 ? c0 : c1;
        } else if (h0) {
            node = c0;
        } else if (h1) {
            node = c1;
        } else {
            if (sp == 0) break;
            --sp;
            node = stack[sp * kTraceBlock + lane];
        }
    }
    return h.leaf >= 0;
}

}  // namespace wf
