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

__device__ __forceinline__ void st_ray(const wf_ray_soa& r, int i, const RayVal& v) {
    r.ox[i] = v.o.x; r.oy[i] = v.o.y; r.oz[i] = v.o.z;
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

}  // namespace wf
