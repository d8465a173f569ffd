// Shading kernels replacing WavefrontRenderer::shadeAll (src/render/WavefrontRenderer.cpp:569-877),
// the miss loop and getSkyColor (cpp:256-265, 879-889), generatePrimaryRays (cpp:401-457),
// the AdaptiveSampler updates (include/render/AdaptiveSampler.h:40-80) and the final resolve
// (cpp:127-137).
//
// One lane per shading-queue entry, in shading order. A path owns exactly one queue entry per
// bounce, so every accumulator update is a plain read-modify-write of its own pixel: no float
// atomics, and the per-pixel order is the reference's (emission, area NEE, directional NEE,
// miss). Shadow rays are resolved by a second kernel over the compacted list of entries that
// emitted one; surviving next rays are compacted in shading order.
#include "wf_fmath.h"
#include "wf_rays.h"
#include "wf_sky.h"
#include "wf_traverse.h"

namespace wf {

struct ShadeArgs {
    int32_t physics;
    int32_t rr_min_depth;
    int32_t cost_aware_rr;
    float firefly;
    uint32_t seed;
    int32_t sample;
    int32_t job;
};

__device__ __forceinline__ V3 texel(const SceneDev& s, int tid, float u, float v) {
    const DevTexture T = s.textures[tid];
    u = u - floorf(u);
    v = v - floorf(v);
    int x = (int)(u * (float)T.width);
    int y = (int)(v * (float)T.height);
    x = x < 0 ? 0 : (x > T.width - 1 ? T.width - 1 : x);
    y = y < 0 ? 0 : (y > T.height - 1 ? T.height - 1 : y);
    const size_t idx = (size_t)y * T.width + x;
    if (T.format == WF_TEX_RGBA8_SRGB) {
        const uint32_t p = s.texels8[T.offset + idx];
        return V3(s.srgb_lut[p & 0xFFu], s.srgb_lut[(p >> 8) & 0xFFu], s.srgb_lut[(p >> 16) & 0xFFu]);
    }
    const float4 f = s.texels32[T.offset + idx];
    return V3(f.x, f.y, f.z);
}

// Material::getSurfaceColor (Material.cpp:143-163) with hashGridCell/hashToUV (SpatialHash.h:18-44).
__device__ __forceinline__ V3 surface_color(const SceneDev& s, const DevMaterial& m, const V3& p,
                                            float u, float v) {
    const int t = m.texture_id;
    if (m.use_spatial_checker && t >= 0 && t < s.n_textures) {
        const int cx = (int)floorf(p.x / m.checker_cell_size);
        const int cy = (int)floorf(p.y / m.checker_cell_size);
        const int cz = (int)floorf(p.z / m.checker_cell_size);
        const uint32_t h = ((uint32_t)cx * 73856093u ^ (uint32_t)cy * 19349663u ^
                            (uint32_t)cz * 83492791u) * 2654435761u;
        const float pu = ((h >> 16) & 0xFFFFu) / 65535.0f;
        const float pv = (h & 0xFFFFu) / 65535.0f;
        return texel(s, t, pu, pv);
    }
    if (t >= 0 && t < s.n_textures) return texel(s, t, u, v);
    return V3(m.albedo[0], m.albedo[1], m.albedo[2]);
}

// Material::emitted (Material.cpp:268-297)
__device__ __forceinline__ V3 emitted(const DevMaterial& m, const V3& ln, const V3& dfl) {
    const V3 alb(m.albedo[0], m.albedo[1], m.albedo[2]);
    if (m.type == WF_MAT_EMISSIVE) return alb * m.emission;
    if (m.type == WF_MAT_SPOTLIGHT) {
        V3 radiance = alb * m.emission;
        const float c = ln.dot(dfl);
        const float deg = fm::acos(smax(-1.0f, smin(1.0f, c))) * (180.0f / kPi);
        if (deg >= m.spot_outer_deg) return V3(0.0f, 0.0f, 0.0f);
        if (deg > m.spot_falloff_deg) {
            const float t = (deg - m.spot_falloff_deg) / (m.spot_outer_deg - m.spot_falloff_deg);
            float f = 1.0f - t;
            f = f * f * (3.0f - 2.0f * f);
            radiance = radiance * f;
        }
        return radiance;
    }
    return V3(0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ float area_to_angle(float area, float dist, float c) {
    const float safe = smax(fabsf(c), 1e-4f);
    return area * (dist * dist) / safe;
}
__device__ __forceinline__ V3 vmin(const V3& a, float c) { return V3(smin(a.x, c), smin(a.y, c), smin(a.z, c)); }

__device__ __forceinline__ V3 reflect(const V3& in, const V3& n) { return in - n * 2.0f * in.dot(n); }
__device__ __forceinline__ V3 refract(const V3& d, const V3& n, float eta) {
    const float ci = smin(1.0f, (d * -1.0f).dot(n));
    const V3 perp = (d + n * ci) * eta;
    const V3 par = n * -__builtin_sqrtf(fabsf(1.0f - perp.dot(perp)));
    return perp + par;
}
__device__ __forceinline__ float schlick(float ci, float eta) {
    float r0 = (1.0f - eta) / (1.0f + eta);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * fm::pow5(1.0f - ci);
}
__device__ __forceinline__ void frame(const V3& n, V3& t, V3& b) {
    const V3 up = fabsf(n.x) > 0.9f ? V3(0.0f, 1.0f, 0.0f) : V3(1.0f, 0.0f, 0.0f);
    t = n.cross(up).normalized();
    b = n.cross(t);
}
__device__ __forceinline__ V3 cosine_hemisphere(const V3& n, float u1, float u2) {
    const float r = __builtin_sqrtf(u1);
    const float ang = 2.0f * kPi * u2;
    float sa, ca;
    fm::sincos(ang, sa, ca);
    const float x = r * ca;
    const float z = r * sa;
    const float y = __builtin_sqrtf(smax(0.0f, 1.0f - r * r));
    V3 t, b;
    frame(n, t, b);
    return (t * x + n * y + b * z).normalized();
}
__device__ __forceinline__ V3 in_unit_sphere(PathRng& rng) {
    V3 c;
    for (int k = 0; k < kSphereTries; ++k) {
        const uint32_t base = D_SPHERE + 3u * (uint32_t)k;
        const float x = rng(base) * 2.0f - 1.0f;
        const float y = rng(base + 1) * 2.0f - 1.0f;
        const float z = rng(base + 2) * 2.0f - 1.0f;
        c = V3(x, y, z);
        if (c.dot(c) < 1.0f) return c;
    }
    return c;
}

// GGX conductor: Heitz 2018 visible-normal sampling, Smith separable G, Schlick Fresnel (F0 =
// surface colour), alpha = roughness^2.
__device__ __forceinline__ float ggx_alpha(float r) { return smax(r * r, 1e-4f); }
__device__ __forceinline__ float ggx_d(const V3& m, float a) {
    const float a2 = a * a;
    const float k = (m.x * m.x + m.y * m.y) / a2 + m.z * m.z;
    return 1.0f / (kPi * a2 * k * k);
}
__device__ __forceinline__ float ggx_g1(const V3& w, float a) {
    const float s2 = a * a * (w.x * w.x + w.y * w.y);
    return 2.0f * w.z / (w.z + __builtin_sqrtf(s2 + w.z * w.z));
}
__device__ __forceinline__ V3 schlick_rgb(const V3& f0, float c) {
    const float k = fm::pow5(1.0f - smax(0.0f, smin(1.0f, c)));
    return V3(f0.x + (1.0f - f0.x) * k, f0.y + (1.0f - f0.y) * k, f0.z + (1.0f - f0.z) * k);
}
__device__ __forceinline__ V3 to_local(const V3& w, const V3& t, const V3& b, const V3& n) {
    return V3(w.dot(t), w.dot(b), w.dot(n));
}
__device__ __forceinline__ void ggx_eval(const V3& f0, float a, const V3& n, const V3& wo,
                                         const V3& wi, V3& fcos, float& pdf) {
    V3 t, b;
    frame(n, t, b);
    const V3 lo = to_local(wo, t, b, n), li = to_local(wi, t, b, n);
    fcos = V3(0.0f, 0.0f, 0.0f);
    pdf = 0.0f;
    if (lo.z <= 0.0f || li.z <= 0.0f) return;
    const V3 m = (lo + li).normalized();
    const float d = ggx_d(m, a);
    const float g1o = ggx_g1(lo, a);
    const float g1i = ggx_g1(li, a);
    const V3 F = schlick_rgb(f0, lo.dot(m));
    const float k = d * g1o * g1i / (4.0f * lo.z);
    fcos = F * k;
    pdf = d * g1o / (4.0f * lo.z);
}
__device__ __forceinline__ bool ggx_sample(const V3& f0, float a, const V3& n, const V3& wo,
                                           float u1, float u2, V3& wi, V3& weight, float& pdf) {
    V3 t, b;
    frame(n, t, b);
    const V3 lo = to_local(wo, t, b, n);
    if (lo.z <= 0.0f) return false;
    const V3 vh = V3(a * lo.x, a * lo.y, lo.z).normalized();
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const V3 t1 = lensq > 0.0f ? V3(-vh.y, vh.x, 0.0f) / __builtin_sqrtf(lensq) : V3(1.0f, 0.0f, 0.0f);
    const V3 t2 = vh.cross(t1);
    const float r = __builtin_sqrtf(u1);
    const float phi = 2.0f * kPi * u2;
    float sp, cp;
    fm::sincos(phi, sp, cp);
    const float p1 = r * cp;
    float p2 = r * sp;
    const float sh = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - sh) * __builtin_sqrtf(smax(0.0f, 1.0f - p1 * p1)) + sh * p2;
    const V3 nh = t1 * p1 + t2 * p2 + vh * __builtin_sqrtf(smax(0.0f, 1.0f - p1 * p1 - p2 * p2));
    const V3 m = V3(a * nh.x, a * nh.y, smax(1e-6f, nh.z)).normalized();
    const float om = lo.dot(m);
    const V3 li = m * (2.0f * om) - lo;
    if (li.z <= 0.0f) return false;
    weight = schlick_rgb(f0, om) * ggx_g1(li, a);
    pdf = ggx_d(m, a) * ggx_g1(lo, a) / (4.0f * lo.z);
    wi = (t * li.x + b * li.y + n * li.z).normalized();
    return true;
}
__device__ __forceinline__ float fresnel_dielectric(float ci, float eta, float& ct) {
    const float s2 = eta * eta * (1.0f - ci * ci);
    if (s2 >= 1.0f) {
        ct = 0.0f;
        return 1.0f;
    }
    ct = __builtin_sqrtf(1.0f - s2);
    const float rs = (eta * ci - ct) / (eta * ci + ct);
    const float rp = (ci - eta * ct) / (ci + eta * ct);
    return 0.5f * (rs * rs + rp * rp);
}

// Material::scatter (Material.cpp:165-266) + the microfacet models.
__device__ __forceinline__ bool scatter(const SceneDev& s, const DevMaterial& m, const V3& d,
                                        const V3& n, const V3& p, float u, float v, bool front,
                                        PathRng& rng, V3& att, V3& dir, float& pdf) {
    const V3 sc = surface_color(s, m, p, u, v);
    switch (m.type) {
    case WF_MAT_DIFFUSE: {
        const float u1 = rng(D_DIR_U1);
        const float u2 = rng(D_DIR_U2);
        dir = cosine_hemisphere(n, u1, u2);
        att = sc;
        pdf = smax(0.0f, n.dot(dir)) / kPi;
        return true;
    }
    case WF_MAT_METAL: {
        if (m.model == WF_MODEL_MICROFACET) {
            const V3 wo = (d * -1.0f).normalized();
            const float u1 = rng(D_DIR_U1);
            const float u2 = rng(D_DIR_U2);
            return ggx_sample(sc, ggx_alpha(m.roughness), n, wo, u1, u2, dir, att, pdf);
        }
        const V3 reflected = reflect(d, n);
        const V3 fuzz = in_unit_sphere(rng) * m.roughness;
        dir = (reflected + fuzz).normalized();
        att = sc;
        pdf = -1.0f;
        return dir.dot(n) > 0.0f;
    }
    case WF_MAT_EMISSIVE:
    case WF_MAT_SPOTLIGHT:
        att = sc * m.emission;
        pdf = -1.0f;
        return false;
    case WF_MAT_GLASS: {
        att = sc;
        pdf = -1.0f;
        const float eta = front ? (1.0f / m.ior) : m.ior;
        const V3 id = d.normalized();
        const float ci = smin(1.0f, (id * -1.0f).dot(n));
        if (m.model == WF_MODEL_MICROFACET) {
            float ct;
            const float F = fresnel_dielectric(ci, eta, ct);
            if (rng(D_LOBE) < F)
                dir = reflect(id, n);
            else
                dir = id * eta + n * (eta * ci - ct);
            return true;
        }
        const float si = __builtin_sqrtf(1.0f - ci * ci);
        const bool tir = eta * si > 1.0f;
        if (tir || schlick(ci, eta) > rng(D_LOBE))
            dir = reflect(id, n);
        else
            dir = refract(id, n, eta);
        return true;
    }
    case WF_MAT_PLASTIC: {
        const float ci = smax(0.0f, (d * -1.0f).dot(n));
        const float R = schlick(ci, 1.5f);
        if (rng(D_LOBE) < R) {
            const V3 reflected = reflect(d.normalized(), n);
            const V3 fuzz = in_unit_sphere(rng) * m.roughness;
            dir = (reflected + fuzz).normalized();
            att = V3(1.0f, 1.0f, 1.0f);
            pdf = -1.0f;
        } else {
            const float u1 = rng(D_DIR_U1);
            const float u2 = rng(D_DIR_U2);
            dir = cosine_hemisphere(n, u1, u2);
            att = sc;
            pdf = smax(0.0f, n.dot(dir)) / kPi;
        }
        return true;
    }
    default:
        pdf = -1.0f;
        return false;
    }
}

// One entry of the shading queue: what shadeAll reads of a hit (ShadingQueue::getHitRecord and
// the ray's state).
struct ShadeIn {
    int32_t pixel, depth;
    float pdf_in;
    V3 P, N, D, T;
    float t, u, v;
    int32_t mat_id;
    bool front;
};

// One entry of shadeAll (cpp:602-797), its outputs at position i of o. With a tag (>= 0) the
// entry's job supplies the RNG stream (seed, job id, job-local pixel, sample) and the depth
// limit; otherwise they come from ShadeArgs and the caller bounds the depth.
__device__ __forceinline__ void shade_entry(const SceneDev& s, const ShadeIn& h, int i,
                                            const ShadeArgs& a,
                                            const double* __restrict__ mat_rel_cost,
                                            float* __restrict__ accum, const ShadeOut& o,
                                            int32_t tag, const JobDev* __restrict__ jobs) {
    const int32_t pixel = h.pixel;
    const int32_t depth = h.depth;
    const float pdf_in = h.pdf_in;
    const V3 P = h.P;
    const V3 N = h.N;
    const V3 D = h.D;
    const V3 T = h.T;
    const float t = h.t;
    const float u = h.u, v = h.v;
    const int mat_id = h.mat_id;
    const bool front = h.front;
    const DevMaterial m = s.materials[mat_id];
    const bool unbiased = a.physics == WF_PHYSICS_UNBIASED;
    uint32_t seed = a.seed, job = (uint32_t)a.job, sample = (uint32_t)a.sample;
    int32_t rng_pixel = pixel;
    int32_t depth_limit = 0x7FFFFFFF;
    if (tag >= 0) {
        const int32_t tg = tag;
        const JobDev& J = jobs[tg >> kTagJobShift];
        seed = J.seed;
        job = (uint32_t)J.tag;
        const int32_t s_job = tg & kTagSampleMask;
        sample = (uint32_t)(s_job + J.first_sample);
        rng_pixel = lane_pixel(J, pixel, s_job) - J.pixel_base;
        depth_limit = J.max_depth;
    }
    PathRng rng{seed, job, (uint32_t)rng_pixel, sample, (uint32_t)depth};
    float* acc = accum + (size_t)pixel * 3;
    uint32_t sflag = 0;

    // emission + MIS (cpp:629-658)
    V3 e = emitted(m, N, D * -1.0f);
    if (e.x > 0.0f || e.y > 0.0f || e.z > 0.0f) {
        float w = 1.0f;
        if (depth > 0 && pdf_in > 0.0f) {
            const float cl = smax(0.0f, (D * -1.0f).dot(N));
            if (cl > 0.0f) {
                if (unbiased) {
                    const float pa = m.light_pdf_area * (1.0f - s.env_choice_prob);
                    w = mis_power(pdf_in, area_to_angle(pa, t, cl));
                } else {
                    const float lpa = s.total_emissive_area <= 0.0f ? 0.0f : 1.0f / s.total_emissive_area;
                    w = mis_balance(pdf_in, area_to_angle(lpa, t, cl));
                }
            }
        }
        e = vmin(e, a.firefly);
        const V3 c = T * e * w;
        acc[0] = acc[0] + c.x;
        acc[1] = acc[1] + c.y;
        acc[2] = acc[2] + c.z;
    }

    // next-event estimation
    const bool ggx = m.type == WF_MAT_METAL && m.model == WF_MODEL_MICROFACET;
    const bool nee = m.type == WF_MAT_DIFFUSE || m.type == WF_MAT_PLASTIC || (unbiased && ggx);
    if (nee && !unbiased) {
        // Scene::sampleLight (Scene.cpp:192-237)
        const int ne = s.n_emissive_meshes;
        if (ne > 0) {
            const int ei = (int)(rng(D_LIGHT_PICK) * (float)ne) % ne;
            const int mi = s.emissive_meshes[ei];
            const int tc = s.mesh_tri_count[mi];
            if (tc > 0) {
                const int k = (int)(rng(D_LIGHT_TRI) * (float)tc) % tc;
                const int slot = s.emissive_light_base[ei] + k;
                float r1 = rng(D_LIGHT_R1);
                float r2 = rng(D_LIGHT_R2);
                if (r1 + r2 > 1.0f) {
                    r1 = 1.0f - r1;
                    r2 = 1.0f - r2;
                }
                const float r3 = 1.0f - r1 - r2;
                const V3 v0 = xyz(s.light_v0[slot]);
                const V3 v1 = xyz(s.light_v1[slot]), v2 = xyz(s.light_v2[slot]);
                const V3 lp = v0 * r3 + v1 * r1 + v2 * r2;
                const V3 ln = (v1 - v0).cross(v2 - v0).normalized();
                const V3 dfl = (P - lp).normalized();
                const V3 le = emitted(s.materials[s.mesh_material[mi]], ln, dfl);
                const V3 tl = lp - P;
                const float dist = tl.length();
                const V3 dl = tl / dist;
                const float cs = N.dot(dl);
                const float cl = (dl * -1.0f).dot(ln);
                if (cs > 0.0f && cl > 0.0f) {
                    const V3 sc = surface_color(s, m, P, u, v);
                    const float geometry = (cs * cl) / (dist * dist);
                    const V3 brdf = sc * (1.0f / kPi);
                    const float lpa = s.total_emissive_area <= 0.0f ? 0.0f : 1.0f / s.total_emissive_area;
                    const float lps = area_to_angle(lpa, dist, cl);
                    const float bps = cs / kPi;
                    const float w = mis_balance(lps, bps);
                    const V3 direct = vmin(T * brdf * le * geometry * s.total_emissive_area * w, a.firefly);
                    st3(o.ldx, o.ldy, o.ldz, i, dl);
                    o.lmax[i] = dist;
                    o.lr[i] = direct.x;
                    o.lg[i] = direct.y;
                    o.lb[i] = direct.z;
                    sflag |= 1u;
                }
            }
        }
    } else if (nee && unbiased) {
        const bool use_env = s.env_choice_prob > 0.0f &&
                             (s.n_light_tris == 0 || rng(D_ENV_CHOICE) < s.env_choice_prob);
        const V3 wo = (D * -1.0f).normalized();
        const V3 sc = surface_color(s, m, P, u, v);
        V3 wi, le;
        float lpdf = 0.0f, maxd = 0.0f;
        bool ok = false;
        if (use_env) {
            const int nt = s.env_w * s.env_h;
            int slot = (int)(rng(D_ENV_SLOT) * (float)nt);
            slot = slot < nt - 1 ? slot : nt - 1;
            if (rng(D_ENV_COIN) >= s.env_prob[slot]) slot = s.env_alias[slot];
            const int tx = slot % s.env_w, ty = slot / s.env_w;
            const float eu = ((float)tx + rng(D_ENV_JX)) / (float)s.env_w;
            const float ev = ((float)ty + rng(D_ENV_JY)) / (float)s.env_h;
            const float phi = (eu - 0.5f) * (2.0f * kPi);
            const float lat = (0.5f - ev) * kPi;
            float sl, cl, sp, cp;
            fm::sincos(lat, sl, cl);
            fm::sincos(phi, sp, cp);
            wi = V3(cl * cp, sl, cl * sp);
            if (N.dot(wi) > 0.0f) {
                float lc;
                const int tx2 = env_texel(s, wi, lc);
                lpdf = env_pdf_sa(s, tx2, lc);
                le = env_radiance(s, tx2);
                maxd = 1e6f;
                ok = lpdf > 0.0f;
            }
        } else if (s.n_light_tris > 0) {
            const int nl = s.n_light_tris;
            int slot = (int)(rng(D_LIGHT_PICK) * (float)nl);
            slot = slot < nl - 1 ? slot : nl - 1;
            if (rng(D_LIGHT_TRI) >= s.light_prob[slot]) slot = s.light_alias[slot];
            float r1 = rng(D_LIGHT_R1);
            float r2 = rng(D_LIGHT_R2);
            if (r1 + r2 > 1.0f) {
                r1 = 1.0f - r1;
                r2 = 1.0f - r2;
            }
            const float r3 = 1.0f - r1 - r2;
            const V3 v0 = xyz(s.light_v0[slot]);
            const V3 v1 = xyz(s.light_v1[slot]), v2 = xyz(s.light_v2[slot]);
            const V3 lp = v0 * r3 + v1 * r1 + v2 * r2;
            const V3 ln = (v1 - v0).cross(v2 - v0).normalized();
            const int lmat = s.light_mat[slot];
            const V3 tl = lp - P;
            const float dist = tl.length();
            const V3 dl = tl / dist;
            const float cs = N.dot(dl);
            const float cl = fabsf((dl * -1.0f).dot(ln));
            if (cs > 0.0f && cl > 0.0f) {
                const DevMaterial lm = s.materials[lmat];
                le = emitted(lm, ln, dl * -1.0f);
                const float pa = lm.light_pdf_area * (1.0f - s.env_choice_prob);
                lpdf = area_to_angle(pa, dist, cl);
                wi = dl;
                maxd = dist;
                ok = lpdf > 0.0f;
            }
        }
        if (ok) {
            V3 fcos;
            float bpdf;
            if (ggx) {
                ggx_eval(sc, ggx_alpha(m.roughness), N, wo, wi, fcos, bpdf);
            } else {
                const float c = N.dot(wi);
                fcos = sc * (1.0f / kPi) * smax(0.0f, c);
                bpdf = smax(0.0f, c) / kPi;
            }
            const float w = mis_power(lpdf, bpdf);
            const V3 direct = vmin(T * fcos * le * (w / lpdf), a.firefly);
            st3(o.ldx, o.ldy, o.ldz, i, wi);
            o.lmax[i] = maxd;
            o.lr[i] = direct.x;
            o.lg[i] = direct.y;
            o.lb[i] = direct.z;
            sflag |= 1u;
        }
    }
    // directional light (cpp:712-737), every material type
    if (s.directional.intensity > 0.0f) {
        const V3 ldir(s.directional.direction[0], s.directional.direction[1], s.directional.direction[2]);
        const V3 tl = ldir * -1.0f;
        const float cs = N.dot(tl);
        if (cs > 0.0f) {
            const V3 sc = surface_color(s, m, P, u, v);
            const V3 brdf = sc * (1.0f / kPi);
            const V3 col(s.directional.color[0], s.directional.color[1], s.directional.color[2]);
            const V3 direct = vmin(T * brdf * col * s.directional.intensity * cs, a.firefly);
            o.dr[i] = direct.x;
            o.dg[i] = direct.y;
            o.db[i] = direct.z;
            sflag |= 2u;
        }
    }
    o.shadow_flag[i] = sflag;

    // scatter + Russian roulette (cpp:739-796); with live cost tracking a quarter of the hits
    // (shouldTimeThisRay, cpp:612) time the scatter: the wave's wall-clock ticks, added to the
    // material's and, when it has one, the texture's sums
    V3 att, dir;
    float spdf = -1.0f;
    const bool timed = o.probe.mat_sum != nullptr && rng(D_COST_COIN) < 0.25f;
    const unsigned long long t_start = timed ? wall_clock64() : 0ull;
    bool alive = scatter(s, m, D, N, P, u, v, front, rng, att, dir, spdf);
    if (timed) {
        const unsigned long long dt = wall_clock64() - t_start;
        atomicAdd(&o.probe.mat_sum[mat_id], dt);
        atomicAdd(&o.probe.mat_cnt[mat_id], 1u);
        if (m.texture_id >= 0 && m.texture_id < s.n_textures) {   // ids beyond the table: untracked
            atomicAdd(&o.probe.tex_sum[m.texture_id], dt);
            atomicAdd(&o.probe.tex_cnt[m.texture_id], 1u);
        }
    }
    V3 nt = T * att;
    if (alive && depth + 1 >= a.rr_min_depth) {
        const float base = smin(0.95f, smax(smax(nt.x, nt.y), nt.z));
        float surv = base;
        if (a.cost_aware_rr) {
            const float rc = (float)mat_rel_cost[mat_id];
            const float x = base / rc;
            surv = x < 0.05f ? 0.05f : (0.95f < x ? 0.95f : x);
        }
        if (rng(D_RR) > surv)
            alive = false;
        else
            nt = nt / surv;
    }
    if (depth + 1 >= depth_limit) alive = false;
    o.alive[i] = alive ? 1u : 0u;
    if (alive) {
        st3(o.ndx, o.ndy, o.ndz, i, dir);
        o.ntr[i] = nt.x;
        o.ntg[i] = nt.y;
        o.ntb[i] = nt.z;
        o.npdf[i] = spdf;
    }
}

// Shading-queue entries (the C-ABI and pass-loop layout).
__global__ void k_shade(const SceneDev s, wf_shade_soa q, int n_host,
                        const int32_t* __restrict__ nd, ShadeArgs a,
                        const double* __restrict__ mat_rel_cost, float* __restrict__ accum,
                        ShadeOut o, const int32_t* __restrict__ tags,
                        const JobDev* __restrict__ jobs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= load_count(n_host, nd)) return;
    ShadeIn h;
    h.pixel = q.pixel[i];
    h.depth = q.depth[i];
    h.pdf_in = q.pdf[i];
    h.P = ld3(q.px, q.py, q.pz, i);
    h.N = ld3(q.nx, q.ny, q.nz, i);
    h.D = ld3(q.dx, q.dy, q.dz, i);
    h.T = V3(q.tr[i], q.tg[i], q.tb[i]);
    h.t = q.t[i];
    h.u = q.u[i];
    h.v = q.v[i];
    h.mat_id = q.mat[i];
    h.front = q.front ? q.front[i] != 0 : true;
    shade_entry(s, h, i, a, mat_rel_cost, accum, o, tags ? tags[i] : -1, jobs);
}

// Index mode: entry j is traced ray h.list[j], mapped (HitRecord) from the ray and its closest
// hit in registers, then shaded; its hit columns are written for the kernels after shading.
__global__ void k_shade_hits(const SceneDev s, HitSource src, int n_host,
                             const int32_t* __restrict__ nd, ShadeArgs a,
                             const double* __restrict__ mat_rel_cost, float* __restrict__ accum,
                             ShadeOut o, const JobDev* __restrict__ jobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= load_count(n_host, nd)) return;
    const int i = hs_traced(src, j);
    const int slot = hs_slot(src, i);
    const RayVal ray = hs_ray(src, slot);
    const float t = src.t[i];
    const MappedHit m = map_hit_vals(s, ray.o, ray.d, src.leaf[i], t, src.u[i], src.v[i]);
    const int32_t tag = src.tags ? src.tags[slot] : -1;
    ShadeIn h;
    h.pixel = ray.pixel;
    h.depth = ray.depth;
    h.pdf_in = ray.pdf;
    h.P = m.P;
    h.N = m.N;
    h.D = ray.d;
    h.T = ray.t;
    h.t = t;
    h.u = m.u;
    h.v = m.v;
    h.mat_id = m.mat;
    h.front = m.front;
    st3(o.hpx, o.hpy, o.hpz, j, m.P);
    o.hpixel[j] = ray.pixel;
    o.hdepth[j] = ray.depth;
    if (o.htag) o.htag[j] = tag;
    shade_entry(s, h, j, a, mat_rel_cost, accum, o, tag, jobs);
}

// Unordered compaction after k_shade: shadow entries and next rays appended at device cursors
// (kAppendItems entries per thread in coalesced rounds, one atomic per block and cursor), shadow
// rays counted (one atomic per block).
constexpr int kAppendBlock = 1024;
constexpr int kAppendItems = 4;
__global__ __launch_bounds__(kAppendBlock) void k_shade_append(HitCols q, ShadeOut o, int n_host,
                                                               const int32_t* __restrict__ nd) {
    __shared__ uint32_t lds[kAppendItems * kAppendBlock / 64];
    __shared__ uint32_t block_rays;
    const int n = load_count(n_host, nd);
    const int base = blockIdx.x * kAppendBlock * kAppendItems;
    if (base >= n) return;
    bool shadow[kAppendItems], alive[kAppendItems];
    int si[kAppendItems], k[kAppendItems];
    int rays = 0;
#pragma unroll
    for (int j = 0; j < kAppendItems; ++j) {
        const int i = base + j * kAppendBlock + (int)threadIdx.x;
        const uint32_t f = i < n ? o.shadow_flag[i] : 0u;
        shadow[j] = f != 0u;
        alive[j] = i < n && o.alive[i] != 0u;
        rays += (int)(f & 1u) + (int)((f >> 1) & 1u);
    }
    if (threadIdx.x == 0) block_rays = 0u;
    block_append_rounds<kAppendItems, kAppendBlock>(o.shadow_entries, shadow, si, lds);
    for (int off = 32; off >= 1; off >>= 1) rays += __shfl_xor(rays, off, 64);
    if (__lane_id() == 0 && rays) atomicAdd(&block_rays, (uint32_t)rays);
    block_append_rounds<kAppendItems, kAppendBlock>(o.next_count, alive, k, lds);
    if (threadIdx.x == 0 && block_rays) atomicAdd(o.shadow_rays, (int32_t)block_rays);
#pragma unroll
    for (int j = 0; j < kAppendItems; ++j) {
        const int i = base + j * kAppendBlock + (int)threadIdx.x;
        if (si[j] >= 0) o.shadow_list[si[j]] = i;
        if (k[j] >= 0) {
            RayVal r;
            r.o = ld3(q.px, q.py, q.pz, i);
            r.d = ld3(o.ndx, o.ndy, o.ndz, i);
            r.t = ld3(o.ntr, o.ntg, o.ntb, i);
            r.pixel = q.pixel[i];
            r.depth = q.depth[i] + 1;
            r.pdf = o.npdf[i];
            if (o.next_packed.a) st_ray(o.next_packed, k[j], r);
            else st_ray(o.next, k[j], r);
            if (o.next_tags) o.next_tags[k[j]] = q.tag[i];
            if (o.next_keys) o.next_keys[k[j]] = ray_sort_word(r.o, r.d, o.key_frame, k[j]);
        }
    }
}

__global__ void k_flags_nonzero(const uint32_t* __restrict__ in, int n_host,
                                const int32_t* __restrict__ nd, uint32_t* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < load_count(n_host, nd)) out[i] = in[i] != 0u ? 1u : 0u;
}

__global__ void k_compact_index(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ prefix,
                                int n_host, const int32_t* __restrict__ nd,
                                int32_t* __restrict__ list) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < load_count(n_host, nd) && flag[i]) list[prefix[i]] = i;
}

__global__ void k_next_rays(wf_shade_soa q, ShadeOut o, const uint32_t* __restrict__ prefix,
                            int n_host, const int32_t* __restrict__ nd, wf_ray_soa next,
                            const int32_t* __restrict__ tags_in, int32_t* __restrict__ tags_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= load_count(n_host, nd) || !o.alive[i]) return;
    const int k = (int)prefix[i];
    if (tags_out) tags_out[k] = tags_in[i];
    next.ox[k] = q.px[i];
    next.oy[k] = q.py[i];
    next.oz[k] = q.pz[i];
    next.dx[k] = o.ndx[i];
    next.dy[k] = o.ndy[i];
    next.dz[k] = o.ndz[i];
    next.tr[k] = o.ntr[i];
    next.tg[k] = o.ntg[i];
    next.tb[k] = o.ntb[i];
    next.pixel[k] = q.pixel[i];
    next.depth[k] = q.depth[i] + 1;
    next.pdf[k] = o.npdf[i];
}

// Miss loop (cpp:256-265): accum += T * getSkyColor(dir); the unbiased track weights environment
// hits of non-delta bounces by the power heuristic against environment NEE.
__global__ void k_miss(const SceneDev s, wf_ray_soa q, int n_host, const int32_t* __restrict__ nd,
                       int physics, float* __restrict__ accum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= load_count(n_host, nd)) return;
    add_miss(s, V3(q.dx[i], q.dy[i], q.dz[i]), V3(q.tr[i], q.tg[i], q.tb[i]), q.pixel[i],
             q.depth[i], q.pdf[i], physics, accum);
}

// EmbreeAccelerator::occluded for the NEE rays of each listed entry, then the contributions in
// the reference order: area/environment light first, directional light second (cpp:678, 720).
// The body of the shadow kernels, run by a persistent block over entries [0, n) of list; stack,
// xchg and spill as for closest_walks.
template <class NODE, int STACK>
__device__ __forceinline__ void shadow_walks(const SceneDev& s, const NODE* __restrict__ nodes,
                                             const HitCols& q, const ShadeOut& o,
                                             const int32_t* __restrict__ list, int n,
                                             float* __restrict__ accum,
                                             int32_t* __restrict__ counter, int refill_lanes,
                                             int32_t* __restrict__ stack, int4* __restrict__ xchg,
                                             volatile int32_t* spill) {
    using W = Walker<true, STACK>;
    int4* wx = xchg + (threadIdx.x >> 6) * kXchgPerWave;
    const V3 tl = V3(s.directional.direction[0], s.directional.direction[1],
                     s.directional.direction[2]) * -1.0f;
    W w;
    // per lane: one shadow entry at a time, its area/environment ray (phase 0) then its
    // directional ray (phase 1); phase 2 = both resolved
    int item = -1, i = 0, phase = 2;
    uint32_t f = 0;
    bool occ_area = false, occ_dir = false;
    WorkQueue wq;
    wq.init(counter, n);
    for (;;) {
        const int r = refill(wq, item < 0, refill_lanes);
        if (r >= 0) {
            item = r;
            i = list[r];
            f = o.shadow_flag[i];
            occ_area = false;
            occ_dir = false;
            phase = 2;
            const V3 P = ld3(q.px, q.py, q.pz, i);
            if (f & 1u) {
                const float tfar = o.lmax[i] - 0.001f;
                if (tfar < 0.0f) {
                    occ_area = true;
                } else {
                    phase = 0;
                    w.start(P, ld3(o.ldx, o.ldy, o.ldz, i), 0.001f, tfar);
                }
            }
            if (phase == 2 && (f & 2u)) {
                phase = 1;
                w.start(P, tl, 0.001f, 1e6f - 0.001f);
            }
        }
        if (__ballot(item >= 0) == 0) {
            if (wq.exhausted()) break;
            continue;
        }
        // every lane of the wave takes part in the step (idle lanes test others' triangles)
        const bool walked = w.step_wave(item >= 0 && phase != 2 && s.n_tris > 0, nodes, s.tri,
                                        stack, spill, wx, refill_lanes);
        if (item < 0) continue;
        if (phase != 2 && (s.n_tris <= 0 || walked)) {
            const bool occ = w.h.leaf >= 0;
            if (phase == 0) {
                occ_area = occ;
                phase = 2;
                if (f & 2u) {
                    phase = 1;
                    w.start(ld3(q.px, q.py, q.pz, i), tl, 0.001f, 1e6f - 0.001f);
                }
            } else {
                occ_dir = occ;
                phase = 2;
            }
        }
        if (phase == 2) {
            // contributions in the reference order: area/environment light, then directional
            float* acc = accum + (size_t)q.pixel[i] * 3;
            if ((f & 1u) && !occ_area) {
                acc[0] = acc[0] + o.lr[i];
                acc[1] = acc[1] + o.lg[i];
                acc[2] = acc[2] + o.lb[i];
            }
            if ((f & 2u) && !occ_dir) {
                acc[0] = acc[0] + o.dr[i];
                acc[1] = acc[1] + o.dg[i];
                acc[2] = acc[2] + o.db[i];
            }
            item = -1;
        }
    }
}

template <class NODE, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_shadow(const SceneDev s, const NODE* __restrict__ nodes,
                                                        HitCols q, ShadeOut o,
                                                        const int32_t* __restrict__ list, int count_host,
                                                        const int32_t* __restrict__ count_dev,
                                                        float* __restrict__ accum,
                                                        int32_t* __restrict__ counter,
                                                        int refill_lanes) {
    __shared__ int32_t stack[lds_entries<STACK>() * kTraceBlock];
    __shared__ int4 xchg[kTraceBlock / 64 * kXchgPerWave];   // candidate tests and stashes
    volatile int32_t spill[Walker<true, STACK>::X > 0 ? Walker<true, STACK>::X : 1];
    shadow_walks<NODE, STACK>(s, nodes, q, o, list, load_count(count_host, count_dev), accum,
                              counter, refill_lanes, stack, xchg, spill);
}

// Closest hits of a bounce's rays, then the deferred shadow walks of the bounce before it, in
// one persistent launch: a wave that finds the ray queue drained moves on to shadow entries
// while the slowest rays of other waves finish. The two queues are independent (the shadow
// entries only add to the accumulators of paths shaded the bounce before).
template <class NODE, int STACK, class R>
__global__ __launch_bounds__(kTraceBlock) void k_closest_shadow(
    const SceneDev s, const NODE* __restrict__ nodes, R rays, int n_host,
    const int32_t* __restrict__ nd, int32_t* __restrict__ out_leaf, float* __restrict__ out_t,
    float* __restrict__ out_u, float* __restrict__ out_v, int32_t* __restrict__ ray_counter,
    int refill_lanes, HitCols q, ShadeOut o, const int32_t* __restrict__ list,
    int shadow_host, const int32_t* __restrict__ shadow_dev, float* __restrict__ accum,
    int32_t* __restrict__ shadow_counter, int shadow_refill_lanes) {
    __shared__ int32_t stack[lds_entries<STACK>() * kTraceBlock];
    __shared__ int4 xchg[kTraceBlock / 64 * kXchgPerWave];   // candidate tests and stashes
    volatile int32_t spill[Walker<false, STACK>::X > 0 ? Walker<false, STACK>::X : 1];
    closest_walks<NODE, STACK>(nodes, s.tri, s.n_tris, rays, load_count(n_host, nd), 0.001f,
                               1e9f, out_leaf, out_t, out_u, out_v, nullptr, ray_counter,
                               refill_lanes, stack, xchg, spill);
    shadow_walks<NODE, STACK>(s, nodes, q, o, list, load_count(shadow_host, shadow_dev), accum,
                              shadow_counter, shadow_refill_lanes, stack, xchg, spill);
}

__global__ void k_count_shadow_rays(const uint32_t* __restrict__ flag, const int32_t* __restrict__ list,
                                    int count_host, const int32_t* __restrict__ count_dev,
                                    int32_t* __restrict__ total) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (k < load_count(count_host, count_dev)) {
        const uint32_t f = flag[list[k]];
        c = (int)(f & 1u) + (int)((f >> 1) & 1u);
    }
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(total, c);
}

static wf_status shade_outputs(wf_context* ctx, int n, ShadeOut& o) {
    void* p = nullptr;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    WF_TRY(workspace(ctx, "shade_out", nn * 4 * 25, &p));
    float* f = (float*)p;
    o.ldx = f + 0 * nn; o.ldy = f + 1 * nn; o.ldz = f + 2 * nn; o.lmax = f + 3 * nn;
    o.lr = f + 4 * nn; o.lg = f + 5 * nn; o.lb = f + 6 * nn;
    o.dr = f + 7 * nn; o.dg = f + 8 * nn; o.db = f + 9 * nn;
    o.shadow_flag = (uint32_t*)(f + 10 * nn);
    o.ndx = f + 11 * nn; o.ndy = f + 12 * nn; o.ndz = f + 13 * nn;
    o.ntr = f + 14 * nn; o.ntg = f + 15 * nn; o.ntb = f + 16 * nn;
    o.npdf = f + 17 * nn;
    o.alive = (uint32_t*)(f + 18 * nn);
    o.hpx = f + 19 * nn; o.hpy = f + 20 * nn; o.hpz = f + 21 * nn;
    o.hpixel = (int32_t*)(f + 22 * nn);
    o.hdepth = (int32_t*)(f + 23 * nn);
    o.htag = (int32_t*)(f + 24 * nn);
    o.append = 0;
    o.shadow_list = nullptr;
    o.shadow_entries = o.shadow_rays = o.next_count = nullptr;
    o.next = wf_ray_soa{};
    o.next_packed = RayPacked{};
    o.next_tags = nullptr;
    o.next_keys = nullptr;
    o.key_frame = RayKeyFrame{};
    o.probe = CostProbe{};
    return WF_OK;
}

// refill threshold of the shadow-ray kernel
static int shadow_refill(const wf_context* ctx) {
    return ctx->tuning.shadow_refill_lanes > 0 ? ctx->tuning.shadow_refill_lanes
                                               : ctx->tuning.refill_lanes;
}

template <class R>
static wf_status closest_deferred(wf_context* ctx, const wf_scene* s, const R* rays, Count n,
                                  int32_t* hit_leaf, float* hit_t, float* hit_bu, float* hit_bv,
                                  const DeferredShadow& d) {
    WF_TRY(scene_ready(s));
    if (n.n <= 0) return WF_OK;
    void *prc, *psc;
    WF_TRY(workspace(ctx, "trace_counter", sizeof(int32_t) * kWorkInts, &prc));
    WF_TRY(workspace(ctx, "shadow_counter", sizeof(int32_t) * kWorkInts, &psc));
    WF_TRY(wf_fill_u32(ctx, prc, 0u, kWorkInts));
    WF_TRY(wf_fill_u32(ctx, psc, 0u, kWorkInts));
    // rays as for the closest-hit kernel (44 B), shadow entries as for the shadow kernel (41 B)
    Timed tm(ctx, "traverse_closest_shadow", 44.0, n, 41.0, d.n);
    with_bvh(s->dev, [&](auto nodes, auto S) {
        k_closest_shadow<std::remove_cv_t<std::remove_pointer_t<decltype(nodes)>>,
                         decltype(S)::value, R>
            <<<persistent_grid(n.n, ctx), kTraceBlock, 0, ctx->stream>>>(
                s->dev, nodes, *rays, n.n, n.dev, hit_leaf, hit_t, hit_bu, hit_bv,
                (int32_t*)prc, ctx->tuning.refill_lanes, d.hc, d.o, d.list, d.n.n, d.n.dev,
                d.accum, (int32_t*)psc, shadow_refill(ctx));
    });
    WF_HIP(hipGetLastError());
    return WF_OK;
}

wf_status trace_closest_deferred(wf_context* ctx, const wf_scene* s, const RayPacked* rays,
                                 Count n, int32_t* hit_leaf, float* hit_t, float* hit_bu,
                                 float* hit_bv, const DeferredShadow& d) {
    return closest_deferred(ctx, s, rays, n, hit_leaf, hit_t, hit_bu, hit_bv, d);
}

wf_status trace_closest_deferred(wf_context* ctx, const wf_scene* s, const RayPackedOrdered* rays,
                                 Count n, int32_t* hit_leaf, float* hit_t, float* hit_bu,
                                 float* hit_bv, const DeferredShadow& d) {
    return closest_deferred(ctx, s, rays, n, hit_leaf, hit_t, hit_bu, hit_bv, d);
}

wf_status shade_bounce(wf_context* ctx, const wf_scene* s, const wf_shade_soa* q, Count n,
                       const wf_shade_params* p, float* accum, wf_ray_soa* next,
                       int32_t* counts_dev, const ShadeJobs& jobs, const uint32_t** alive_out,
                       const CostProbe* probe) {
    WF_TRY(scene_ready(s));
    if (n.n <= 0) {
        WF_TRY(wf_fill_u32(ctx, counts_dev, 0u, 3));
        return WF_OK;
    }
    ShadeOut o;
    WF_TRY(shade_outputs(ctx, n.n, o));
    if (probe) o.probe = *probe;
    void *pprefix, *plist, *pshflag;
    const size_t nn = (size_t)n.n;
    WF_TRY(workspace(ctx, "shade_prefix", nn * 4, &pprefix));
    WF_TRY(workspace(ctx, "shade_list", nn * 4, &plist));
    WF_TRY(workspace(ctx, "shade_sflag", nn * 4, &pshflag));
    ShadeArgs a{p->physics, p->rr_min_depth, p->cost_aware_rr, p->firefly_threshold, p->seed,
                p->sample_index, p->job};
    if (jobs.unordered) {
        // queue order is free: one shading kernel appends the shadow entries and next rays
        o.append = 1;
        o.shadow_list = (int32_t*)plist;
        o.next_count = counts_dev;
        o.shadow_rays = counts_dev + 1;
        o.shadow_entries = counts_dev + 2;
        if (next) o.next = *next;
        o.next_packed = jobs.next_packed;
        o.next_tags = jobs.next_tags;
        o.next_keys = jobs.next_keys;
        o.key_frame = jobs.key_frame;
        WF_TRY(wf_fill_u32(ctx, counts_dev, 0u, 3));
        {
            Timed tm(ctx, "shade", 120.0, n);
            k_shade<<<grid_for(n.n, 128), 128, 0, ctx->stream>>>(s->dev, *q, n.n, n.dev, a,
                                                                 s->dev.mat_rel_cost, accum, o,
                                                                 jobs.tags, jobs.jobs);
            WF_HIP(hipGetLastError());
        }
        const HitCols hc{q->px, q->py, q->pz, q->pixel, q->depth, jobs.tags};
        {
            Timed tm(ctx, "next_rays", o.next_keys ? 64.0 : 56.0, n);
            k_shade_append<<<grid_for(n.n, kAppendBlock * kAppendItems), kAppendBlock, 0,
                             ctx->stream>>>(hc, o, n.n, n.dev);
            WF_HIP(hipGetLastError());
        }
        const Count ns{n.n, counts_dev + 2};
        if (jobs.defer) {
            // sized here so that a captured iteration finds the launch's counters allocated
            void* pc = nullptr;
            WF_TRY(workspace(ctx, "shadow_counter", sizeof(int32_t) * kWorkInts, &pc));
            jobs.defer->armed = true;
            jobs.defer->hc = hc;
            jobs.defer->o = o;
            jobs.defer->list = (int32_t*)plist;
            jobs.defer->n = ns;
            jobs.defer->accum = accum;
            if (alive_out) *alive_out = o.alive;
            return WF_OK;
        }
        void* pc = nullptr;
        WF_TRY(workspace(ctx, "shadow_counter", sizeof(int32_t) * kWorkInts, &pc));
        WF_TRY(wf_fill_u32(ctx, pc, 0u, kWorkInts));
        Timed tm(ctx, "shadow", 41.0, ns);
        with_bvh(s->dev, [&](auto nodes, auto S) {
            k_shadow<std::remove_cv_t<std::remove_pointer_t<decltype(nodes)>>, decltype(S)::value>
                <<<persistent_grid(ns.n, ctx, ctx->tuning.shadow_blocks_per_cu), kTraceBlock, 0,
                   ctx->stream>>>(
                    s->dev, nodes, hc, o, (int32_t*)plist, ns.n, ns.dev, accum, (int32_t*)pc,
                    shadow_refill(ctx));
        });
        WF_HIP(hipGetLastError());
        if (alive_out) *alive_out = o.alive;
        return WF_OK;
    }
    {
        Timed tm(ctx, "shade", 120.0, n);
        k_shade<<<grid_for(n.n, 128), 128, 0, ctx->stream>>>(s->dev, *q, n.n, n.dev, a,
                                                             s->dev.mat_rel_cost, accum, o,
                                                             jobs.tags, jobs.jobs);
        WF_HIP(hipGetLastError());
    }
    // shadow rays: compact the entries that emitted one, trace, accumulate
    k_flags_nonzero<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(o.shadow_flag, n.n, n.dev,
                                                                 (uint32_t*)pshflag);
    WF_HIP(hipGetLastError());
    int32_t* counts = counts_dev;
    WF_TRY(exclusive_scan_u32(ctx, (uint32_t*)pshflag, (uint32_t*)pprefix, n, counts + 2,
                              "compact"));
    k_compact_index<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(
        (uint32_t*)pshflag, (uint32_t*)pprefix, n.n, n.dev, (int32_t*)plist);
    WF_HIP(hipGetLastError());
    // a host count reads the number of shadow entries back to size the launch; a device count
    // launches over the capacity and lets entries past the count exit at once
    Count ns{n.n, counts + 2};
    if (!n.dev) {
        int32_t n_shadow = 0;
        WF_TRY(readback(ctx, &n_shadow, counts + 2, sizeof(int32_t)));
        ns = host_count(n_shadow);
    }
    WF_TRY(wf_fill_u32(ctx, counts + 1, 0u, 1));
    if (ns.n > 0) {
        k_count_shadow_rays<<<grid_for(ns.n, 256), 256, 0, ctx->stream>>>(
            o.shadow_flag, (int32_t*)plist, ns.n, ns.dev, counts + 1);
        WF_HIP(hipGetLastError());
        void* pc = nullptr;
        WF_TRY(workspace(ctx, "shadow_counter", sizeof(int32_t) * kWorkInts, &pc));
        WF_TRY(wf_fill_u32(ctx, pc, 0u, kWorkInts));
        Timed tm(ctx, "shadow", 41.0, ns);
        with_bvh(s->dev, [&](auto nodes, auto S) {
            k_shadow<std::remove_cv_t<std::remove_pointer_t<decltype(nodes)>>, decltype(S)::value>
                <<<persistent_grid(ns.n, ctx, ctx->tuning.shadow_blocks_per_cu), kTraceBlock, 0,
                   ctx->stream>>>(
                    s->dev, nodes, HitCols{q->px, q->py, q->pz, q->pixel, q->depth, jobs.tags},
                    o, (int32_t*)plist, ns.n, ns.dev, accum, (int32_t*)pc, shadow_refill(ctx));
        });
        WF_HIP(hipGetLastError());
    }
    // next rays: stable compaction in shading order
    WF_TRY(exclusive_scan_u32(ctx, o.alive, (uint32_t*)pprefix, n, counts, "compact"));
    {
        Timed tm(ctx, "next_rays", 76.0, n);
        k_next_rays<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(
            *q, o, (uint32_t*)pprefix, n.n, n.dev, *next, jobs.tags, jobs.next_tags);
        WF_HIP(hipGetLastError());
    }
    if (alive_out) *alive_out = o.alive;
    return WF_OK;
}

wf_status shade_misses(wf_context* ctx, const wf_scene* s, const wf_ray_soa* q, Count n,
                       int32_t physics, float* accum) {
    if (n.n <= 0) return WF_OK;
    Timed tm(ctx, "miss", 52.0, n);
    k_miss<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(s->dev, *q, n.n, n.dev, physics, accum);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

wf_status shade_hits(wf_context* ctx, const wf_scene* s, const HitSource& h, Count n,
                     const wf_shade_params* p, float* accum, int32_t* counts_dev,
                     const ShadeJobs& jobs, const uint32_t** alive_out, HitCols* hc_out) {
    WF_TRY(scene_ready(s));
    WF_REQUIRE(jobs.unordered && jobs.defer, "shade_hits: index mode needs deferred shadow walks");
    ShadeOut o;
    WF_TRY(shade_outputs(ctx, n.n, o));
    void* plist;
    WF_TRY(workspace(ctx, "shade_list", (size_t)(n.n > 0 ? n.n : 1) * 4, &plist));
    // sized here so that a captured iteration finds the next launch's counters allocated
    void* pc = nullptr;
    WF_TRY(workspace(ctx, "shadow_counter", sizeof(int32_t) * kWorkInts, &pc));
    ShadeArgs a{p->physics, p->rr_min_depth, p->cost_aware_rr, p->firefly_threshold, p->seed,
                p->sample_index, p->job};
    o.append = 1;
    o.shadow_list = (int32_t*)plist;
    o.next_count = counts_dev;
    o.shadow_rays = counts_dev + 1;
    o.shadow_entries = counts_dev + 2;
    o.next_packed = jobs.next_packed;
    o.next_tags = jobs.next_tags;
    o.next_keys = jobs.next_keys;
    o.key_frame = jobs.key_frame;
    if (!h.tags) o.htag = nullptr;
    WF_TRY(wf_fill_u32(ctx, counts_dev, 0u, 3));
    const HitCols hc{o.hpx, o.hpy, o.hpz, o.hpixel, o.hdepth, o.htag};
    if (n.n > 0) {
        {
            // list 4, hit buffers 16, packed ray 48, tag 4, triangle 48, UVs 32, mesh material
            // 4 read; hit columns 24, shadow record 44, next-ray state 32 written
            Timed tm(ctx, "shade", 256.0, n);
            k_shade_hits<<<grid_for(n.n, 128), 128, 0, ctx->stream>>>(
                s->dev, h, n.n, n.dev, a, s->dev.mat_rel_cost, accum, o, jobs.jobs);
            WF_HIP(hipGetLastError());
        }
        {
            Timed tm(ctx, "next_rays", o.next_keys ? 64.0 : 56.0, n);
            k_shade_append<<<grid_for(n.n, kAppendBlock * kAppendItems), kAppendBlock, 0,
                             ctx->stream>>>(hc, o, n.n, n.dev);
            WF_HIP(hipGetLastError());
        }
    }
    jobs.defer->armed = true;
    jobs.defer->hc = hc;
    jobs.defer->o = o;
    jobs.defer->list = (int32_t*)plist;
    jobs.defer->n = Count{n.n, counts_dev + 2};
    jobs.defer->accum = accum;
    if (alive_out) *alive_out = o.alive;
    if (hc_out) *hc_out = hc;
    return WF_OK;
}

}  // namespace wf

using namespace wf;

extern "C" {

wf_status wf_shade(const wf_scene* s, const wf_shade_soa* queue, int32_t n,
                   const wf_shade_params* params, float* accum_rgb, wf_ray_soa* next,
                   int32_t* n_next) {
    clear_error();
    WF_REQUIRE(s && queue && params && accum_rgb && next && n_next, "wf_shade: NULL argument");
    WF_REQUIRE(n >= 0, "wf_shade: negative count");
    WF_REQUIRE(params->physics == WF_PHYSICS_REFERENCE || params->physics == WF_PHYSICS_UNBIASED,
               "wf_shade: unknown physics %d", params->physics);
    void* pc = nullptr;
    WF_TRY(workspace(s->ctx, "shade_counts", sizeof(int32_t) * 4, &pc));
    WF_TRY(shade_bounce(s->ctx, s, queue, host_count(n), params, accum_rgb, next, (int32_t*)pc));
    int32_t c[2] = {0, 0};
    WF_TRY(readback(s->ctx, c, pc, sizeof(c)));
    *n_next = c[0];
    return WF_OK;
}

wf_status wf_shade_misses(const wf_scene* s, const wf_ray_soa* misses, int32_t n,
                          int32_t physics, float* accum_rgb) {
    clear_error();
    WF_REQUIRE(s && misses && accum_rgb, "wf_shade_misses: NULL argument");
    WF_REQUIRE(n >= 0, "wf_shade_misses: negative count");
    WF_TRY(shade_misses(s->ctx, s, misses, host_count(n), physics, accum_rgb));
    WF_HIP(hipStreamSynchronize(s->ctx->stream));
    return WF_OK;
}

}  // extern "C"
