// ORACLE -- test infrastructure only (see orc_core.h). Restatement of the reference's shading and
// frame loop: Material::{getSurfaceColor,scatter,emitted} (src/shading/Material.cpp:143-297),
// Texture::sample (include/shading/Texture.h:34-46), the spatial hash (include/math/SpatialHash.h),
// Scene::sampleLight (src/core/Scene.cpp:192-237), MIS helpers (include/shading/MIS.h:6-31),
// WavefrontRenderer::{renderScene,generatePrimaryRays,intersectAll,shadeAll,getSkyColor}
// (src/render/WavefrontRenderer.cpp:94-889), EnvironmentMap::sample (src/core/EnvironmentMap.cpp:30-48)
// and AdaptiveSampler (include/render/AdaptiveSampler.h:29-105).
//
// The reference draws uniforms from a thread_local mt19937; here every draw is addressed by
// (seed, job, pixel, sample, depth, dimension) through orc::uniform, the same stream the GPU uses.
// Per pixel, contributions are accumulated in the order emission, area NEE, directional NEE,
// miss, bounce after bounce -- the order the reference produces for one path.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>

#include "orc_core.h"
#include "wf_fmath.h"

namespace orc {
void hit_record(const Scene& s, const V3& o, const V3& d, float t, int32_t g, float bu, float bv,
                V3& point, V3& normal, float& u, float& v, int32_t& mat, int32_t& tex,
                uint8_t& front);
bool trace_closest(const Scene& s, const V3& o, const V3& d, float tnear, float tfar, float& t,
                   int32_t& g, float& bu, float& bv);
bool trace_occluded(const Scene& s, const V3& o, const V3& d, float max_dist);
}  // namespace orc

using namespace orc;

namespace {

const float PI = 3.14159265f;  // Material.cpp:6 and WavefrontRenderer.h:22

// Transcendentals. The restatement calls the single-precision sin/cos/asin/acos/atan2/x^5 that
// the device kernels call (wf_fmath.h), so a path's directions and weights are the same bits on
// both sides. While a replay entry point pins a sampling map against the reference's compiled
// Material.cpp, it calls the C library's functions instead, as the reference does
// (std::cos/std::sin/std::pow/std::acos, Material.cpp:26-37, 40-53, 268-297).
thread_local bool t_libm = false;
struct ReferenceLibm {
    bool saved;
    ReferenceLibm() : saved(t_libm) { t_libm = true; }
    ~ReferenceLibm() { t_libm = saved; }
};
float m_cos(float x) { return t_libm ? std::cos(x) : wf::fm::cos(x); }
float m_sin(float x) { return t_libm ? std::sin(x) : wf::fm::sin(x); }
float m_pow5(float x) { return t_libm ? std::pow(x, 5.0f) : wf::fm::pow5(x); }
float m_acos(float x) { return t_libm ? std::acos(x) : wf::fm::acos(x); }
float m_asin(float x) { return t_libm ? std::asin(x) : wf::fm::asin(x); }
float m_atan2(float y, float x) { return t_libm ? std::atan2(y, x) : wf::fm::atan2(y, x); }

// Draw source: the counter-based stream, or (for pinning against the compiled reference) a
// replayed sequence consumed in request order. Every call site below requests its draws in the
// reference's order, one statement per draw.
struct Rng {
    uint32_t seed, job, pixel, sample, depth;
    const float* replay = nullptr;
    int replay_n = 0;
    mutable int consumed = 0;
    float operator()(uint32_t dim) const {
        if (replay) {
            float x = consumed < replay_n ? replay[consumed] : 0.5f;
            ++consumed;
            return x;
        }
        return uniform(seed, job, pixel, sample, depth, dim);
    }
};

// ---- Material.cpp helpers ---------------------------------------------------------------------
V3 reflect(const V3& in, const V3& n) { return in - n * 2.0f * in.dot(n); }

V3 refract(const V3& d, const V3& n, float eta) {
    float cos_i = std::min(1.0f, (d * -1.0f).dot(n));
    V3 perp = (d + n * cos_i) * eta;
    V3 par = n * -std::sqrt(std::fabs(1.0f - perp.dot(perp)));
    return perp + par;
}

float schlick(float cos_i, float eta) {
    float r0 = (1.0f - eta) / (1.0f + eta);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * m_pow5(1.0f - cos_i);
}

void frame(const V3& n, V3& t, V3& b) {
    V3 up = std::fabs(n.x) > 0.9f ? V3(0, 1, 0) : V3(1, 0, 0);
    t = n.cross(up).normalized();
    b = n.cross(t);
}

V3 cosine_hemisphere(const V3& n, float u1, float u2) {
    float r = std::sqrt(u1);
    float ang = 2.0f * PI * u2;
    float x = r * m_cos(ang);
    float z = r * m_sin(ang);
    float y = std::sqrt(std::max(0.0f, 1.0f - r * r));
    V3 t, b;
    frame(n, t, b);
    return (t * x + n * y + b * z).normalized();
}

V3 in_unit_sphere(const Rng& rng) {
    V3 c;
    for (int k = 0; k < kSphereTries; ++k) {
        uint32_t base = DIM_SPHERE + 3u * (uint32_t)k;
        float x = rng(base) * 2.0f - 1.0f;
        float y = rng(base + 1) * 2.0f - 1.0f;
        float z = rng(base + 2) * 2.0f - 1.0f;
        c = V3(x, y, z);
        if (c.dot(c) < 1.0f) return c;
    }
    return c;
}

V3 tex_sample(const Texture& t, float u, float v) {
    u = u - std::floor(u);
    v = v - std::floor(v);
    int x = static_cast<int>(u * t.width);
    int y = static_cast<int>(v * t.height);
    x = std::max(0, std::min(x, t.width - 1));
    y = std::max(0, std::min(y, t.height - 1));
    return t.texels[(size_t)y * t.width + x];
}

uint32_t hash_grid_cell(int cx, int cy, int cz) {
    uint32_t h = static_cast<uint32_t>(cx) * 73856093u ^ static_cast<uint32_t>(cy) * 19349663u ^
                 static_cast<uint32_t>(cz) * 83492791u;
    return h * 2654435761u;
}

V3 surface_color(const Scene& s, const wf_material& m, const V3& p, float u, float v) {
    int nt = (int)s.textures.size();
    if (m.use_spatial_checker && m.texture_id >= 0 && m.texture_id < nt) {
        int cx = static_cast<int>(std::floor(p.x / m.checker_cell_size));
        int cy = static_cast<int>(std::floor(p.y / m.checker_cell_size));
        int cz = static_cast<int>(std::floor(p.z / m.checker_cell_size));
        uint32_t h = hash_grid_cell(cx, cy, cz);
        float pu = ((h >> 16) & 0xFFFF) / 65535.0f;
        float pv = (h & 0xFFFF) / 65535.0f;
        return tex_sample(s.textures[m.texture_id], pu, pv);
    }
    if (m.texture_id >= 0 && m.texture_id < nt) return tex_sample(s.textures[m.texture_id], u, v);
    return V3(m.albedo[0], m.albedo[1], m.albedo[2]);
}

V3 emitted(const wf_material& m, const V3& ln, const V3& dir_from_light) {
    V3 alb(m.albedo[0], m.albedo[1], m.albedo[2]);
    if (m.type == WF_MAT_EMISSIVE) return alb * m.emission;
    if (m.type == WF_MAT_SPOTLIGHT) {
        V3 radiance = alb * m.emission;
        float c = ln.dot(dir_from_light);
        float deg = m_acos(std::max(-1.0f, std::min(1.0f, c))) * (180.0f / PI);
        if (deg >= m.spot_outer_deg) return V3(0, 0, 0);
        if (deg > m.spot_falloff_deg) {
            float t = (deg - m.spot_falloff_deg) / (m.spot_outer_deg - m.spot_falloff_deg);
            float f = 1.0f - t;
            f = f * f * (3.0f - 2.0f * f);
            radiance = radiance * f;
        }
        return radiance;
    }
    return V3(0, 0, 0);
}

float area_to_angle(float area, float dist, float cosine) {
    float safe = std::max(std::fabs(cosine), 1e-4f);
    return area * (dist * dist) / safe;
}
float mis_balance(float a, float b) {
    float den = a + b;
    if (den <= 0.0f) return 0.0f;
    return a / den;
}
float mis_power(float a, float b) {
    float a2 = a * a, b2 = b * b;
    float den = a2 + b2;
    if (den <= 0.0f) return 0.0f;
    return a2 / den;
}
float light_area_pdf(const Scene& s) {
    return s.total_emissive_area <= 0.0f ? 0.0f : 1.0f / s.total_emissive_area;
}
V3 vmin(const V3& a, float c) { return V3(std::min(a.x, c), std::min(a.y, c), std::min(a.z, c)); }
bool any_positive(const V3& e) { return e.x > 0.0f || e.y > 0.0f || e.z > 0.0f; }

// ---- GGX conductor (Heitz 2018 visible-normal sampling; Walter et al. 2007 microfacet model) ----
float ggx_alpha(float roughness) { return std::max(roughness * roughness, 1e-4f); }
float ggx_d(const V3& m, float a) {
    float a2 = a * a;
    float k = (m.x * m.x + m.y * m.y) / a2 + m.z * m.z;
    return 1.0f / (PI * a2 * k * k);
}
float ggx_g1(const V3& w, float a) {
    float s2 = a * a * (w.x * w.x + w.y * w.y);
    return 2.0f * w.z / (w.z + std::sqrt(s2 + w.z * w.z));
}
V3 schlick_rgb(const V3& f0, float c) {
    float k = m_pow5(1.0f - std::max(0.0f, std::min(1.0f, c)));
    return V3(f0.x + (1.0f - f0.x) * k, f0.y + (1.0f - f0.y) * k, f0.z + (1.0f - f0.z) * k);
}
V3 to_local(const V3& w, const V3& t, const V3& b, const V3& n) {
    return V3(w.dot(t), w.dot(b), w.dot(n));
}
V3 to_world(const V3& w, const V3& t, const V3& b, const V3& n) {
    return t * w.x + b * w.y + n * w.z;
}
// f*cos and pdf for direction wi (world) seen from wo (world), both unit.
void ggx_eval(const V3& f0, float a, const V3& n, const V3& wo, const V3& wi, V3& fcos,
              float& pdf) {
    V3 t, b;
    frame(n, t, b);
    V3 lo = to_local(wo, t, b, n), li = to_local(wi, t, b, n);
    fcos = V3(0, 0, 0);
    pdf = 0.0f;
    if (lo.z <= 0.0f || li.z <= 0.0f) return;
    V3 m = (lo + li).normalized();
    float d = ggx_d(m, a);
    float g1o = ggx_g1(lo, a);
    float g1i = ggx_g1(li, a);
    V3 F = schlick_rgb(f0, lo.dot(m));
    float k = d * g1o * g1i / (4.0f * lo.z);
    fcos = F * k;
    pdf = d * g1o / (4.0f * lo.z);
}
bool ggx_sample(const V3& f0, float a, const V3& n, const V3& wo, float u1, float u2, V3& wi,
                V3& weight, float& pdf) {
    V3 t, b;
    frame(n, t, b);
    V3 lo = to_local(wo, t, b, n);
    if (lo.z <= 0.0f) return false;
    V3 vh = V3(a * lo.x, a * lo.y, lo.z).normalized();
    float lensq = vh.x * vh.x + vh.y * vh.y;
    V3 t1 = lensq > 0.0f ? V3(-vh.y, vh.x, 0.0f) / std::sqrt(lensq) : V3(1, 0, 0);
    V3 t2 = vh.cross(t1);
    float r = std::sqrt(u1);
    float phi = 2.0f * PI * u2;
    float p1 = r * m_cos(phi);
    float p2 = r * m_sin(phi);
    float sh = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - sh) * std::sqrt(std::max(0.0f, 1.0f - p1 * p1)) + sh * p2;
    V3 nh = t1 * p1 + t2 * p2 + vh * std::sqrt(std::max(0.0f, 1.0f - p1 * p1 - p2 * p2));
    V3 m = V3(a * nh.x, a * nh.y, std::max(1e-6f, nh.z)).normalized();
    float om = lo.dot(m);
    V3 li = m * (2.0f * om) - lo;
    if (li.z <= 0.0f) return false;
    V3 F = schlick_rgb(f0, om);
    weight = F * ggx_g1(li, a);
    pdf = ggx_d(m, a) * ggx_g1(lo, a) / (4.0f * lo.z);
    wi = to_world(li, t, b, n).normalized();
    return true;
}

// Exact dielectric Fresnel reflectance for eta = n_incident / n_transmitted.
float fresnel_dielectric(float cos_i, float eta, float& cos_t) {
    float sin2t = eta * eta * (1.0f - cos_i * cos_i);
    if (sin2t >= 1.0f) {
        cos_t = 0.0f;
        return 1.0f;
    }
    cos_t = std::sqrt(1.0f - sin2t);
    float rs = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float rp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (rs * rs + rp * rp);
}

// Material::scatter, Material.cpp:165-266 (reference model) plus the microfacet models.
bool scatter(const Scene& s, const wf_material& m, const V3& d, const V3& n, const V3& p, float u,
             float v, uint8_t front, const Rng& rng, V3& att, V3& out_dir, float& out_pdf) {
    V3 sc = surface_color(s, m, p, u, v);
    switch (m.type) {
    case WF_MAT_DIFFUSE: {
        float u1 = rng(DIM_DIR_U1);
        float u2 = rng(DIM_DIR_U2);
        V3 dir = cosine_hemisphere(n, u1, u2);
        out_dir = dir;
        att = sc;
        float c = std::max(0.0f, n.dot(dir));
        out_pdf = c / PI;
        return true;
    }
    case WF_MAT_METAL: {
        if (m.model == WF_MODEL_MICROFACET) {
            V3 wo = (d * -1.0f).normalized();
            V3 w;
            float pdf;
            float u1 = rng(DIM_DIR_U1);
            float u2 = rng(DIM_DIR_U2);
            if (!ggx_sample(sc, ggx_alpha(m.roughness), n, wo, u1, u2, out_dir, w, pdf))
                return false;
            att = w;
            out_pdf = pdf;
            return true;
        }
        V3 reflected = reflect(d, n);
        V3 fuzz = in_unit_sphere(rng) * m.roughness;
        out_dir = (reflected + fuzz).normalized();
        att = sc;
        out_pdf = -1.0f;
        return out_dir.dot(n) > 0.0f;
    }
    case WF_MAT_EMISSIVE:
    case WF_MAT_SPOTLIGHT:
        att = sc * m.emission;
        out_pdf = -1.0f;
        return false;
    case WF_MAT_GLASS: {
        att = sc;
        out_pdf = -1.0f;
        float eta = front ? (1.0f / m.ior) : m.ior;
        V3 id = d.normalized();
        float cos_i = std::min(1.0f, (id * -1.0f).dot(n));
        if (m.model == WF_MODEL_MICROFACET) {
            float cos_t;
            float F = fresnel_dielectric(cos_i, eta, cos_t);
            if (rng(DIM_LOBE) < F)
                out_dir = reflect(id, n);
            else
                out_dir = id * eta + n * (eta * cos_i - cos_t);
            return true;
        }
        float sin_i = std::sqrt(1.0f - cos_i * cos_i);
        bool tir = eta * sin_i > 1.0f;
        if (tir || schlick(cos_i, eta) > rng(DIM_LOBE))
            out_dir = reflect(id, n);
        else
            out_dir = refract(id, n, eta);
        return true;
    }
    case WF_MAT_PLASTIC: {
        float cos_i = std::max(0.0f, (d * -1.0f).dot(n));
        float R = schlick(cos_i, 1.5f);
        if (rng(DIM_LOBE) < R) {
            V3 reflected = reflect(d.normalized(), n);
            V3 fuzz = in_unit_sphere(rng) * m.roughness;
            out_dir = (reflected + fuzz).normalized();
            att = V3(1.0f, 1.0f, 1.0f);
            out_pdf = -1.0f;
        } else {
            float u1 = rng(DIM_DIR_U1);
            float u2 = rng(DIM_DIR_U2);
            V3 dir = cosine_hemisphere(n, u1, u2);
            out_dir = dir;
            att = sc;
            float c = std::max(0.0f, n.dot(dir));
            out_pdf = c / PI;
        }
        return true;
    }
    default:
        out_pdf = -1.0f;
        return false;
    }
}

// getSkyColor, cpp:879-889, with EnvironmentMap::sample, EnvironmentMap.cpp:30-48.
int env_texel(const Scene& s, const V3& dir, float& lat_cos) {
    V3 d = dir.normalized();
    float u = 0.5f + m_atan2(d.z, d.x) / (2.0f * static_cast<float>(M_PI));
    float v = 0.5f - m_asin(std::max(-1.0f, std::min(1.0f, d.y))) / static_cast<float>(M_PI);
    int x = std::min(s.env_w - 1, std::max(0, static_cast<int>(u * s.env_w)));
    int y = std::min(s.env_h - 1, std::max(0, static_cast<int>(v * s.env_h)));
    lat_cos = std::sqrt(std::max(0.0f, 1.0f - d.y * d.y));
    return y * s.env_w + x;
}
V3 env_radiance(const Scene& s, int texel) {
    size_t i = (size_t)texel * 3;
    V3 raw(s.env_rgb[i], s.env_rgb[i + 1], s.env_rgb[i + 2]);
    return raw * s.env_intensity * s.env_tint;
}
V3 sky(const Scene& s, const V3& dir) {
    if (s.env_w > 0) {
        float lc;
        return env_radiance(s, env_texel(s, dir, lc));
    }
    float blend = 0.5f * (dir.normalized().y + 1.0f);
    V3 dark(0.02f, 0.02f, 0.05f), horizon(0.05f, 0.04f, 0.03f);
    return horizon * (1.0f - blend) + dark * blend;
}
// Solid-angle density of environment NEE for a direction with texel i and cos(latitude) lc.
float env_pdf_sa(const Scene& s, int texel, float lc) {
    if (s.env_choice_prob <= 0.0f || lc <= 0.0f) return 0.0f;
    float p = s.env_choice_prob * s.env_pdf[texel] * (float)s.env_w * (float)s.env_h;
    return p / (2.0f * PI * PI * lc);
}

// shadeAll body for one hit (cpp:602-797). acc: the pixel's accumulator.
template <typename Acc>
bool shade_one(const Scene& s, const wf_shade_params& sp, const double* mat_cost, const V3& P,
               const V3& N, float t, float u, float v, int32_t mat_id, uint8_t front, const V3& O,
               const V3& D, const V3& T, int32_t pixel, int32_t depth, float pdf_in, Acc&& acc,
               V3& next_o, V3& next_d, V3& next_t, float& next_pdf, int64_t& shadow_rays) {
    (void)O;
    const wf_material& m = s.materials[mat_id];
    const bool unbiased = sp.physics == WF_PHYSICS_UNBIASED;
    Rng rng{sp.seed, (uint32_t)sp.job, (uint32_t)pixel, (uint32_t)sp.sample_index,
            (uint32_t)depth};
    const float clampv = sp.firefly_threshold;
    // emission + MIS (cpp:629-658)
    V3 e = emitted(m, N, D * -1.0f);
    if (any_positive(e)) {
        float w = 1.0f;
        if (depth > 0 && pdf_in > 0.0f) {
            float cos_l = std::max(0.0f, (D * -1.0f).dot(N));
            if (cos_l > 0.0f) {
                if (unbiased) {
                    float pa = s.mat_light_pdf_area[mat_id] * (1.0f - s.env_choice_prob);
                    w = mis_power(pdf_in, area_to_angle(pa, t, cos_l));
                } else {
                    w = mis_balance(pdf_in, area_to_angle(light_area_pdf(s), t, cos_l));
                }
            }
        }
        e = vmin(e, clampv);
        acc(T * e * w);
    }
    // next-event estimation
    const bool ggx = m.type == WF_MAT_METAL && m.model == WF_MODEL_MICROFACET;
    const bool nee = m.type == WF_MAT_DIFFUSE || m.type == WF_MAT_PLASTIC || (unbiased && ggx);
    if (nee && !unbiased) {
        // Scene::sampleLight, Scene.cpp:192-237
        if (!s.emissive_meshes.empty()) {
            int ne = (int)s.emissive_meshes.size();
            int mi = s.emissive_meshes[static_cast<int>(rng(DIM_LIGHT_PICK) * ne) % ne];
            int tc = s.mesh_tri_count[mi];
            if (tc > 0) {
                int k = static_cast<int>(rng(DIM_LIGHT_TRI) * tc) % tc;
                int g = s.mesh_tri_offset[mi] + k;
                float r1 = rng(DIM_LIGHT_R1), r2 = rng(DIM_LIGHT_R2);
                if (r1 + r2 > 1.0f) {
                    r1 = 1.0f - r1;
                    r2 = 1.0f - r2;
                }
                float r3 = 1.0f - r1 - r2;
                V3 lp = s.v0[g] * r3 + s.v1[g] * r1 + s.v2[g] * r2;
                V3 ln = (s.v1[g] - s.v0[g]).cross(s.v2[g] - s.v0[g]).normalized();
                V3 dfl = (P - lp).normalized();
                V3 le = emitted(s.materials[s.mesh_material[mi]], ln, dfl);
                float larea = s.total_emissive_area;
                V3 to_l = lp - P;
                float dist = to_l.length();
                V3 dir_l = to_l / dist;
                float cos_s = N.dot(dir_l);
                float cos_l = (dir_l * -1.0f).dot(ln);
                if (cos_s > 0.0f && cos_l > 0.0f) {
                    ++shadow_rays;
                    if (!trace_occluded(s, P, dir_l, dist)) {
                        V3 sc = surface_color(s, m, P, u, v);
                        float geometry = (cos_s * cos_l) / (dist * dist);
                        V3 brdf = sc * (1.0f / PI);
                        float lps = area_to_angle(light_area_pdf(s), dist, cos_l);
                        float bps = cos_s / PI;
                        float w = mis_balance(lps, bps);
                        V3 direct = T * brdf * le * geometry * larea * w;
                        acc(vmin(direct, clampv));
                    }
                }
            }
        }
    } else if (nee && unbiased) {
        bool use_env = s.env_choice_prob > 0.0f &&
                       (s.light_tris.empty() || rng(DIM_ENV_CHOICE) < s.env_choice_prob);
        V3 wo = (D * -1.0f).normalized();
        V3 sc = surface_color(s, m, P, u, v);
        auto eval = [&](const V3& wi, V3& fcos, float& bpdf) {
            if (ggx) {
                ggx_eval(sc, ggx_alpha(m.roughness), N, wo, wi, fcos, bpdf);
            } else {
                float c = N.dot(wi);
                fcos = sc * (1.0f / PI) * std::max(0.0f, c);
                bpdf = std::max(0.0f, c) / PI;
            }
        };
        if (use_env) {
            int nt = s.env_w * s.env_h;
            int slot = std::min(nt - 1, static_cast<int>(rng(DIM_ENV_SLOT) * nt));
            if (rng(DIM_ENV_COIN) >= s.env_prob[slot]) slot = s.env_alias[slot];
            int tx = slot % s.env_w, ty = slot / s.env_w;
            float eu = ((float)tx + rng(DIM_ENV_JX)) / (float)s.env_w;
            float ev = ((float)ty + rng(DIM_ENV_JY)) / (float)s.env_h;
            float phi = (eu - 0.5f) * (2.0f * PI);
            float lat = (0.5f - ev) * PI;
            float cl = m_cos(lat);
            V3 wi(cl * m_cos(phi), m_sin(lat), cl * m_sin(phi));
            float cos_s = N.dot(wi);
            if (cos_s > 0.0f) {
                float lc;
                int texel = env_texel(s, wi, lc);
                float lpdf = env_pdf_sa(s, texel, lc);
                if (lpdf > 0.0f) {
                    V3 fcos;
                    float bpdf;
                    eval(wi, fcos, bpdf);
                    ++shadow_rays;
                    if (!trace_occluded(s, P, wi, 1e6f)) {
                        float w = mis_power(lpdf, bpdf);
                        V3 direct = T * fcos * env_radiance(s, texel) * (w / lpdf);
                        acc(vmin(direct, clampv));
                    }
                }
            }
        } else if (!s.light_tris.empty()) {
            int nl = (int)s.light_tris.size();
            int slot = std::min(nl - 1, static_cast<int>(rng(DIM_LIGHT_PICK) * nl));
            if (rng(DIM_LIGHT_TRI) >= s.light_prob[slot]) slot = s.light_alias[slot];
            int g = s.light_tris[slot];
            float r1 = rng(DIM_LIGHT_R1), r2 = rng(DIM_LIGHT_R2);
            if (r1 + r2 > 1.0f) {
                r1 = 1.0f - r1;
                r2 = 1.0f - r2;
            }
            float r3 = 1.0f - r1 - r2;
            V3 lp = s.v0[g] * r3 + s.v1[g] * r1 + s.v2[g] * r2;
            V3 ln = (s.v1[g] - s.v0[g]).cross(s.v2[g] - s.v0[g]).normalized();
            int lmat = s.mesh_material[s.tri_geom[g]];
            V3 to_l = lp - P;
            float dist = to_l.length();
            V3 dir_l = to_l / dist;
            float cos_s = N.dot(dir_l);
            float cos_l = std::fabs((dir_l * -1.0f).dot(ln));
            if (cos_s > 0.0f && cos_l > 0.0f) {
                V3 le = emitted(s.materials[lmat], ln, dir_l * -1.0f);
                float pa = s.mat_light_pdf_area[lmat] * (1.0f - s.env_choice_prob);
                float lpdf = area_to_angle(pa, dist, cos_l);
                if (lpdf > 0.0f) {
                    V3 fcos;
                    float bpdf;
                    eval(dir_l, fcos, bpdf);
                    ++shadow_rays;
                    if (!trace_occluded(s, P, dir_l, dist)) {
                        float w = mis_power(lpdf, bpdf);
                        V3 direct = T * fcos * le * (w / lpdf);
                        acc(vmin(direct, clampv));
                    }
                }
            }
        }
    }
    // directional light (cpp:712-737), applied to every material type
    if (s.directional.intensity > 0.0f) {
        V3 ldir(s.directional.direction[0], s.directional.direction[1],
                s.directional.direction[2]);
        V3 to_l = ldir * -1.0f;
        float cos_s = N.dot(to_l);
        if (cos_s > 0.0f) {
            ++shadow_rays;
            if (!trace_occluded(s, P, to_l, 1e6f)) {
                V3 sc = surface_color(s, m, P, u, v);
                V3 brdf = sc * (1.0f / PI);
                V3 col(s.directional.color[0], s.directional.color[1], s.directional.color[2]);
                V3 direct = T * brdf * col * s.directional.intensity * cos_s;
                acc(vmin(direct, clampv));
            }
        }
    }
    // scatter + Russian roulette (cpp:739-796)
    V3 att, dir;
    float spdf;
    if (!scatter(s, m, D, N, P, u, v, front, rng, att, dir, spdf)) return false;
    V3 nt = T * att;
    if (depth + 1 >= sp.rr_min_depth) {
        float base = std::min(0.95f, std::max({nt.x, nt.y, nt.z}));
        float surv = base;
        if (sp.cost_aware_rr) {
            float rc = static_cast<float>(mat_cost ? mat_cost[mat_id] : 1.0);
            surv = std::clamp(base / rc, 0.05f, 0.95f);
        }
        if (rng(DIM_RR) > surv) return false;
        nt = nt / surv;
    }
    next_o = P;
    next_d = dir;
    next_t = nt;
    next_pdf = spdf;
    return true;
}

template <typename F>
void parallel_for(int n, int threads, F&& fn) {
    if (threads <= 1 || n < 4096) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<int> next{0};
    const int chunk = 1024;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            for (;;) {
                int b = next.fetch_add(chunk);
                if (b >= n) break;
                fn(b, std::min(n, b + chunk));
            }
        });
    for (auto& th : pool) th.join();
}

struct RayQ {
    std::vector<float> ox, oy, oz, dx, dy, dz, tr, tg, tb, pdf;
    std::vector<int32_t> pixel, depth;
    void resize(size_t n) {
        for (auto* v : {&ox, &oy, &oz, &dx, &dy, &dz, &tr, &tg, &tb, &pdf}) v->resize(n);
        pixel.resize(n);
        depth.resize(n);
    }
    size_t size() const { return ox.size(); }
    void push(const V3& o, const V3& d, const V3& t, int32_t p, int32_t dp, float pd) {
        ox.push_back(o.x);
        oy.push_back(o.y);
        oz.push_back(o.z);
        dx.push_back(d.x);
        dy.push_back(d.y);
        dz.push_back(d.z);
        tr.push_back(t.x);
        tg.push_back(t.y);
        tb.push_back(t.z);
        pixel.push_back(p);
        depth.push_back(dp);
        pdf.push_back(pd);
    }
    void gather(const std::vector<int32_t>& idx) {
        RayQ o;
        o.resize(idx.size());
        for (size_t i = 0; i < idx.size(); ++i) {
            size_t j = (size_t)idx[i];
            o.ox[i] = ox[j];
            o.oy[i] = oy[j];
            o.oz[i] = oz[j];
            o.dx[i] = dx[j];
            o.dy[i] = dy[j];
            o.dz[i] = dz[j];
            o.tr[i] = tr[j];
            o.tg[i] = tg[j];
            o.tb[i] = tb[j];
            o.pixel[i] = pixel[j];
            o.depth[i] = depth[j];
            o.pdf[i] = pdf[j];
        }
        *this = std::move(o);
    }
};

}  // namespace

extern "C" {
void orc_ray_schedule(int32_t n, const float* ox, const float* oy, const float* oz,
                      const float* dx, const float* dy, const float* dz, const float* bmin,
                      const float* bmax, int32_t* sorted);
int32_t orc_shading_schedule(int32_t policy, int32_t n, const int32_t* mat, const int32_t* tex,
                             const float* tr, const float* tg, const float* tb, const float* px,
                             const float* py, const float* pz, int32_t n_mats,
                             const int32_t* mat_type, const int32_t* mat_tex, int32_t n_tex,
                             const uint64_t* tex_mem, const double* mat_cost,
                             const double* tex_cost, int32_t* sorted);

// Material::scatter with the draws replayed in request order (pins the uniform->direction maps
// against the compiled reference). Returns didScatter; *consumed = draws used.
int32_t orc_scatter_replay(void* sp, const wf_material* m, const float* d, const float* n,
                           const float* p, float u, float v, int32_t front, const float* draws,
                           int32_t n_draws, float* att, float* dir, float* pdf,
                           int32_t* consumed) {
    const Scene& s = *(Scene*)sp;
    ReferenceLibm libm;
    Rng rng{0, 0, 0, 0, 0, draws, n_draws, 0};
    V3 a, o;
    float pd = 0.0f;
    bool ok = scatter(s, *m, V3(d[0], d[1], d[2]), V3(n[0], n[1], n[2]), V3(p[0], p[1], p[2]), u,
                      v, (uint8_t)front, rng, a, o, pd);
    att[0] = a.x; att[1] = a.y; att[2] = a.z;
    dir[0] = o.x; dir[1] = o.y; dir[2] = o.z;
    *pdf = pd;
    *consumed = rng.consumed;
    return ok ? 1 : 0;
}

// Unbiased-track building blocks, exported for closed-form tests.
int32_t orc_ggx_sample(const float* f0, float roughness, const float* n, const float* wo, float u1,
                       float u2, float* wi, float* weight, float* pdf) {
    V3 w, wt;
    float p = 0.0f;
    bool ok = ggx_sample(V3(f0[0], f0[1], f0[2]), ggx_alpha(roughness), V3(n[0], n[1], n[2]),
                         V3(wo[0], wo[1], wo[2]), u1, u2, w, wt, p);
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    weight[0] = wt.x; weight[1] = wt.y; weight[2] = wt.z;
    *pdf = p;
    return ok ? 1 : 0;
}

void orc_ggx_eval(const float* f0, float roughness, const float* n, const float* wo,
                  const float* wi, float* fcos, float* pdf) {
    V3 f;
    float p;
    ggx_eval(V3(f0[0], f0[1], f0[2]), ggx_alpha(roughness), V3(n[0], n[1], n[2]),
             V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2]), f, p);
    fcos[0] = f.x; fcos[1] = f.y; fcos[2] = f.z;
    *pdf = p;
}

float orc_fresnel_dielectric(float cos_i, float eta) {
    float ct;
    return fresnel_dielectric(cos_i, eta, ct);
}

float orc_mis_power(float a, float b) { return mis_power(a, b); }
float orc_area_to_angle(float area, float dist, float c) { return area_to_angle(area, dist, c); }
float orc_mis_balance(float a, float b) { return mis_balance(a, b); }

void orc_cosine_hemisphere(const float* n, float u1, float u2, float* out) {
    V3 d = cosine_hemisphere(V3(n[0], n[1], n[2]), u1, u2);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

// Light tables of the unbiased track: emitter alias table and per-material area pdf.
int32_t orc_light_tables(void* sp, float* prob, int32_t* alias, int32_t* tris, float* mat_pdf,
                         int32_t cap) {
    const Scene& s = *(Scene*)sp;
    int n = (int)s.light_tris.size();
    for (int i = 0; i < n && i < cap; ++i) {
        prob[i] = s.light_prob[i];
        alias[i] = s.light_alias[i];
        tris[i] = s.light_tris[i];
    }
    for (size_t m = 0; m < s.mat_light_pdf_area.size(); ++m) mat_pdf[m] = s.mat_light_pdf_area[m];
    return n;
}

int32_t orc_env_tables(void* sp, float* prob, int32_t* alias, float* pdf, float* choice) {
    const Scene& s = *(Scene*)sp;
    int n = (int)s.env_prob.size();
    for (int i = 0; i < n; ++i) {
        prob[i] = s.env_prob[i];
        alias[i] = s.env_alias[i];
        pdf[i] = s.env_pdf[i];
    }
    *choice = s.env_choice_prob;
    return n;
}

// Material::emitted (Material.cpp:268-297).
void orc_emitted(const wf_material* m, const float* ln, const float* dfl, float* out) {
    ReferenceLibm libm;
    V3 e = emitted(*m, V3(ln[0], ln[1], ln[2]), V3(dfl[0], dfl[1], dfl[2]));
    out[0] = e.x; out[1] = e.y; out[2] = e.z;
}

// Material::getSurfaceColor (Material.cpp:143-163) against the scene's textures.
void orc_surface_color(void* sp, const wf_material* m, const float* p, float u, float v,
                       float* out) {
    V3 c = surface_color(*(Scene*)sp, *m, V3(p[0], p[1], p[2]), u, v);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

uint32_t orc_hash_grid_cell(int32_t x, int32_t y, int32_t z) { return hash_grid_cell(x, y, z); }

// Markstein's reciprocal correction y + y (1 - b y) (two fmas) over every significand of the
// binade [2^e, 2^(e+1)) and every faithful approximation y of 1/b (|y - 1/b| < 1 ulp): the number
// of (b, y) pairs whose corrected value differs from the correctly rounded 1/b, and the last such
// b's bits. The device traversal corrects the hardware reciprocal (faithful) this way.
int64_t orc_reciprocal_correction_failures(int32_t e, uint32_t* last_bad_bits) {
    int64_t bad = 0;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        const uint32_t bits = ((uint32_t)(e + 127) << 23) | m;
        float b;
        std::memcpy(&b, &bits, 4);
        const float cr = 1.0f / b;
        const double exact = 1.0 / (double)b;
        uint32_t cb;
        std::memcpy(&cb, &cr, 4);
        for (int k = -2; k <= 2; ++k) {
            const uint32_t yb = cb + (uint32_t)k;
            float y, y1;
            std::memcpy(&y, &yb, 4);
            const uint32_t yb1 = yb + 1u;
            std::memcpy(&y1, &yb1, 4);
            if (std::fabs((double)y - exact) >= (double)(y1 - y)) continue;   // not faithful
            const float z = std::fma(std::fma(-b, y, 1.0f), y, y);
            if (z != cr) {
                ++bad;
                *last_bad_bits = bits;
            }
        }
    }
    return bad;
}

// The shared single-precision functions, exported for their accuracy tests.
void orc_fm_sincos(int32_t n, const float* x, float* s, float* c) {
    for (int32_t i = 0; i < n; ++i) wf::fm::sincos(x[i], s[i], c[i]);
}
void orc_fm_inverse(int32_t n, const float* x, const float* y, float* as, float* ac, float* at) {
    for (int32_t i = 0; i < n; ++i) {
        as[i] = wf::fm::asin(x[i]);
        ac[i] = wf::fm::acos(x[i]);
        at[i] = wf::fm::atan2(y[i], x[i]);
    }
}

float orc_uniform(uint32_t seed, uint32_t job, uint32_t pixel, uint32_t sample, uint32_t depth,
                  uint32_t dim) {
    return uniform(seed, job, pixel, sample, depth, dim);
}

void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    philox4x32_10(ctr, key, out);
}

// Camera constructor (include/core/Camera.h:13-33).
struct CamBasis {
    V3 from, tl, horizontal, vertical;
};
static CamBasis camera_basis(const wf_camera* cam, int32_t width, int32_t height) {
    V3 from(cam->eye[0], cam->eye[1], cam->eye[2]), at(cam->target[0], cam->target[1],
                                                       cam->target[2]);
    V3 upv(cam->up[0], cam->up[1], cam->up[2]);
    float aspect = static_cast<float>(width) / height;
    float theta = cam->vfov_deg * 3.14159265f / 180.0f;
    float vh = 2.0f * std::tan(theta / 2.0f);
    float vw = aspect * vh;
    V3 fwd = (from - at).normalized();
    V3 right = upv.cross(fwd).normalized();
    V3 up = fwd.cross(right);
    V3 horizontal = right * vw, vertical = up * vh;
    V3 tl = from - horizontal / 2.0f - vertical / 2.0f - fwd;
    return CamBasis{from, tl, horizontal, vertical};
}

// Camera::getRay (Camera.h:35-40).
void orc_camera_ray(const wf_camera* cam, int32_t width, int32_t height, float u, float v,
                    float* o, float* d) {
    CamBasis c = camera_basis(cam, width, height);
    V3 dir = (c.tl + c.horizontal * u + c.vertical * v - c.from).normalized();
    o[0] = c.from.x; o[1] = c.from.y; o[2] = c.from.z;
    d[0] = dir.x; d[1] = dir.y; d[2] = dir.z;
}

// generatePrimaryRays (cpp:401-457): jittered camera rays for the given pixels, in order.
void orc_raygen(const wf_camera* cam, int32_t width, int32_t height, int32_t n,
                const int32_t* pixels, uint32_t seed, int32_t job, int32_t sample, float* ox,
                float* oy, float* oz, float* dx, float* dy, float* dz) {
    CamBasis c = camera_basis(cam, width, height);
    const V3 from = c.from, tl = c.tl, horizontal = c.horizontal, vertical = c.vertical;
    for (int32_t i = 0; i < n; ++i) {
        int32_t p = pixels[i];
        int x = p % width, y = p / width;
        float jx = uniform(seed, (uint32_t)job, (uint32_t)p, (uint32_t)sample, kRaygenDepth, 0);
        float jy = uniform(seed, (uint32_t)job, (uint32_t)p, (uint32_t)sample, kRaygenDepth, 1);
        float u = (static_cast<float>(x) + jx) / (width - 1);
        float v = 1.0f - (static_cast<float>(y) + jy) / (height - 1);
        V3 d = (tl + horizontal * u + vertical * v - from).normalized();
        ox[i] = from.x;
        oy[i] = from.y;
        oz[i] = from.z;
        dx[i] = d.x;
        dy[i] = d.y;
        dz[i] = d.z;
    }
}

// WavefrontRenderer::intersectAll (cpp:459-567) for one queue: closest hit through the CPU BVH,
// HitRecord mapping, then the hit ShadingQueue and miss RayQueue in ascending input order.
// Output arrays are sized n by the caller; counts returned.
void orc_intersect_split(void* sp, int32_t n, const wf_ray_soa* in, wf_shade_soa* hits,
                         int32_t* n_hits, wf_ray_soa* miss, int32_t* n_miss, int32_t threads) {
    const Scene& s = *(Scene*)sp;
    std::vector<float> t(n), bu(n), bv(n);
    std::vector<int32_t> g(n);
    parallel_for(n, threads, [&](int b, int e) {
        for (int i = b; i < e; ++i) {
            V3 o(in->ox[i], in->oy[i], in->oz[i]), d(in->dx[i], in->dy[i], in->dz[i]);
            if (!trace_closest(s, o, d, 0.001f, 1e9f, t[i], g[i], bu[i], bv[i])) g[i] = -1;
        }
    });
    int32_t h = 0, m = 0;
    for (int i = 0; i < n; ++i) {
        if (g[i] >= 0) {
            V3 o(in->ox[i], in->oy[i], in->oz[i]), d(in->dx[i], in->dy[i], in->dz[i]);
            V3 P, N;
            float u, v;
            int32_t mat, tex;
            uint8_t front;
            hit_record(s, o, d, t[i], g[i], bu[i], bv[i], P, N, u, v, mat, tex, front);
            hits->px[h] = P.x; hits->py[h] = P.y; hits->pz[h] = P.z;
            hits->nx[h] = N.x; hits->ny[h] = N.y; hits->nz[h] = N.z;
            hits->t[h] = t[i]; hits->u[h] = u; hits->v[h] = v;
            hits->tri[h] = s.tri_prim[g[i]];
            hits->mat[h] = mat;
            hits->tex[h] = tex;
            hits->ox[h] = in->ox[i]; hits->oy[h] = in->oy[i]; hits->oz[h] = in->oz[i];
            hits->dx[h] = in->dx[i]; hits->dy[h] = in->dy[i]; hits->dz[h] = in->dz[i];
            hits->tr[h] = in->tr[i]; hits->tg[h] = in->tg[i]; hits->tb[h] = in->tb[i];
            hits->pixel[h] = in->pixel[i];
            hits->depth[h] = in->depth[i];
            hits->pdf[h] = in->pdf[i];
            if (hits->geom) hits->geom[h] = s.tri_geom[g[i]];
            if (hits->front) hits->front[h] = front;
            ++h;
        } else {
            miss->ox[m] = in->ox[i]; miss->oy[m] = in->oy[i]; miss->oz[m] = in->oz[i];
            miss->dx[m] = in->dx[i]; miss->dy[m] = in->dy[i]; miss->dz[m] = in->dz[i];
            miss->tr[m] = in->tr[i]; miss->tg[m] = in->tg[i]; miss->tb[m] = in->tb[i];
            miss->pixel[m] = in->pixel[i];
            miss->depth[m] = in->depth[i];
            miss->pdf[m] = in->pdf[i];
            ++m;
        }
    }
    *n_hits = h;
    *n_miss = m;
}

// shadeAll (cpp:569-877) over a shading queue already in shading order: accumulates into
// accum_rgb (pixel-indexed float RGB) and appends survivors to next in queue order.
void orc_shade(void* sp, const wf_shade_soa* q, int32_t n, const wf_shade_params* prm,
               const double* mat_cost, float* accum_rgb, double* accum64, wf_ray_soa* next,
               int32_t* n_next, int64_t* shadow_rays, int32_t threads) {
    const Scene& s = *(Scene*)sp;
    std::vector<uint8_t> alive(n);
    std::vector<float> no(n * 3), nd(n * 3), ntt(n * 3), npdf(n);
    std::atomic<int64_t> shadows{0};
    parallel_for(n, threads, [&](int b, int e) {
        int64_t local_shadow = 0;
        for (int i = b; i < e; ++i) {
            int32_t pix = q->pixel[i];
            float* a = accum_rgb + (size_t)pix * 3;
            double* a64 = accum64 ? accum64 + (size_t)pix * 3 : nullptr;
            V3 o, d, t;
            float pd = 0.0f;   // set when the path continues
            alive[i] = shade_one(
                s, *prm, mat_cost, V3(q->px[i], q->py[i], q->pz[i]),
                V3(q->nx[i], q->ny[i], q->nz[i]), q->t[i], q->u[i], q->v[i], q->mat[i],
                q->front ? q->front[i] : 1, V3(q->ox[i], q->oy[i], q->oz[i]),
                V3(q->dx[i], q->dy[i], q->dz[i]), V3(q->tr[i], q->tg[i], q->tb[i]), pix,
                q->depth[i], q->pdf[i],
                [&](const V3& c) {
                    a[0] = a[0] + c.x;
                    a[1] = a[1] + c.y;
                    a[2] = a[2] + c.z;
                    if (a64) {
                        a64[0] += c.x;
                        a64[1] += c.y;
                        a64[2] += c.z;
                    }
                },
                o, d, t, pd, local_shadow);
            if (alive[i]) {
                no[i * 3] = o.x; no[i * 3 + 1] = o.y; no[i * 3 + 2] = o.z;
                nd[i * 3] = d.x; nd[i * 3 + 1] = d.y; nd[i * 3 + 2] = d.z;
                ntt[i * 3] = t.x; ntt[i * 3 + 1] = t.y; ntt[i * 3 + 2] = t.z;
                npdf[i] = pd;
            }
        }
        shadows += local_shadow;
    });
    int32_t k = 0;
    for (int i = 0; i < n; ++i) {
        if (!alive[i]) continue;
        next->ox[k] = no[i * 3]; next->oy[k] = no[i * 3 + 1]; next->oz[k] = no[i * 3 + 2];
        next->dx[k] = nd[i * 3]; next->dy[k] = nd[i * 3 + 1]; next->dz[k] = nd[i * 3 + 2];
        next->tr[k] = ntt[i * 3]; next->tg[k] = ntt[i * 3 + 1]; next->tb[k] = ntt[i * 3 + 2];
        next->pixel[k] = q->pixel[i];
        next->depth[k] = q->depth[i] + 1;
        next->pdf[k] = npdf[i];
        ++k;
    }
    *n_next = k;
    if (shadow_rays) *shadow_rays = shadows.load();
}

// Miss loop (cpp:256-265) with the unbiased track's environment MIS.
void orc_shade_misses(void* sp, const wf_ray_soa* q, int32_t n, int32_t physics,
                      float* accum_rgb, double* accum64) {
    const Scene& s = *(Scene*)sp;
    for (int i = 0; i < n; ++i) {
        V3 d(q->dx[i], q->dy[i], q->dz[i]);
        V3 T(q->tr[i], q->tg[i], q->tb[i]);
        V3 c;
        if (physics == WF_PHYSICS_UNBIASED && s.env_w > 0) {
            float lc;
            int texel = env_texel(s, d, lc);
            float w = 1.0f;
            if (q->depth[i] > 0 && q->pdf[i] > 0.0f) w = mis_power(q->pdf[i], env_pdf_sa(s, texel, lc));
            c = T * env_radiance(s, texel) * w;
        } else {
            c = T * sky(s, d);
        }
        float* a = accum_rgb + (size_t)q->pixel[i] * 3;
        a[0] = a[0] + c.x;
        a[1] = a[1] + c.y;
        a[2] = a[2] + c.z;
        if (accum64) {
            double* b = accum64 + (size_t)q->pixel[i] * 3;
            b[0] += c.x;
            b[1] += c.y;
            b[2] += c.z;
        }
    }
}

// computeAverageRunLength (src/render/WavefrontRenderer.cpp:67-92): entries without a texture
// are skipped; a run starts where the ID differs from the previous kept entry's.
double orc_run_length(int32_t n, const int32_t* ids, const int32_t* texture_ids) {
    if (n == 0) return 0.0;
    int runs = 0, kept = 0, prev = -2;
    for (int i = 0; i < n; ++i) {
        if (texture_ids[i] < 0) continue;
        if (ids[i] != prev) ++runs;
        prev = ids[i];
        ++kept;
    }
    return kept == 0 ? 0.0 : static_cast<double>(kept) / runs;
}

// cacheLineHomogeneity (src/render/WavefrontRenderer.cpp:210-234): lines of 16 IDs holding one
// ID over all lines.
double orc_line_homogeneity(int32_t n, const int32_t* ids) {
    if (n == 0) return 0.0;
    int mono = 0, lines = 0;
    for (int start = 0; start < n; start += 16) {
        const int end = std::min(start + 16, n);
        bool same = true;
        for (int j = start + 1; j < end; ++j)
            if (ids[j] != ids[start]) {
                same = false;
                break;
            }
        mono += same ? 1 : 0;
        ++lines;
    }
    return static_cast<double>(mono) / lines;
}

// AdaptiveSampler::addSample (AdaptiveSampler.h:40-53).
void orc_adaptive_add(float* mean, float* m2, int32_t* count, int32_t i, float x) {
    count[i] += 1;
    int c = count[i];
    float before = mean[i];
    float dev = x - before;
    mean[i] += dev / static_cast<float>(c);
    float dev2 = x - mean[i];
    m2[i] += dev * dev2;
}

// AdaptiveSampler::updateConvergence (AdaptiveSampler.h:55-80).
void orc_adaptive_update(int32_t n, const float* mean, const float* m2, const int32_t* count,
                         uint8_t* active, float rel, int32_t min_samples) {
    for (int i = 0; i < n; ++i) {
        if (!active[i]) continue;
        int c = count[i];
        if (c < min_samples) continue;
        float var = m2[i] / static_cast<float>(c - 1);
        float se = std::sqrt(var / static_cast<float>(c));
        float thr = rel * (mean[i] + 0.01f);
        if (se < thr) active[i] = 0;
    }
}

// WavefrontRenderer::renderScene (cpp:94-399) on the CPU with the same queues and policies.
// image: float RGB out; counts: per-pixel samples out (may be NULL); image64: FP64-accumulated
// image out (may be NULL); stats: ray counts and phase wall times (milliseconds).
static int32_t render_masked(void* sp, const wf_render_params* rp, const uint8_t* pixel_mask,
                             const double* mat_cost, const double* tex_cost, float* image,
                             int32_t* counts, double* image64, wf_render_stats* stats,
                             int32_t threads);

int32_t orc_render(void* sp, const wf_render_params* rp, const double* mat_cost,
                   const double* tex_cost, float* image, int32_t* counts, double* image64,
                   wf_render_stats* stats, int32_t threads) {
    return render_masked(sp, rp, nullptr, mat_cost, tex_cost, image, counts, image64, stats,
                         threads);
}

// The frame of orc_render restricted to the pixels of a window [x0, x0 + w) x [y0, y0 + h) of the
// full image: every other pixel is never traced. A pixel's paths are those of the full frame
// (their draws are addressed by the full-frame pixel index). Outputs are full-frame sized.
int32_t orc_render_window(void* sp, const wf_render_params* rp, int32_t x0, int32_t y0,
                          int32_t w, int32_t h, float* image, int32_t* counts, double* image64,
                          wf_render_stats* stats, int32_t threads) {
    std::vector<uint8_t> mask((size_t)rp->width * rp->height, 0);
    for (int y = std::max(0, y0); y < std::min(rp->height, y0 + h); ++y)
        for (int x = std::max(0, x0); x < std::min(rp->width, x0 + w); ++x)
            mask[(size_t)y * rp->width + x] = 1;
    return render_masked(sp, rp, mask.data(), nullptr, nullptr, image, counts, image64, stats,
                         threads);
}

static int32_t render_masked(void* sp, const wf_render_params* rp, const uint8_t* pixel_mask,
                             const double* mat_cost, const double* tex_cost, float* image,
                             int32_t* counts, double* image64, wf_render_stats* stats,
                             int32_t threads) {
    const Scene& s = *(Scene*)sp;
    const int W = rp->width, H = rp->height, P = W * H;
    std::vector<float> accum((size_t)P * 3, 0.0f), before;
    std::vector<double> acc64(image64 ? (size_t)P * 3 : 0, 0.0);
    std::vector<float> mean(P, 0.0f), m2(P, 0.0f);
    std::vector<int32_t> cnt(P, 0);
    std::vector<uint8_t> active(P, 1);
    wf_render_stats st{};
    int samples_run = rp->spp;
    const int nmat = (int)s.materials.size(), ntex = (int)s.textures.size();
    std::vector<int32_t> mtype(nmat), mtex(nmat);
    for (int m = 0; m < nmat; ++m) {
        mtype[m] = s.materials[m].type;
        mtex[m] = s.materials[m].texture_id;
    }
    std::vector<uint64_t> tmem(ntex);
    for (int t = 0; t < ntex; ++t) tmem[t] = s.textures[t].memory_bytes();
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    auto t_frame = now();
    // coherence batch sums (cpp:154-160, 187-246)
    double coh[4] = {0.0, 0.0, 0.0, 0.0};
    long long coh_weight = 0;
    for (int sample = 0; sample < rp->spp; ++sample) {
        if (rp->adaptive) before = accum;
        std::vector<int32_t> pix;
        for (int p = 0; p < P; ++p)
            if ((!pixel_mask || pixel_mask[p]) && (!rp->adaptive || active[p])) pix.push_back(p);
        RayQ q;
        q.resize(pix.size());
        orc_raygen(&rp->camera, W, H, (int)pix.size(), pix.data(), rp->shade.seed, rp->shade.job,
                   rp->first_sample + sample, q.ox.data(), q.oy.data(), q.oz.data(), q.dx.data(), q.dy.data(),
                   q.dz.data());
        for (size_t i = 0; i < pix.size(); ++i) {
            q.tr[i] = q.tg[i] = q.tb[i] = 1.0f;
            q.pixel[i] = pix[i];
            q.depth[i] = 0;
            q.pdf[i] = -1.0f;
        }
        for (int depth = 0; depth < rp->max_depth; ++depth) {
            const int n = (int)q.size();
            if (n == 0) break;
            st.rays_traced += n;
            auto t0 = now();
            if (rp->ray_sort) {
                std::vector<int32_t> idx(n);
                float bmin[3] = {s.bounds_min.x, s.bounds_min.y, s.bounds_min.z};
                float bmax[3] = {s.bounds_max.x, s.bounds_max.y, s.bounds_max.z};
                orc_ray_schedule(n, q.ox.data(), q.oy.data(), q.oz.data(), q.dx.data(),
                                 q.dy.data(), q.dz.data(), bmin, bmax, idx.data());
                q.gather(idx);
            }
            auto t1 = now();
            st.sort_ms += ms(t0, t1);
            // intersectAll
            std::vector<float> hb[21];
            for (auto& v : hb) v.resize(n);
            std::vector<int32_t> hi[6];
            for (auto& v : hi) v.resize(n);
            std::vector<uint8_t> hfront(n);
            wf_shade_soa hs{hb[0].data(), hb[1].data(), hb[2].data(), hb[3].data(), hb[4].data(),
                            hb[5].data(), hb[6].data(), hb[7].data(), hb[8].data(), hi[0].data(),
                            hi[1].data(), hi[2].data(), hb[9].data(), hb[10].data(),
                            hb[11].data(), hb[12].data(), hb[13].data(), hb[14].data(),
                            hb[15].data(), hb[16].data(), hb[17].data(), hi[3].data(),
                            hi[4].data(), hb[18].data(), hi[5].data(), hfront.data()};
            RayQ mq;
            mq.resize(n);
            wf_ray_soa in{q.ox.data(), q.oy.data(), q.oz.data(), q.dx.data(), q.dy.data(),
                          q.dz.data(), q.tr.data(), q.tg.data(), q.tb.data(), q.pixel.data(),
                          q.depth.data(), q.pdf.data()};
            wf_ray_soa ms_{mq.ox.data(), mq.oy.data(), mq.oz.data(), mq.dx.data(), mq.dy.data(),
                           mq.dz.data(), mq.tr.data(), mq.tg.data(), mq.tb.data(),
                           mq.pixel.data(), mq.depth.data(), mq.pdf.data()};
            int32_t nh = 0, nm = 0;
            orc_intersect_split(sp, n, &in, &hs, &nh, &ms_, &nm, threads);
            auto t2 = now();
            st.intersect_ms += ms(t1, t2);
            st.shaded_hits += nh;
            st.misses += nm;
            // ShadingQueue::schedule + materialize
            std::vector<int32_t> order(nh);
            orc_shading_schedule(rp->policy, nh, hs.mat, hs.tex, hs.tr, hs.tg, hs.tb, hs.px,
                                 hs.py, hs.pz, nmat, mtype.data(), mtex.data(), ntex,
                                 tmem.data(), mat_cost, tex_cost, order.data());
            std::vector<float> sb[19];
            std::vector<int32_t> si[6];
            std::vector<uint8_t> sfront(nh);
            float* fsrc[19] = {hs.px, hs.py, hs.pz, hs.nx, hs.ny, hs.nz, hs.t, hs.u, hs.v, hs.ox,
                               hs.oy, hs.oz, hs.dx, hs.dy, hs.dz, hs.tr, hs.tg, hs.tb, hs.pdf};
            int32_t* isrc[6] = {hs.tri, hs.mat, hs.tex, hs.pixel, hs.depth, hs.geom};
            for (int c = 0; c < 19; ++c) {
                sb[c].resize(nh);
                for (int i = 0; i < nh; ++i) sb[c][i] = fsrc[c][order[i]];
            }
            for (int c = 0; c < 6; ++c) {
                si[c].resize(nh);
                for (int i = 0; i < nh; ++i) si[c][i] = isrc[c][order[i]];
            }
            for (int i = 0; i < nh; ++i) sfront[i] = hfront[order[i]];
            if (rp->coherence_stats && nh > 0) {
                coh[0] += orc_run_length(nh, si[1].data(), si[2].data()) * nh;
                coh[1] += orc_run_length(nh, si[2].data(), si[2].data()) * nh;
                if (rp->coherence_stats >= WF_COHERENCE_CACHE_LINES) {
                    coh[2] += orc_line_homogeneity(nh, si[1].data()) * nh;
                    coh[3] += orc_line_homogeneity(nh, si[2].data()) * nh;
                }
                coh_weight += nh;
            }
            wf_shade_soa ss{sb[0].data(), sb[1].data(), sb[2].data(), sb[3].data(), sb[4].data(),
                            sb[5].data(), sb[6].data(), sb[7].data(), sb[8].data(), si[0].data(),
                            si[1].data(), si[2].data(), sb[9].data(), sb[10].data(),
                            sb[11].data(), sb[12].data(), sb[13].data(), sb[14].data(),
                            sb[15].data(), sb[16].data(), sb[17].data(), si[3].data(),
                            si[4].data(), sb[18].data(), si[5].data(), sfront.data()};
            auto t3 = now();
            st.schedule_ms += ms(t2, t3);
            // shadeAll
            RayQ nq;
            nq.resize(nh);
            wf_ray_soa nx{nq.ox.data(), nq.oy.data(), nq.oz.data(), nq.dx.data(), nq.dy.data(),
                          nq.dz.data(), nq.tr.data(), nq.tg.data(), nq.tb.data(),
                          nq.pixel.data(), nq.depth.data(), nq.pdf.data()};
            int32_t nn = 0;
            int64_t shadows = 0;
            wf_shade_params sp2 = rp->shade;
            sp2.sample_index = rp->first_sample + sample;
            orc_shade(sp, &ss, nh, &sp2, mat_cost, accum.data(), image64 ? acc64.data() : nullptr,
                      &nx, &nn, &shadows, threads);
            nq.resize(nn);
            st.shadow_rays += shadows;
            auto t4 = now();
            st.shade_ms += ms(t3, t4);
            // miss loop
            mq.resize(nm);
            wf_ray_soa mq2{mq.ox.data(), mq.oy.data(), mq.oz.data(), mq.dx.data(), mq.dy.data(),
                           mq.dz.data(), mq.tr.data(), mq.tg.data(), mq.tb.data(),
                           mq.pixel.data(), mq.depth.data(), mq.pdf.data()};
            orc_shade_misses(sp, &mq2, nm, rp->shade.physics, accum.data(),
                             image64 ? acc64.data() : nullptr);
            st.miss_ms += ms(t4, now());
            q = std::move(nq);
        }
        if (rp->adaptive) {
            for (int p = 0; p < P; ++p) {
                if (!active[p]) continue;
                V3 c(accum[p * 3] - before[p * 3], accum[p * 3 + 1] - before[p * 3 + 1],
                     accum[p * 3 + 2] - before[p * 3 + 2]);
                orc_adaptive_add(mean.data(), m2.data(), cnt.data(), p, luminance(c));
            }
            if ((sample + 1) % rp->adaptive_interval == 0) {
                orc_adaptive_update(P, mean.data(), m2.data(), cnt.data(), active.data(),
                                    rp->adaptive_threshold, rp->adaptive_min_samples);
                int act = 0;
                for (int p = 0; p < P; ++p) act += active[p];
                if (act == 0) {
                    samples_run = sample + 1;
                    break;
                }
            }
        }
    }
    for (int p = 0; p < P; ++p) {
        int c = rp->adaptive ? cnt[p] : samples_run;
        c = std::max(c, 1);
        for (int k = 0; k < 3; ++k) {
            image[p * 3 + k] = accum[p * 3 + k] / static_cast<float>(c);
            if (image64) image64[p * 3 + k] = acc64[p * 3 + k] / (double)c;
        }
        if (counts) counts[p] = c;
    }
    int act = 0;
    for (int p = 0; p < P; ++p) act += active[p];
    st.samples_run = samples_run;
    st.active_pixels_end = rp->adaptive ? act : 0;
    st.frame_ms = ms(t_frame, now());
    if (coh_weight > 0) {
        st.material_run_length = coh[0] / coh_weight;
        st.texture_run_length = coh[1] / coh_weight;
        st.material_homogeneity = coh[2] / coh_weight;
        st.texture_homogeneity = coh[3] / coh_weight;
    }
    if (stats) *stats = st;
    return 0;
}

}  // extern "C"
