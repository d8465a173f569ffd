// ORACLE -- test infrastructure only (see orc_core.h). Restatement of the reference's geometry
// path: scene flattening (Scene::addMesh, src/core/Scene.cpp:36-56), bounds (Scene.cpp:59-77),
// Moller-Trumbore (include/geometry/Triangle.h:6-42), closest hit and occlusion with Embree's
// tnear/tfar conventions (src/accelerator/Embree.cpp:122-279), and the HitRecord mapping
// (Embree.cpp:240-252, src/geometry/Mesh.cpp:46-63, include/core/HitRecord.h:17-20).
//
// Embree itself cannot run here. Closest hit is defined as the lexicographic minimum of
// (t, global triangle index) over every triangle the Moller-Trumbore test accepts with
// t in [tnear, tfar]; the CPU BVH below returns exactly that minimum (checked against brute force).
#include <algorithm>
#include <cfloat>
#include <limits>
#include <memory>
#include <numeric>

#include "orc_core.h"

using namespace orc;

namespace {

// intersectTriangle, Triangle.h:6-42, with the edges precomputed (same float ops).
inline bool mt_intersect(const V3& o, const V3& d, const V3& v0, const V3& e1, const V3& e2,
                         float tmin, float tmax, float& t, float& bu, float& bv) {
    const float EPSILON = 0.0000001f;
    V3 pvec = d.cross(e2);
    float det = e1.dot(pvec);
    if (det > -EPSILON && det < EPSILON) return false;
    float inv = 1.0f / det;
    V3 tvec = o - v0;
    float u = inv * tvec.dot(pvec);
    if (u < 0.0f || u > 1.0f) return false;
    V3 qvec = tvec.cross(e1);
    float v = inv * d.dot(qvec);
    if (v < 0.0f || u + v > 1.0f) return false;
    float tt = inv * e2.dot(qvec);
    if (tt < tmin || tt > tmax) return false;
    t = tt;
    bu = u;
    bv = v;
    return true;
}

struct Hit {
    float t;
    int32_t g;
    float bu, bv;
};

inline bool better(float t, int32_t g, const Hit& h) {
    return h.g < 0 || t < h.t || (t == h.t && g < h.g);
}

void closest_brute(const Scene& s, const V3& o, const V3& d, float tnear, float tfar, Hit& h) {
    h.g = -1;
    h.t = tfar;
    const int nt = s.n_tris();
    for (int g = 0; g < nt; ++g) {
        float t, bu, bv;
        V3 e1 = s.v1[g] - s.v0[g], e2 = s.v2[g] - s.v0[g];
        if (mt_intersect(o, d, s.v0[g], e1, e2, tnear, tfar, t, bu, bv) && better(t, g, h)) {
            h = Hit{t, g, bu, bv};
        }
    }
}

bool any_brute(const Scene& s, const V3& o, const V3& d, float tnear, float tfar) {
    const int nt = s.n_tris();
    for (int g = 0; g < nt; ++g) {
        float t, bu, bv;
        V3 e1 = s.v1[g] - s.v0[g], e2 = s.v2[g] - s.v0[g];
        if (mt_intersect(o, d, s.v0[g], e1, e2, tnear, tfar, t, bu, bv)) return true;
    }
    return false;
}

// Conservative slab test: boxes are padded at build time, so a triangle the MT test accepts is
// never culled.
inline bool box_hit(const Scene::Node& nd, const V3& o, const V3& inv, float tmax, float& tenter) {
    float t0 = 0.0f, t1 = tmax;
    const float oo[3] = {o.x, o.y, o.z}, ii[3] = {inv.x, inv.y, inv.z};
    for (int a = 0; a < 3; ++a) {
        float tn = (nd.lo[a] - oo[a]) * ii[a];
        float tf = (nd.hi[a] - oo[a]) * ii[a];
        if (tn > tf) std::swap(tn, tf);
        t0 = tn > t0 ? tn : t0;
        t1 = tf < t1 ? tf : t1;
    }
    tenter = t0;
    return t0 <= t1 * 1.00000024f;
}

inline V3 safe_inv(const V3& d) {
    auto f = [](float v) {
        float a = std::fabs(v) < 1e-20f ? std::copysign(1e-20f, v) : v;
        return 1.0f / a;
    };
    return V3(f(d.x), f(d.y), f(d.z));
}

void closest_bvh(const Scene& s, const V3& o, const V3& d, float tnear, float tfar, Hit& h) {
    h.g = -1;
    h.t = tfar;
    if (s.nodes.empty()) return;
    V3 inv = safe_inv(d);
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Scene::Node& nd = s.nodes[stack[--sp]];
        float te;
        if (!box_hit(nd, o, inv, h.t, te)) continue;
        if (nd.count > 0) {
            int first = -nd.left - 1;
            for (int k = 0; k < nd.count; ++k) {
                int g = s.bvh_tris[first + k];
                float t, bu, bv;
                V3 e1 = s.v1[g] - s.v0[g], e2 = s.v2[g] - s.v0[g];
                if (mt_intersect(o, d, s.v0[g], e1, e2, tnear, tfar, t, bu, bv) &&
                    better(t, g, h))
                    h = Hit{t, g, bu, bv};
            }
        } else {
            stack[sp++] = nd.right;
            stack[sp++] = nd.left;
        }
    }
}

bool any_bvh(const Scene& s, const V3& o, const V3& d, float tnear, float tfar) {
    if (s.nodes.empty()) return false;
    V3 inv = safe_inv(d);
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Scene::Node& nd = s.nodes[stack[--sp]];
        float te;
        if (!box_hit(nd, o, inv, tfar, te)) continue;
        if (nd.count > 0) {
            int first = -nd.left - 1;
            for (int k = 0; k < nd.count; ++k) {
                int g = s.bvh_tris[first + k];
                float t, bu, bv;
                V3 e1 = s.v1[g] - s.v0[g], e2 = s.v2[g] - s.v0[g];
                if (mt_intersect(o, d, s.v0[g], e1, e2, tnear, tfar, t, bu, bv)) return true;
            }
        } else {
            stack[sp++] = nd.right;
            stack[sp++] = nd.left;
        }
    }
    return false;
}

}  // namespace

namespace orc {

// Binned-SAH BVH2 over triangle bounds, boxes padded by a relative epsilon.
void build_bvh(Scene& s) {
    const int nt = s.n_tris();
    s.nodes.clear();
    s.bvh_tris.resize(nt);
    std::iota(s.bvh_tris.begin(), s.bvh_tris.end(), 0);
    if (nt == 0) return;
    std::vector<float> blo(nt * 3), bhi(nt * 3), cen(nt * 3);
    for (int g = 0; g < nt; ++g) {
        const V3* p[3] = {&s.v0[g], &s.v1[g], &s.v2[g]};
        for (int a = 0; a < 3; ++a) {
            float lo = FLT_MAX, hi = -FLT_MAX;
            for (int k = 0; k < 3; ++k) {
                float c = a == 0 ? p[k]->x : (a == 1 ? p[k]->y : p[k]->z);
                lo = std::min(lo, c);
                hi = std::max(hi, c);
            }
            float pad = (std::fabs(lo) + std::fabs(hi) + (hi - lo)) * 4e-6f + 1e-30f;
            blo[g * 3 + a] = lo - pad;
            bhi[g * 3 + a] = hi + pad;
            cen[g * 3 + a] = 0.5f * (lo + hi);
        }
    }
    struct Task {
        int node, first, count;
    };
    s.nodes.reserve(2 * nt);
    s.nodes.push_back(Scene::Node{});
    std::vector<Task> tasks{{0, 0, nt}};
    while (!tasks.empty()) {
        Task tk = tasks.back();
        tasks.pop_back();
        Scene::Node nd;
        for (int a = 0; a < 3; ++a) {
            nd.lo[a] = FLT_MAX;
            nd.hi[a] = -FLT_MAX;
        }
        float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int k = tk.first; k < tk.first + tk.count; ++k) {
            int g = s.bvh_tris[k];
            for (int a = 0; a < 3; ++a) {
                nd.lo[a] = std::min(nd.lo[a], blo[g * 3 + a]);
                nd.hi[a] = std::max(nd.hi[a], bhi[g * 3 + a]);
                clo[a] = std::min(clo[a], cen[g * 3 + a]);
                chi[a] = std::max(chi[a], cen[g * 3 + a]);
            }
        }
        int best_axis = -1, best_split = 0;
        if (tk.count > 4) {
            const int B = 16;
            double best_cost = 1e300;
            for (int a = 0; a < 3; ++a) {
                float ext = chi[a] - clo[a];
                if (!(ext > 0.0f)) continue;
                int cnt[B] = {0};
                float lo[B][3], hi[B][3];
                for (int b = 0; b < B; ++b)
                    for (int c = 0; c < 3; ++c) {
                        lo[b][c] = FLT_MAX;
                        hi[b][c] = -FLT_MAX;
                    }
                for (int k = tk.first; k < tk.first + tk.count; ++k) {
                    int g = s.bvh_tris[k];
                    int b = std::min(B - 1, (int)((cen[g * 3 + a] - clo[a]) / ext * B));
                    cnt[b]++;
                    for (int c = 0; c < 3; ++c) {
                        lo[b][c] = std::min(lo[b][c], blo[g * 3 + c]);
                        hi[b][c] = std::max(hi[b][c], bhi[g * 3 + c]);
                    }
                }
                auto area = [](const float* l, const float* h) {
                    double dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
                    return dx * dy + dy * dz + dz * dx;
                };
                for (int sp = 1; sp < B; ++sp) {
                    float l0[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h0[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
                    float l1[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h1[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
                    int c0 = 0, c1 = 0;
                    for (int b = 0; b < B; ++b) {
                        if (!cnt[b]) continue;
                        float* L = b < sp ? l0 : l1;
                        float* H = b < sp ? h0 : h1;
                        (b < sp ? c0 : c1) += cnt[b];
                        for (int c = 0; c < 3; ++c) {
                            L[c] = std::min(L[c], lo[b][c]);
                            H[c] = std::max(H[c], hi[b][c]);
                        }
                    }
                    if (!c0 || !c1) continue;
                    double cost = c0 * area(l0, h0) + c1 * area(l1, h1);
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_split = sp;
                    }
                }
            }
        }
        int mid = -1;
        if (best_axis >= 0) {
            const int B = 16;
            float ext = chi[best_axis] - clo[best_axis];
            auto it = std::partition(
                s.bvh_tris.begin() + tk.first, s.bvh_tris.begin() + tk.first + tk.count,
                [&](int g) {
                    int b = std::min(B - 1, (int)((cen[g * 3 + best_axis] - clo[best_axis]) / ext * B));
                    return b < best_split;
                });
            mid = (int)(it - s.bvh_tris.begin());
            if (mid == tk.first || mid == tk.first + tk.count) mid = -1;
        }
        if (mid < 0 && tk.count > 8) mid = tk.first + tk.count / 2;  // degenerate centroids
        if (mid < 0) {
            nd.left = -(tk.first + 1);
            nd.right = 0;
            nd.count = tk.count;
            s.nodes[tk.node] = nd;
            continue;
        }
        int l = (int)s.nodes.size();
        s.nodes.push_back(Scene::Node{});
        int r = (int)s.nodes.size();
        s.nodes.push_back(Scene::Node{});
        nd.left = l;
        nd.right = r;
        nd.count = 0;
        s.nodes[tk.node] = nd;
        tasks.push_back({r, mid, tk.first + tk.count - mid});
        tasks.push_back({l, tk.first, mid - tk.first});
    }
}

}  // namespace orc

// Walker's alias method on a weight vector; prob[i] in [0,1], alias[i] the other outcome.
static void build_alias(const std::vector<double>& w, std::vector<float>& prob,
                        std::vector<int32_t>& alias) {
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
        size_t sidx = small.back();
        small.pop_back();
        size_t l = large.back();
        prob[sidx] = (float)q[sidx];
        alias[sidx] = (int32_t)l;
        q[l] = (q[l] + q[sidx]) - 1.0;
        if (q[l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (size_t i : large) prob[i] = 1.0f;
    for (size_t i : small) prob[i] = 1.0f;
}

namespace orc {

// Unbiased-track light tables: every emissive triangle weighted by area * luminance(radiance),
// environment texels weighted by luminance * cos(latitude) (equirectangular solid angle).
void build_light_tables(Scene& s) {
    s.light_tris.clear();
    std::vector<double> w;
    double total = 0.0;
    for (int32_t m : s.emissive_meshes) {
        const wf_material& mt = s.materials[s.mesh_material[m]];
        V3 rad = V3(mt.albedo[0], mt.albedo[1], mt.albedo[2]) * mt.emission;
        double lum = luminance(rad);
        for (int k = 0; k < s.mesh_tri_count[m]; ++k) {
            int g = s.mesh_tri_offset[m] + k;
            V3 e1 = s.v1[g] - s.v0[g], e2 = s.v2[g] - s.v0[g];
            double area = (double)(e1.cross(e2).length() * 0.5f);
            s.light_tris.push_back(g);
            w.push_back(area * lum);
            total += area * lum;
        }
    }
    build_alias(w, s.light_prob, s.light_alias);
    s.mat_light_pdf_area.assign(s.materials.size(), 0.0f);
    if (total > 0.0) {
        for (size_t m = 0; m < s.materials.size(); ++m) {
            const wf_material& mt = s.materials[m];
            if (mt.type != WF_MAT_EMISSIVE && mt.type != WF_MAT_SPOTLIGHT) continue;
            V3 rad = V3(mt.albedo[0], mt.albedo[1], mt.albedo[2]) * mt.emission;
            s.mat_light_pdf_area[m] = (float)((double)luminance(rad) / total);
        }
    }
    s.env_prob.clear();
    s.env_alias.clear();
    s.env_pdf.clear();
    s.env_choice_prob = 0.0f;
    if (s.env_w > 0) {
        const int W = s.env_w, H = s.env_h;
        std::vector<double> ew((size_t)W * H);
        double etot = 0.0;
        for (int y = 0; y < H; ++y) {
            double lat = (0.5 - (y + 0.5) / H) * M_PI;
            double c = std::cos(lat);
            for (int x = 0; x < W; ++x) {
                size_t i = (size_t)y * W + x;
                V3 raw(s.env_rgb[i * 3], s.env_rgb[i * 3 + 1], s.env_rgb[i * 3 + 2]);
                ew[i] = (double)luminance(raw) * c;
                etot += ew[i];
            }
        }
        build_alias(ew, s.env_prob, s.env_alias);
        s.env_pdf.resize(ew.size());
        for (size_t i = 0; i < ew.size(); ++i)
            s.env_pdf[i] = etot > 0.0 ? (float)(ew[i] / etot) : 0.0f;
        if (etot > 0.0) s.env_choice_prob = s.light_tris.empty() ? 1.0f : 0.5f;
    }
}

}  // namespace orc

extern "C" {

// Scene::addMesh (Scene.cpp:36-56) + buildAccelerator (Scene.cpp:79-113) for the oracle.
void* orc_scene_create(const wf_scene_desc* desc) {
    auto* s = new Scene();
    for (int m = 0; m < desc->mesh_count; ++m) {
        const wf_mesh_desc& md = desc->meshes[m];
        s->mesh_tri_offset.push_back(s->n_tris());
        s->mesh_tri_count.push_back(md.triangle_count);
        s->mesh_material.push_back(md.material_id);
        for (int k = 0; k < md.triangle_count; ++k) {
            const float* p = md.positions;
            int i0 = md.indices[k * 3], i1 = md.indices[k * 3 + 1], i2 = md.indices[k * 3 + 2];
            s->v0.push_back(V3(p[i0 * 3], p[i0 * 3 + 1], p[i0 * 3 + 2]));
            s->v1.push_back(V3(p[i1 * 3], p[i1 * 3 + 1], p[i1 * 3 + 2]));
            s->v2.push_back(V3(p[i2 * 3], p[i2 * 3 + 1], p[i2 * 3 + 2]));
            s->tri_geom.push_back(m);
            s->tri_prim.push_back(k);
            bool has = md.uv_u != nullptr && md.uv_v != nullptr;
            s->tri_has_uv.push_back(has ? 1 : 0);
            for (int c = 0; c < 3; ++c) {
                s->tri_uv.push_back(has ? md.uv_u[k * 3 + c] : 0.0f);
                s->tri_uv.push_back(has ? md.uv_v[k * 3 + c] : 0.0f);
            }
        }
    }
    s->materials.assign(desc->materials, desc->materials + desc->material_count);
    // Texture storage: linear float RGB (Texture::load decodes 8-bit sources, Texture.cpp:30-58).
    float lut[256];
    for (int i = 0; i < 256; ++i) {
        float c = (float)i / 255.0f;
        lut[i] = c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
    }
    for (int t = 0; t < desc->texture_count; ++t) {
        const wf_texture_desc& td = desc->textures[t];
        Texture tx;
        tx.width = td.width;
        tx.height = td.height;
        size_t n = (size_t)td.width * td.height;
        tx.texels.resize(n);
        if (td.format == WF_TEX_RGBA8_SRGB) {
            const uint8_t* b = (const uint8_t*)td.data;
            for (size_t i = 0; i < n; ++i)
                tx.texels[i] = V3(lut[b[i * 4]], lut[b[i * 4 + 1]], lut[b[i * 4 + 2]]);
        } else {
            const float* f = (const float*)td.data;
            for (size_t i = 0; i < n; ++i) tx.texels[i] = V3(f[i * 3], f[i * 3 + 1], f[i * 3 + 2]);
        }
        s->textures.push_back(std::move(tx));
    }
    s->directional = desc->directional;
    if (desc->environment.width > 0 && desc->environment.rgb) {
        s->env_w = desc->environment.width;
        s->env_h = desc->environment.height;
        s->env_rgb.assign(desc->environment.rgb,
                          desc->environment.rgb + (size_t)s->env_w * s->env_h * 3);
        s->env_intensity = desc->environment.intensity;
        s->env_tint = V3(desc->environment.tint[0], desc->environment.tint[1],
                         desc->environment.tint[2]);
    }
    // Emitter cache and total emissive area (Scene.cpp:84-93, 162-190): float sum in mesh order.
    float total = 0.0f;
    for (int m = 0; m < desc->mesh_count; ++m) {
        int ty = s->materials[s->mesh_material[m]].type;
        if (ty == WF_MAT_EMISSIVE || ty == WF_MAT_SPOTLIGHT) {
            s->emissive_meshes.push_back(m);
            for (int k = 0; k < s->mesh_tri_count[m]; ++k) {
                int g = s->mesh_tri_offset[m] + k;
                V3 e1 = s->v1[g] - s->v0[g], e2 = s->v2[g] - s->v0[g];
                total += e1.cross(e2).length() * 0.5f;
            }
        }
    }
    s->total_emissive_area = total;
    // computeBounds over all vertices (Scene.cpp:59-77): every vertex of every mesh.
    V3 lo(FLT_MAX, FLT_MAX, FLT_MAX), hi(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (int m = 0; m < desc->mesh_count; ++m) {
        const wf_mesh_desc& md = desc->meshes[m];
        for (int v = 0; v < md.vertex_count; ++v) {
            const float* p = md.positions + v * 3;
            lo = V3(std::min(lo.x, p[0]), std::min(lo.y, p[1]), std::min(lo.z, p[2]));
            hi = V3(std::max(hi.x, p[0]), std::max(hi.y, p[1]), std::max(hi.z, p[2]));
        }
    }
    s->bounds_min = lo;
    s->bounds_max = hi;
    build_bvh(*s);
    build_light_tables(*s);
    return s;
}

void orc_scene_destroy(void* s) { delete (Scene*)s; }

void orc_scene_info(void* sp, float* bmin, float* bmax, float* total_emissive_area,
                    int32_t* n_tris) {
    Scene* s = (Scene*)sp;
    bmin[0] = s->bounds_min.x;
    bmin[1] = s->bounds_min.y;
    bmin[2] = s->bounds_min.z;
    bmax[0] = s->bounds_max.x;
    bmax[1] = s->bounds_max.y;
    bmax[2] = s->bounds_max.z;
    *total_emissive_area = s->total_emissive_area;
    *n_tris = s->n_tris();
}

// intersectTriangle (Triangle.h:6-42) on explicit vertices (for the TriangleTest known answers).
int32_t orc_intersect_triangle(const float* o, const float* d, const float* a, const float* b,
                               const float* c, float tmin, float tmax, float* t, float* u,
                               float* v) {
    V3 v0(a[0], a[1], a[2]), v1(b[0], b[1], b[2]), v2(c[0], c[1], c[2]);
    return mt_intersect(V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), v0, v1 - v0, v2 - v0, tmin,
                        tmax, *t, *u, *v)
               ? 1
               : 0;
}

// Closest hit for n rays; use_bvh=0 brute force over every triangle, 1 the CPU BVH.
void orc_closest_hit(void* sp, int32_t n, const float* ox, const float* oy, const float* oz,
                     const float* dx, const float* dy, const float* dz, float tnear, float tfar,
                     int32_t use_bvh, float* t, int32_t* geom, int32_t* prim, float* bu,
                     float* bv) {
    const Scene& s = *(Scene*)sp;
    for (int32_t i = 0; i < n; ++i) {
        Hit h;
        V3 o(ox[i], oy[i], oz[i]), d(dx[i], dy[i], dz[i]);
        if (use_bvh)
            closest_bvh(s, o, d, tnear, tfar, h);
        else
            closest_brute(s, o, d, tnear, tfar, h);
        t[i] = h.g >= 0 ? h.t : tfar;
        geom[i] = h.g >= 0 ? s.tri_geom[h.g] : -1;
        prim[i] = h.g >= 0 ? s.tri_prim[h.g] : -1;
        bu[i] = h.g >= 0 ? h.bu : 0.0f;
        bv[i] = h.g >= 0 ? h.bv : 0.0f;
    }
}

// EmbreeAccelerator::occluded, Embree.cpp:258-279: tnear 0.001, tfar = maxDistance - 0.001;
// reports occluded when tfar ends up negative, which includes maxDistance < 0.001.
void orc_occluded(void* sp, int32_t n, const float* ox, const float* oy, const float* oz,
                  const float* dx, const float* dy, const float* dz, const float* maxd,
                  int32_t use_bvh, uint8_t* out) {
    const Scene& s = *(Scene*)sp;
    for (int32_t i = 0; i < n; ++i) {
        float tfar = maxd[i] - 0.001f;
        V3 o(ox[i], oy[i], oz[i]), d(dx[i], dy[i], dz[i]);
        bool occ = tfar < 0.0f ||
                   (use_bvh ? any_bvh(s, o, d, 0.001f, tfar) : any_brute(s, o, d, 0.001f, tfar));
        out[i] = occ ? 1 : 0;
    }
}

}  // extern "C"

namespace orc {

// HitRecord mapping of EmbreeAccelerator::intersect4, Embree.cpp:240-252: point = o + d*t,
// UVs through Mesh::interpolateUV (Mesh.cpp:46-63), geometric normal from the edges, flipped to
// face the ray by HitRecord::setFaceNormal (HitRecord.h:17-20).
void hit_record(const Scene& s, const V3& o, const V3& d, float t, int32_t g, float bu, float bv,
                V3& point, V3& normal, float& u, float& v, int32_t& mat, int32_t& tex,
                uint8_t& front) {
    point = o + d * t;
    int32_t m = s.tri_geom[g];
    mat = s.mesh_material[m];
    tex = s.materials[mat].texture_id;
    if (s.tri_has_uv[g]) {
        const float* uv = &s.tri_uv[(size_t)g * 6];
        float w0 = 1.0f - bu - bv;
        u = w0 * uv[0] + bu * uv[2] + bv * uv[4];
        v = w0 * uv[1] + bu * uv[3] + bv * uv[5];
    } else {
        u = bu;
        v = bv;
    }
    V3 e1 = s.v1[g] - s.v0[g], e2 = s.v2[g] - s.v0[g];
    V3 outward = e1.cross(e2).normalized();
    bool ff = d.dot(outward) < 0.0f;
    normal = ff ? outward : outward * -1.0f;
    front = ff ? 1 : 0;
}

bool trace_closest(const Scene& s, const V3& o, const V3& d, float tnear, float tfar, float& t,
                   int32_t& g, float& bu, float& bv) {
    Hit h;
    closest_bvh(s, o, d, tnear, tfar, h);
    t = h.t;
    g = h.g;
    bu = h.bu;
    bv = h.bv;
    return h.g >= 0;
}

bool trace_occluded(const Scene& s, const V3& o, const V3& d, float max_dist) {
    float tfar = max_dist - 0.001f;
    return tfar < 0.0f || any_bvh(s, o, d, 0.001f, tfar);
}

}  // namespace orc
