// ORACLE -- test infrastructure only. C entry points over the parts of the reference renderer
// that compile from the reference tree by themselves (no Embree, TBB, OIIO or USD needed):
//   src/shading/Material.cpp, src/geometry/Mesh.cpp, include/geometry/Triangle.h,
//   include/core/{Camera,HitRecord}.h, include/math/{SpatialHash,Random}.h,
//   include/shading/Texture.h, include/render/{AdaptiveSampler,CostTracker}.h.
// Built by `make ref` into oracle/_ref/ with the reference include directory on the path; it is
// used only to check the oracle's restatement. RayQueue.h and ShadingQueue.h include
// <tbb/parallel_sort.h>, whose headers are absent here, so they are not built.
// shading/Material.h comes first: it pulls in Texture.h before SpatialHash.h, the order the
// reference's own Material.cpp relies on (the two headers include each other).
#include <shading/Material.h>
#include <shading/Texture.h>
#include <core/Camera.h>
#include <core/HitRecord.h>
#include <geometry/Mesh.h>
#include <geometry/Triangle.h>
#include <math/Random.h>
#include <math/SpatialHash.h>
#include <render/AdaptiveSampler.h>
#include <render/CostTracker.h>

#include <cstdint>
#include <random>
#include <vector>

extern "C" {

int32_t ref_intersect_triangle(const float* o, const float* d, const float* a, const float* b,
                               const float* c, float tmin, float tmax, float* t, float* u,
                               float* v) {
    Ray ray(Point3(o[0], o[1], o[2]), Vec3(d[0], d[1], d[2]));
    return intersectTriangle(ray, Point3(a[0], a[1], a[2]), Point3(b[0], b[1], b[2]),
                             Point3(c[0], c[1], c[2]), tmin, tmax, *t, *u, *v)
               ? 1
               : 0;
}

// Mesh::interpolateUV (uv_u/uv_v may be NULL: no authored UVs).
void ref_interpolate_uv(const float* uv_u, const float* uv_v, float bu, float bv, float* u,
                        float* v) {
    Mesh mesh;
    mesh.vertexPositions = {Point3(0, 0, 0), Point3(1, 0, 0), Point3(0, 1, 0)};
    mesh.triangleIndices = {0, 1, 2};
    if (uv_u && uv_v) {
        mesh.triangleUVsU = {uv_u[0], uv_u[1], uv_u[2]};
        mesh.triangleUVsV = {uv_v[0], uv_v[1], uv_v[2]};
    }
    mesh.interpolateUV(0, bu, bv, *u, *v);
}

// Mesh::hit brute force (the reference's fallback when no accelerator is built, Scene.cpp:145-157).
int32_t ref_mesh_hit(const float* positions, int32_t n_vertices, const int32_t* indices,
                     int32_t n_tris, const float* o, const float* d, float tmin, float tmax,
                     float* t, int32_t* prim, float* u, float* v, float* point, float* normal,
                     int32_t* front) {
    Mesh mesh;
    for (int i = 0; i < n_vertices; ++i)
        mesh.vertexPositions.push_back(
            Point3(positions[i * 3], positions[i * 3 + 1], positions[i * 3 + 2]));
    mesh.triangleIndices.assign(indices, indices + n_tris * 3);
    mesh.materialID = 0;
    HitRecord rec;
    Ray ray(Point3(o[0], o[1], o[2]), Vec3(d[0], d[1], d[2]));
    if (!mesh.hit(ray, tmin, tmax, rec)) return 0;
    *t = rec.distance;
    *prim = rec.triangleIndex;
    *u = rec.u;
    *v = rec.v;
    point[0] = rec.point.x; point[1] = rec.point.y; point[2] = rec.point.z;
    normal[0] = rec.normal.x; normal[1] = rec.normal.y; normal[2] = rec.normal.z;
    *front = rec.frontFace ? 1 : 0;
    return 1;
}

uint32_t ref_hash_grid_cell(int32_t x, int32_t y, int32_t z) { return hashGridCell(x, y, z); }

void ref_hash_to_uv(uint32_t h, float* u, float* v) { hashToUV(h, *u, *v); }

void ref_texture_sample(int32_t w, int32_t h, const float* rgb, float u, float v, float* out) {
    Texture tex(w, h);
    for (int i = 0; i < w * h; ++i) tex.pixels[i] = Color(rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2]);
    Color c = tex.sample(u, v);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

uint64_t ref_texture_memory_bytes(int32_t w, int32_t h) { return Texture(w, h).memoryBytes(); }

// Camera(lookFrom, lookAt, up, vfov, w, h).getRay(u, v)
void ref_camera_ray(const float* eye, const float* target, const float* up, float vfov,
                    int32_t w, int32_t h, float u, float v, float* o, float* d) {
    Camera cam(Point3(eye[0], eye[1], eye[2]), Point3(target[0], target[1], target[2]),
               Vec3(up[0], up[1], up[2]), vfov, w, h);
    Ray r = cam.getRay(u, v);
    o[0] = r.origin.x; o[1] = r.origin.y; o[2] = r.origin.z;
    d[0] = r.direction.x; d[1] = r.direction.y; d[2] = r.direction.z;
}

static Material make_material(int32_t type, const float* albedo, float roughness, float emission,
                              int32_t texture_id, float spot_outer, float spot_falloff, float ior,
                              int32_t checker, float cell) {
    Material m;
    m.type = static_cast<MaterialType>(type);
    m.albedo = Color(albedo[0], albedo[1], albedo[2]);
    m.roughness = roughness;
    m.emission = emission;
    m.textureID = texture_id;
    m.spotOuterAngle = spot_outer;
    m.spotFalloffAngle = spot_falloff;
    m.indexOfRefraction = ior;
    m.useSpatialChecker = checker != 0;
    m.spatialCheckerCellSize = cell;
    return m;
}

static std::vector<Texture> make_textures(int32_t n, const int32_t* ws, const int32_t* hs,
                                          const float* rgb) {
    std::vector<Texture> out;
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        Texture t(ws[i], hs[i]);
        for (int k = 0; k < ws[i] * hs[i]; ++k, ++off)
            t.pixels[k] = Color(rgb[off * 3], rgb[off * 3 + 1], rgb[off * 3 + 2]);
        out.push_back(t);
    }
    return out;
}

// The first k uniforms randomFloat() returns after getThreadRNG() is seeded with `seed`.
void ref_draws(uint32_t seed, int32_t k, float* out) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> dist(0.0f, 1.0f);
    for (int i = 0; i < k; ++i) out[i] = dist(rng);
}

// Material::scatter with the thread RNG reseeded to `seed`; *consumed = draws the call used.
int32_t ref_scatter(uint32_t seed, int32_t type, const float* albedo, float roughness,
                    float emission, int32_t texture_id, float ior, int32_t checker, float cell,
                    int32_t n_tex, const int32_t* tex_w, const int32_t* tex_h,
                    const float* tex_rgb, const float* o, const float* d, const float* point,
                    const float* normal, int32_t front, float u, float v, float* att,
                    float* dir, float* pdf, int32_t* consumed) {
    Material m = make_material(type, albedo, roughness, emission, texture_id, 45.0f, 30.0f, ior,
                               checker, cell);
    std::vector<Texture> textures = make_textures(n_tex, tex_w, tex_h, tex_rgb);
    HitRecord rec;
    rec.point = Point3(point[0], point[1], point[2]);
    rec.normal = Vec3(normal[0], normal[1], normal[2]);
    rec.frontFace = front != 0;
    rec.u = u;
    rec.v = v;
    rec.distance = 1.0f;
    rec.materialID = 0;
    rec.textureID = texture_id;
    Ray in(Point3(o[0], o[1], o[2]), Vec3(d[0], d[1], d[2]));
    getThreadRNG().seed(seed);
    std::mt19937 probe(seed);
    Color a;
    Ray out;
    float p = 0.0f;
    bool ok = m.scatter(in, rec, textures, a, out, p);
    int k = 0;
    while (!(probe == getThreadRNG()) && k < 100000) {
        probe.discard(1);
        ++k;
    }
    att[0] = a.x; att[1] = a.y; att[2] = a.z;
    dir[0] = out.direction.x; dir[1] = out.direction.y; dir[2] = out.direction.z;
    *pdf = p;
    *consumed = k;
    return ok ? 1 : 0;
}

void ref_emitted(int32_t type, const float* albedo, float emission, float spot_outer,
                 float spot_falloff, const float* ln, const float* dfl, float* out) {
    Material m = make_material(type, albedo, 0.0f, emission, -1, spot_outer, spot_falloff, 1.5f,
                               0, 1.0f);
    Color c = m.emitted(Vec3(ln[0], ln[1], ln[2]), Vec3(dfl[0], dfl[1], dfl[2]));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

void ref_surface_color(int32_t type, const float* albedo, int32_t texture_id, int32_t checker,
                       float cell, int32_t n_tex, const int32_t* tex_w, const int32_t* tex_h,
                       const float* tex_rgb, const float* point, float u, float v, float* out) {
    Material m = make_material(type, albedo, 0.5f, 0.0f, texture_id, 45.0f, 30.0f, 1.5f, checker,
                               cell);
    std::vector<Texture> textures = make_textures(n_tex, tex_w, tex_h, tex_rgb);
    HitRecord rec;
    rec.point = Point3(point[0], point[1], point[2]);
    rec.u = u;
    rec.v = v;
    Color c = m.getSurfaceColor(rec, textures);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

// HitRecord::setFaceNormal
void ref_set_face_normal(const float* d, const float* outward, float* normal, int32_t* front) {
    HitRecord rec;
    Ray r(Point3(0, 0, 0), Vec3(d[0], d[1], d[2]));
    rec.setFaceNormal(r, Vec3(outward[0], outward[1], outward[2]));
    normal[0] = rec.normal.x; normal[1] = rec.normal.y; normal[2] = rec.normal.z;
    *front = rec.frontFace ? 1 : 0;
}

float ref_luminance(float r, float g, float b) { return luminance(Color(r, g, b)); }

// AdaptiveSampler: feed samples[pixel_of[i]] in order, calling updateConvergence after the
// sample indices listed in checks; report the final state.
void ref_adaptive_run(int32_t n_pixels, float threshold, int32_t n_samples,
                      const int32_t* pixel_of, const float* values, int32_t n_checks,
                      const int32_t* checks, float* mean, float* m2, int32_t* count,
                      uint8_t* active) {
    AdaptiveSampler s(n_pixels, threshold);
    int c = 0;
    for (int i = 0; i < n_samples; ++i) {
        s.addSample(pixel_of[i], values[i]);
        while (c < n_checks && checks[c] == i) {
            s.updateConvergence();
            ++c;
        }
    }
    for (int p = 0; p < n_pixels; ++p) {
        mean[p] = s.pixelMean[p];
        m2[p] = s.pixelM2[p];
        count[p] = s.pixelSampleCount[p];
        active[p] = s.isActive(p) ? 1 : 0;
    }
}

// CostTracker: record (ids[i], costs[i]) in order, then relativeCost for every item.
void ref_cost_tracker_run(int32_t n_items, int32_t n_records, const int32_t* ids,
                          const double* costs, double* rel) {
    CostTracker t(n_items);
    for (int i = 0; i < n_records; ++i) t.record(ids[i], costs[i]);
    for (int k = 0; k < n_items; ++k) rel[k] = t.relativeCost(k);
}

}  // extern "C"
