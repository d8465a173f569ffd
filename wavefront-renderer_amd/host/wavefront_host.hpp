// Header-only C++ adapter with the reference renderer's class surfaces over the C-ABI of
// libwavefront_hip.so. A renderer written against the reference's EmbreeAccelerator, RayQueue,
// ShadingQueue and WavefrontRenderer (include/accelerator/Embree.h:16-40,
// include/scheduling/RayQueue.h:63-245, include/scheduling/ShadingQueue.h:27-230,
// include/render/WavefrontRenderer.h:16-57) switches to the MI355X path by using these types.
//
//   wfhost::Device dev(0);
//   wfhost::SceneBuilder sb;                 // Scene::addMaterial / addTexture / addMesh
//   ...
//   wfhost::GpuAccelerator accel(dev, sb);   // EmbreeAccelerator::build
//   accel.intersect(ray, 0.001f, 1e9f, rec); // EmbreeAccelerator::intersect
//   wfhost::WavefrontRenderer r(16, 8, WF_POLICY_NONE);
//   r.renderScene(accel, camera, width, height, pixels);
//
// Errors raise wfhost::Error carrying the wf_status and wf_last_error() text (the reference
// ignores Embree errors; callers that want the reference's behaviour can catch and continue).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/wavefront_hip.h"

namespace wfhost {

struct Error : std::runtime_error {
    wf_status status;
    Error(wf_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(wf_status s, const char* call) {
    if (s != WF_OK) throw Error(s, std::string(call) + ": " + wf_last_error());
}

// ---- plain host types with the reference's field names ----------------------------------------
struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
using Point3 = Vec3;
using Color = Vec3;

struct Ray {                          // include/core/Ray.h:7-22
    Point3 origin;
    Vec3 direction;
};

struct HitRecord {                    // include/core/HitRecord.h:6-21 (+ geomID)
    Point3 point;
    Vec3 normal;
    float distance = 0.0f;
    int materialID = -1;
    int textureID = -1;
    int triangleIndex = -1;
    int geomID = -1;
    float u = 0.0f, v = 0.0f;
    bool frontFace = true;
};

// ---- device memory ---------------------------------------------------------------------------
class Device {
  public:
    explicit Device(int device = 0) { check(wf_context_create(device, &ctx_), "wf_context_create"); }
    ~Device() { wf_context_destroy(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    wf_context* ctx() const { return ctx_; }

  private:
    wf_context* ctx_ = nullptr;
};

template <typename T>
class DeviceArray {
  public:
    DeviceArray(const Device& d, size_t n) : ctx_(d.ctx()), n_(n) {
        void* p = nullptr;
        check(wf_device_alloc(ctx_, n * sizeof(T), &p), "wf_device_alloc");
        ptr_ = (T*)p;
    }
    ~DeviceArray() { wf_device_free(ctx_, ptr_); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    T* data() const { return ptr_; }
    size_t size() const { return n_; }
    void upload(const std::vector<T>& h) {
        check(wf_memcpy_htod(ctx_, ptr_, h.data(), h.size() * sizeof(T)), "wf_memcpy_htod");
    }
    std::vector<T> download(size_t n) const {
        std::vector<T> h(n);
        check(wf_memcpy_dtoh(ctx_, h.data(), ptr_, n * sizeof(T)), "wf_memcpy_dtoh");
        return h;
    }

  private:
    wf_context* ctx_;
    size_t n_;
    T* ptr_ = nullptr;
};

inline std::vector<uint8_t> DeviceArrayView(const Device& d, const uint8_t* p, size_t n) {
    std::vector<uint8_t> h(n);
    if (n) check(wf_memcpy_dtoh(d.ctx(), h.data(), p, n), "wf_memcpy_dtoh");
    return h;
}

// ---- scene description (Scene::addMaterial / addTexture / addMesh) ------------------------------
struct MeshData {                     // include/geometry/Mesh.h:8-20
    std::vector<float> positions;     // xyz per vertex
    std::vector<int32_t> triangleIndices;
    std::vector<float> triangleUVsU, triangleUVsV;
    int materialID = 0;
};

class SceneBuilder {
  public:
    int addMaterial(const wf_material& m) {
        materials_.push_back(m);
        return (int)materials_.size() - 1;
    }
    int addTexture(int w, int h, int format, std::vector<uint8_t> bytes) {
        texBytes_.push_back(std::move(bytes));
        wf_texture_desc t{w, h, format, nullptr};
        textures_.push_back(t);
        return (int)textures_.size() - 1;
    }
    void addMesh(MeshData m) { meshes_.push_back(std::move(m)); }
    void setDirectionalLight(const wf_directional_light& l) { directional_ = l; }

    // The wf_scene_desc view; valid while this builder lives.
    wf_scene_desc desc() {
        meshDesc_.clear();
        for (auto& m : meshes_) {
            wf_mesh_desc d{};
            d.positions = m.positions.data();
            d.vertex_count = (int32_t)(m.positions.size() / 3);
            d.indices = m.triangleIndices.data();
            d.triangle_count = (int32_t)(m.triangleIndices.size() / 3);
            d.uv_u = m.triangleUVsU.empty() ? nullptr : m.triangleUVsU.data();
            d.uv_v = m.triangleUVsV.empty() ? nullptr : m.triangleUVsV.data();
            d.material_id = m.materialID;
            meshDesc_.push_back(d);
        }
        for (size_t i = 0; i < textures_.size(); ++i) textures_[i].data = texBytes_[i].data();
        wf_scene_desc s{};
        s.meshes = meshDesc_.data();
        s.mesh_count = (int32_t)meshDesc_.size();
        s.materials = materials_.data();
        s.material_count = (int32_t)materials_.size();
        s.textures = textures_.data();
        s.texture_count = (int32_t)textures_.size();
        s.directional = directional_;
        return s;
    }

  private:
    std::vector<MeshData> meshes_;
    std::vector<wf_mesh_desc> meshDesc_;
    std::vector<wf_material> materials_;
    std::vector<wf_texture_desc> textures_;
    std::vector<std::vector<uint8_t>> texBytes_;
    wf_directional_light directional_{};
};

// ---- EmbreeAccelerator (include/accelerator/Embree.h:16-40) -------------------------------------
class GpuAccelerator {
  public:
    GpuAccelerator(const Device& dev, SceneBuilder& sb) : dev_(dev) {
        wf_scene_desc d = sb.desc();
        check(wf_scene_create(dev.ctx(), &d, &scene_), "wf_scene_create");
        check(wf_accel_build(scene_, &stats_), "wf_accel_build");
    }
    ~GpuAccelerator() { wf_scene_destroy(scene_); }
    GpuAccelerator(const GpuAccelerator&) = delete;
    GpuAccelerator& operator=(const GpuAccelerator&) = delete;

    const wf_build_stats& stats() const { return stats_; }
    wf_scene* handle() const { return scene_; }
    const Device& device() const { return dev_; }

    // EmbreeAccelerator::intersect (Embree.cpp:122-177)
    bool intersect(const Ray& ray, float tmin, float tmax, HitRecord& rec) const {
        std::vector<HitRecord> out(1);
        std::vector<int> hit = intersectBatch(std::vector<Ray>{ray}, tmin, tmax, out);
        if (hit[0]) rec = out[0];
        return hit[0] != 0;
    }
    // EmbreeAccelerator::intersect4 over any number of rays: returns a hit flag per ray and
    // fills the HitRecords the way Embree.cpp:240-252 does.
    std::vector<int> intersectBatch(const std::vector<Ray>& rays, float tmin, float tmax,
                                    std::vector<HitRecord>& records) const;
    // EmbreeAccelerator::occluded (Embree.cpp:258-279)
    bool occluded(const Point3& o, const Vec3& d, float maxDist) const {
        float v[7] = {o.x, o.y, o.z, d.x, d.y, d.z, maxDist};
        DeviceArray<float> buf(dev_, 7);
        buf.upload(std::vector<float>(v, v + 7));
        DeviceArray<uint8_t> out(dev_, 1);
        float* p = buf.data();
        check(wf_occluded(scene_, p, p + 1, p + 2, p + 3, p + 4, p + 5, p + 6, 1, out.data()),
              "wf_occluded");
        return out.download(1)[0] != 0;
    }

  private:
    const Device& dev_;
    wf_scene* scene_ = nullptr;
    wf_build_stats stats_{};
};

// SoA storage helper: one device array per field, the field order of the C structs.
template <typename SoA, int NF, int NI, int NB>
struct SoAStore {
    std::vector<std::unique_ptr<DeviceArray<float>>> f;
    std::vector<std::unique_ptr<DeviceArray<int32_t>>> i;
    std::vector<std::unique_ptr<DeviceArray<uint8_t>>> b;
    SoAStore(const Device& d, size_t cap) {
        for (int k = 0; k < NF; ++k) f.emplace_back(new DeviceArray<float>(d, cap ? cap : 1));
        for (int k = 0; k < NI; ++k) i.emplace_back(new DeviceArray<int32_t>(d, cap ? cap : 1));
        for (int k = 0; k < NB; ++k) b.emplace_back(new DeviceArray<uint8_t>(d, cap ? cap : 1));
    }
};

// Device RayQueue with the reference's schedule()/materialize() pair (RayQueue.h:165-215).
class DeviceRayQueue {
  public:
    DeviceRayQueue(const Device& dev, size_t capacity)
        : dev_(dev), cap_(capacity), a_(dev, capacity), b_(dev, capacity),
          sorted_(dev, capacity ? capacity : 1) {}

    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }

    // RayQueue::add for a whole batch (host -> device), field order of RayQueue.h:66-90.
    void assign(const std::vector<Ray>& rays, const std::vector<Color>& throughput,
                const std::vector<int>& pixel, const std::vector<int>& depth,
                const std::vector<float>& pdf) {
        n_ = rays.size();
        if (n_ > cap_) throw Error(WF_ERR_CAPACITY, "DeviceRayQueue: capacity exceeded");
        std::vector<float> cols[9];
        for (auto& c : cols) c.resize(n_);
        for (size_t k = 0; k < n_; ++k) {
            cols[0][k] = rays[k].origin.x; cols[1][k] = rays[k].origin.y; cols[2][k] = rays[k].origin.z;
            cols[3][k] = rays[k].direction.x; cols[4][k] = rays[k].direction.y; cols[5][k] = rays[k].direction.z;
            cols[6][k] = throughput[k].x; cols[7][k] = throughput[k].y; cols[8][k] = throughput[k].z;
        }
        for (int c = 0; c < 9; ++c) a_.f[c]->upload(cols[c]);
        a_.i[0]->upload(std::vector<int32_t>(pixel.begin(), pixel.end()));
        a_.i[1]->upload(std::vector<int32_t>(depth.begin(), depth.end()));
        a_.f[9]->upload(pdf);
    }

    // RayQueue::schedule: sortedIndices by octant + Morton key (stable on ties).
    void schedule(const Point3& sceneMin, const Point3& sceneMax) {
        float lo[3] = {sceneMin.x, sceneMin.y, sceneMin.z};
        float hi[3] = {sceneMax.x, sceneMax.y, sceneMax.z};
        wf_ray_soa s = soa(a_);
        check(wf_ray_schedule(dev_.ctx(), &s, (int32_t)n_, lo, hi, sorted_.data(), nullptr),
              "wf_ray_schedule");
    }
    // RayQueue::materialize: physical reorder; sortedIndices becomes the identity.
    void materialize() {
        wf_ray_soa s = soa(a_), d = soa(b_);
        check(wf_materialize_rays(dev_.ctx(), &s, sorted_.data(), (int32_t)n_, &d),
              "wf_materialize_rays");
        std::swap(a_, b_);
    }
    std::vector<int32_t> sortedIndices() const { return sorted_.download(n_); }
    std::vector<int32_t> pixelIndices() const { return a_.i[0]->download(n_); }
    std::vector<float> originsX() const { return a_.f[0]->download(n_); }
    wf_ray_soa view() const { return soa(a_); }
    void setSize(size_t n) { n_ = n; }

  private:
    using Store = SoAStore<wf_ray_soa, 10, 2, 0>;
    static wf_ray_soa soa(const Store& s) {
        return wf_ray_soa{s.f[0]->data(), s.f[1]->data(), s.f[2]->data(), s.f[3]->data(),
                          s.f[4]->data(), s.f[5]->data(), s.f[6]->data(), s.f[7]->data(),
                          s.f[8]->data(), s.i[0]->data(), s.i[1]->data(), s.f[9]->data()};
    }
    const Device& dev_;
    size_t cap_, n_ = 0;
    Store a_, b_;
    DeviceArray<int32_t> sorted_;
};

// Device ShadingQueue (ShadingQueue.h:27-275): filled by the hit/miss split of intersectAll,
// ordered by schedule(policy), physically reordered by materialize().
class DeviceShadingQueue {
  public:
    DeviceShadingQueue(const Device& dev, size_t capacity, int policy)
        : dev_(dev), cap_(capacity), policy_(policy), a_(dev, capacity), b_(dev, capacity),
          sorted_(dev, capacity ? capacity : 1) {}
    size_t size() const { return n_; }
    int policy() const { return policy_; }

    // WavefrontRenderer::intersectAll (cpp:459-567): this queue receives the hits, `misses` the
    // rest, both in input order.
    void intersectAll(const GpuAccelerator& accel, const DeviceRayQueue& rays,
                      DeviceRayQueue& misses) {
        wf_ray_soa in = rays.view(), out_m = misses.view();
        wf_shade_soa out_h = soa(a_);
        int32_t nh = 0, nm = 0;
        check(wf_intersect_split(accel.handle(), &in, (int32_t)rays.size(), &out_h, &nh, &out_m, &nm),
              "wf_intersect_split");
        n_ = (size_t)nh;
        misses.setSize((size_t)nm);
    }
    // ShadingQueue::schedule (ShadingQueue.h:90-141): cost tables live on the scene
    // (wf_scene_set_costs, the CostTracker::relativeCost values).
    void schedule(const GpuAccelerator& accel) {
        wf_shade_soa s = soa(a_);
        check(wf_shading_schedule(accel.handle(), &s, (int32_t)n_, policy_, sorted_.data()),
              "wf_shading_schedule");
    }
    void materialize() {
        wf_shade_soa s = soa(a_), d = soa(b_);
        check(wf_materialize_shading(dev_.ctx(), &s, sorted_.data(), (int32_t)n_, &d),
              "wf_materialize_shading");
        std::swap(a_, b_);
    }
    std::vector<int32_t> sortedIndices() const { return sorted_.download(n_); }
    std::vector<int32_t> materialIDs() const { return a_.i[1]->download(n_); }
    wf_shade_soa view() const { return soa(a_); }

  private:
    using Store = SoAStore<wf_shade_soa, 19, 6, 1>;
    static wf_shade_soa soa(const Store& s) {
        wf_shade_soa q;
        float** fl[19] = {&q.px, &q.py, &q.pz, &q.nx, &q.ny, &q.nz, &q.t, &q.u, &q.v, &q.ox,
                          &q.oy, &q.oz, &q.dx, &q.dy, &q.dz, &q.tr, &q.tg, &q.tb, &q.pdf};
        for (int k = 0; k < 19; ++k) *fl[k] = s.f[k]->data();
        int32_t** in[6] = {&q.tri, &q.mat, &q.tex, &q.pixel, &q.depth, &q.geom};
        for (int k = 0; k < 6; ++k) *in[k] = s.i[k]->data();
        q.front = s.b[0]->data();
        return q;
    }
    const Device& dev_;
    size_t cap_, n_ = 0;
    int policy_;
    Store a_, b_;
    DeviceArray<int32_t> sorted_;
};

inline std::vector<int> GpuAccelerator::intersectBatch(const std::vector<Ray>& rays, float tmin,
                                                       float tmax,
                                                       std::vector<HitRecord>& records) const {
    const size_t n = rays.size();
    std::vector<int> flags(n, 0);
    records.assign(n, HitRecord{});
    if (n == 0) return flags;
    DeviceRayQueue q(dev_, n);
    q.assign(rays, std::vector<Color>(n, Color{1, 1, 1}), std::vector<int>(n, 0),
             std::vector<int>(n, 0), std::vector<float>(n, -1.0f));
    DeviceShadingQueue rec(dev_, n, WF_POLICY_NONE);
    DeviceArray<uint8_t> hit(dev_, n);
    wf_ray_soa in = q.view();
    wf_shade_soa out = rec.view();
    check(wf_intersect_records(handle(), &in, (int32_t)n, tmin, tmax, &out, hit.data()),
          "wf_intersect_records");
    auto dl = [&](float* p) {
        std::vector<float> v(n);
        check(wf_memcpy_dtoh(dev_.ctx(), v.data(), p, sizeof(float) * n), "wf_memcpy_dtoh");
        return v;
    };
    auto dli = [&](int32_t* p) {
        std::vector<int32_t> v(n);
        check(wf_memcpy_dtoh(dev_.ctx(), v.data(), p, sizeof(int32_t) * n), "wf_memcpy_dtoh");
        return v;
    };
    auto px = dl(out.px), py = dl(out.py), pz = dl(out.pz);
    auto nx = dl(out.nx), ny = dl(out.ny), nz = dl(out.nz);
    auto tt = dl(out.t), uu = dl(out.u), vv = dl(out.v);
    auto tri = dli(out.tri), mat = dli(out.mat), tex = dli(out.tex), geom = dli(out.geom);
    std::vector<uint8_t> front = DeviceArrayView(dev_, out.front, n);
    std::vector<uint8_t> h = hit.download(n);
    for (size_t k = 0; k < n; ++k) {
        if (!h[k]) continue;
        HitRecord& r = records[k];
        r.point = {px[k], py[k], pz[k]};
        r.normal = {nx[k], ny[k], nz[k]};
        r.distance = tt[k];
        r.u = uu[k];
        r.v = vv[k];
        r.triangleIndex = tri[k];
        r.materialID = mat[k];
        r.textureID = tex[k];
        r.geomID = geom[k];
        r.frontFace = front[k] != 0;
        flags[k] = 1;
    }
    return flags;
}

// WavefrontRenderer (include/render/WavefrontRenderer.h:16-57): same public members; the
// per-bounce loop runs on the device.
class WavefrontRenderer {
  public:
    int samplesPerPixel;
    int maxDepth;
    int rrMinDepth = 3;
    int policy;
    bool enableCostAwareRR = false;
    bool enableRaySort = true;
    bool enableAdaptiveSampling = true;
    float adaptiveSamplingThreshold = 0.05f;
    float fireflyThreshold = 10.0f;
    int physics = WF_PHYSICS_REFERENCE;
    uint32_t seed = 0;

    WavefrontRenderer(int samples, int maxDepth_, int policy_)
        : samplesPerPixel(samples), maxDepth(maxDepth_), policy(policy_) {}

    // renderScene (cpp:94-399): returns device shade time in ms like the reference returns
    // totalShadeOnlyMs; pixels receives width*height RGB floats.
    double renderScene(const GpuAccelerator& accel, const wf_camera& camera, int width,
                       int height, std::vector<float>& pixels, wf_render_stats* statsOut = nullptr) {
        wf_render_params p{};
        p.width = width;
        p.height = height;
        p.spp = samplesPerPixel;
        p.max_depth = maxDepth;
        p.policy = policy;
        p.ray_sort = enableRaySort ? 1 : 0;
        p.adaptive = enableAdaptiveSampling ? 1 : 0;
        p.adaptive_threshold = adaptiveSamplingThreshold;
        p.adaptive_min_samples = 32;
        p.adaptive_interval = 16;
        p.shade = wf_shade_params{physics, rrMinDepth, enableCostAwareRR ? 1 : 0,
                                  fireflyThreshold, seed, 0, 0};
        p.camera = camera;
        DeviceArray<float> img(accel.device(), (size_t)width * height * 3);
        wf_render_stats st{};
        check(wf_render(accel.handle(), &p, img.data(), nullptr, &st), "wf_render");
        pixels = img.download((size_t)width * height * 3);
        if (statsOut) *statsOut = st;
        return st.shade_ms;
    }
};

}  // namespace wfhost
