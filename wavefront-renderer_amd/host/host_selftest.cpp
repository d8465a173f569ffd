// The reference's own unit tests, re-expressed against the C++ adapter (wavefront_host.hpp) so
// that the drop-in surfaces are exercised by a C++ caller exactly as the reference renderer uses
// them. Each check names the reference test it mirrors. Prints one PASS/FAIL line per check and
// exits non-zero on any failure. Built with plain g++ against libwavefront_hip.so.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "wavefront_host.hpp"

using namespace wfhost;

static int failures = 0;
static void expect(bool ok, const char* name) {
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
    std::fflush(stdout);
    if (!ok) ++failures;
}

static wf_material diffuse(float r, float g, float b) {
    wf_material m{};
    m.type = WF_MAT_DIFFUSE;
    m.albedo[0] = r;
    m.albedo[1] = g;
    m.albedo[2] = b;
    m.texture_id = -1;
    m.ior = 1.5f;
    m.checker_cell_size = 1.0f;
    return m;
}

static MeshData quad(std::vector<float> p, int material) {
    MeshData m;
    m.positions = std::move(p);
    m.triangleIndices = {0, 1, 2, 0, 2, 3};
    m.materialID = material;
    return m;
}

static Ray ray(float ox, float oy, float oz, float dx, float dy, float dz) {
    return Ray{{ox, oy, oz}, {dx, dy, dz}};
}

int main() {
    Device dev(0);

    // tests/accelerator/EmbreeTest.cpp:8-26 buildSingleQuadScene
    SceneBuilder sb;
    sb.addMaterial(diffuse(0.8f, 0.8f, 0.8f));
    sb.addMesh(quad({-1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1}, 0));
    GpuAccelerator accel(dev, sb);
    {
        HitRecord r;
        expect(accel.intersect(ray(0, 0, 0, 0, 0, -1), 0.001f, 1e9f, r), "EmbreeTest.RayHitsQuadReturnsTrue");
        expect(std::fabs(r.distance - 1.0f) < 1e-3f, "EmbreeTest.RayHitDistanceIsCorrect");
        expect(r.materialID == 0, "EmbreeTest.HitRecordMaterialIDMatchesMesh");
        expect(r.normal.z > 0.0f && r.frontFace, "EmbreeTest.HitRecordNormalPointsTowardRay");
        HitRecord r2;
        expect(!accel.intersect(ray(0, 0, 0, 0, 0, 1), 0.001f, 1e9f, r2), "EmbreeTest.RayMissReturnsNoHit");
        expect(!accel.intersect(ray(5, 5, 0, 0, 0, -1), 0.001f, 1e9f, r2), "EmbreeTest.RayMissesOutsideQuadBoundary");
        expect(!accel.intersect(ray(0, 0, 0, 0, 0, -1), 0.001f, 0.5f, r2), "EmbreeTest.HitBeyondMaxDistanceReturnsFalse");
        expect(accel.occluded({0, 0, 0}, {0, 0, -1}, 5.0f), "occluded: quad between origin and target");
        expect(!accel.occluded({0, 0, 0}, {0, 0, -1}, 0.5f), "occluded: target before the quad");
        expect(accel.occluded({0, 0, 0}, {0, 0, 1}, 0.0005f), "occluded: negative tfar reports occluded");
    }
    {   // EmbreeTest.cpp:113-124
        SceneBuilder empty;
        empty.addMaterial(diffuse(1, 1, 1));
        GpuAccelerator e(dev, empty);
        HitRecord r;
        expect(!e.intersect(ray(0, 0, 0, 0, 0, -1), 0.001f, 1e9f, r), "EmbreeTest.EmptySceneReturnsNoHit");
    }
    {   // EmbreeTest.cpp:126-156
        SceneBuilder two;
        int a = two.addMaterial(diffuse(1, 0, 0));
        int b = two.addMaterial(diffuse(0, 1, 0));
        MeshData ma, mb;
        ma.positions = {-1, -1, -2, 1, -1, -2, 0, 1, -2};
        ma.triangleIndices = {0, 1, 2};
        ma.materialID = a;
        mb.positions = {-1, -1, -1, 1, -1, -1, 0, 1, -1};
        mb.triangleIndices = {0, 1, 2};
        mb.materialID = b;
        two.addMesh(ma);
        two.addMesh(mb);
        GpuAccelerator t(dev, two);
        HitRecord r;
        t.intersect(ray(0, 0, 0, 0, 0, -1), 0.001f, 1e9f, r);
        expect(r.materialID == b && std::fabs(r.distance - 1.0f) < 1e-2f && r.geomID == 1,
               "EmbreeTest.MultiMeshSceneHitsCorrectMaterial");
    }
    {   // EmbreeTest.cpp:164-292 floor scene, intersect4 packets
        SceneBuilder fl;
        fl.addMaterial(diffuse(0.8f, 0.8f, 0.8f));
        fl.addMesh(quad({-5, 0, -5, 5, 0, -5, 5, 0, 5, -5, 0, 5}, 0));
        GpuAccelerator f(dev, fl);
        std::vector<HitRecord> rec;
        auto h = f.intersectBatch({ray(-2, 5, -2, 0, -1, 0), ray(0, 5, 0, 0, -1, 0),
                                   ray(2, 5, 2, 0, -1, 0), ray(-1, 5, 1, 0, -1, 0)},
                                  0.001f, 1e9f, rec);
        expect(h == std::vector<int>({1, 1, 1, 1}), "EmbreeTest.Intersect4AllFourRaysHit");
        bool mats = true;
        for (auto& x : rec) mats &= x.materialID == 0 && std::fabs(x.distance - 5.0f) < 1e-3f;
        expect(mats, "EmbreeTest.Intersect4MaterialIDCorrect");
        h = f.intersectBatch({ray(0, 5, 0, 0, -1, 0), ray(0, 5, 0, 0, 1, 0), ray(2, 5, 2, 0, -1, 0),
                              ray(0, 5, 0, 0, 1, 0)}, 0.001f, 1e9f, rec);
        expect(h == std::vector<int>({1, 0, 1, 0}), "EmbreeTest.Intersect4PartialHitMask");
    }
    {   // tests/scheduling/RayQueueTest.cpp:156-171 and 173-203
        DeviceRayQueue q(dev, 3);
        q.assign({ray(5, 0, 0, 0, 0, 1), ray(-5, 0, 0, 0, 0, 1), ray(0, 0, 0, 0, 0, 1)},
                 {{1, 1, 1}, {1, 1, 1}, {1, 1, 1}}, {100, 200, 300}, {0, 0, 0}, {-1, -1, -1});
        q.schedule({-10, -10, -10}, {10, 10, 10});
        q.materialize();
        expect(q.pixelIndices() == std::vector<int32_t>({200, 300, 100}),
               "RayQueueTest.ScheduleOrdersRaysBySpatialProximityWithinSameOctant");
        auto ox = q.originsX();
        expect(ox[0] == -5.0f && ox[1] == 0.0f && ox[2] == 5.0f,
               "RayQueueTest.MaterializeKeepsOriginDirectionAndPixelIndexPaired");
    }
    {   // tests/scheduling/ShadingQueueTest.cpp:92-130 via a real hit queue
        SceneBuilder ms;
        for (int t = 0; t < 6; ++t) {
            wf_material m = diffuse(1, 1, 1);
            m.type = t;
            ms.addMaterial(m);
        }
        // three stacked quads carrying materials 2, 0, 1 hit by three rays in that order
        ms.addMesh(quad({-1, 0, -1, 1, 0, -1, 1, 0, 1, -1, 0, 1}, 2));
        ms.addMesh(quad({9, 0, -1, 11, 0, -1, 11, 0, 1, 9, 0, 1}, 0));
        ms.addMesh(quad({19, 0, -1, 21, 0, -1, 21, 0, 1, 19, 0, 1}, 1));
        GpuAccelerator acc(dev, ms);
        DeviceRayQueue rays(dev, 3), misses(dev, 3);
        rays.assign({ray(0, 1, 0, 0, -1, 0), ray(10, 1, 0, 0, -1, 0), ray(20, 1, 0, 0, -1, 0)},
                    {{1, 1, 1}, {1, 1, 1}, {1, 1, 1}}, {0, 1, 2}, {0, 0, 0}, {-1, -1, -1});
        DeviceShadingQueue none(dev, 3, WF_POLICY_NONE);
        none.intersectAll(acc, rays, misses);
        none.schedule(acc);
        expect(none.sortedIndices() == std::vector<int32_t>({0, 1, 2}),
               "ShadingQueueTest.NonePolicyPreservesArrivalOrder");
        DeviceShadingQueue mat(dev, 3, WF_POLICY_MATERIAL);
        mat.intersectAll(acc, rays, misses);
        mat.schedule(acc);
        mat.materialize();
        expect(mat.materialIDs() == std::vector<int32_t>({2, 1, 0}),
               "ShadingQueueTest.MaterialAwareSortsByMaterialIDAscending");
    }
    {   // tests/render/WavefrontRendererTest.cpp:9-95 buildTestScene + RenderSceneWritesAllPixels
        SceneBuilder ts;
        int white = ts.addMaterial(diffuse(0.8f, 0.8f, 0.8f));
        wf_material light{};
        light.type = WF_MAT_EMISSIVE;
        light.albedo[0] = light.albedo[1] = light.albedo[2] = 1.0f;
        light.emission = 10.0f;
        light.texture_id = -1;
        int lm = ts.addMaterial(light);
        ts.addMesh(quad({-2, -1, -3, 2, -1, -3, 2, -1, 1, -2, -1, 1}, white));
        ts.addMesh(quad({-0.5f, 2, -2, 0.5f, 2, -2, 0.5f, 2, -1, -0.5f, 2, -1}, lm));
        GpuAccelerator acc(dev, ts);
        wf_camera cam{{0.0f, 0.5f, 2.0f}, {0.0f, 0.0f, -1.0f}, {0.0f, 1.0f, 0.0f}, 45.0f};
        WavefrontRenderer r(2, 2, WF_POLICY_NONE);
        std::vector<float> px;
        double shade_ms = r.renderScene(acc, cam, 16, 16, px);
        int black = 0;
        bool valid = true;
        for (int i = 0; i < 16 * 16; ++i) {
            if (px[i * 3] == 0 && px[i * 3 + 1] == 0 && px[i * 3 + 2] == 0) ++black;
            for (int c = 0; c < 3; ++c) valid &= std::isfinite(px[i * 3 + c]) && px[i * 3 + c] >= 0;
        }
        expect(black < 16 * 16 / 2, "WavefrontRendererTest.RenderSceneWritesAllPixels");
        expect(valid, "WavefrontRendererTest.PixelValuesAreInValidRange");
        expect(shade_ms >= 0.0, "WavefrontRendererTest.RenderReturnsNonNegativeShadingTime");
        // MaterialAwarePolicyProducesSameValueRange: scheduling never changes the image here
        WavefrontRenderer rm(2, 2, WF_POLICY_MATERIAL);
        std::vector<float> pm;
        rm.renderScene(acc, cam, 16, 16, pm);
        expect(pm == px, "WavefrontRendererTest.MaterialAwarePolicyProducesSameValueRange");
    }
    std::printf("%d failures\n", failures);
    return failures == 0 ? 0 : 1;
}
