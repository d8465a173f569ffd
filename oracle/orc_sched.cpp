// ORACLE -- test infrastructure only (see orc_core.h). Restatement of the reference's queue
// scheduling: RayQueue::sortKey/schedule/materialize and ShadingQueue::schedule for the four
// policies. Where the reference uses an unstable sort (tbb::parallel_sort) the oracle sorts
// stably, i.e. ties keep ascending queue index: that order is always one of the outputs the
// reference comparator admits.
#include <algorithm>
#include <map>
#include <numeric>

#include "orc_core.h"

using namespace orc;

namespace {

// spreadBits10, include/scheduling/RayQueue.h:45-55 (masks RayQueue.h:32-40).
uint32_t spread_bits10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x30000FFu;
    v = (v | (v << 8)) & 0x300F00Fu;
    v = (v | (v << 4)) & 0x30C30C3u;
    v = (v | (v << 2)) & 0x9249249u;
    return v;
}

// mortonEncode3, RayQueue.h:57-60.
uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return spread_bits10(x) | (spread_bits10(y) << 1) | (spread_bits10(z) << 2);
}

// quantizeToGrid lambda of RayQueue::sortKey, RayQueue.h:147-153.
uint32_t quantize_ray(float value, float min_value, float extent) {
    float normalized = (value - min_value) / extent;
    normalized = std::clamp(normalized, 0.0f, 1.0f);
    return static_cast<uint32_t>(normalized * (float)(1024u - 1u));
}

// computeMortonCode quantise lambda, ShadingQueue.h:523-529.
uint32_t quantize_hit(float value, float lo, float hi) {
    float normalized = (hi > lo) ? (value - lo) / (hi - lo) : 0.0f;
    return static_cast<uint32_t>(std::clamp(normalized, 0.0f, 1.0f) * (float)1023u);
}

}  // namespace

extern "C" {

// RayQueue::sortKey, RayQueue.h:135-162.
void orc_ray_sort_keys(int32_t n, const float* ox, const float* oy, const float* oz,
                       const float* dx, const float* dy, const float* dz, const float* bmin,
                       const float* bmax, uint64_t* keys) {
    float ex = std::max(bmax[0] - bmin[0], 1e-6f);
    float ey = std::max(bmax[1] - bmin[1], 1e-6f);
    float ez = std::max(bmax[2] - bmin[2], 1e-6f);
    for (int32_t i = 0; i < n; ++i) {
        uint64_t octant = (dx[i] < 0.0f ? 1u : 0u) | (dy[i] < 0.0f ? 2u : 0u) |
                          (dz[i] < 0.0f ? 4u : 0u);
        uint32_t gx = quantize_ray(ox[i], bmin[0], ex);
        uint32_t gy = quantize_ray(oy[i], bmin[1], ey);
        uint32_t gz = quantize_ray(oz[i], bmin[2], ez);
        keys[i] = (octant << 30) | (uint64_t)morton3(gx, gy, gz);
    }
}

// RayQueue::schedule, RayQueue.h:165-179, with a stable sort.
void orc_ray_schedule(int32_t n, const float* ox, const float* oy, const float* oz,
                      const float* dx, const float* dy, const float* dz, const float* bmin,
                      const float* bmax, int32_t* sorted) {
    std::vector<uint64_t> keys(n);
    orc_ray_sort_keys(n, ox, oy, oz, dx, dy, dz, bmin, bmax, keys.data());
    std::iota(sorted, sorted + n, 0);
    std::stable_sort(sorted, sorted + n,
                     [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
}

// ShadingQueue::computeMortonCode, ShadingQueue.h:513-550 (exported for tests).
uint64_t orc_hit_morton(float x, float y, float z, float min_x, float min_y, float min_z,
                        float max_x, float max_y, float max_z) {
    auto spread = [](uint32_t value) -> uint64_t {
        uint64_t e = value & 0x3ffull;
        e = (e | e << 16) & 0x30000ffull;
        e = (e | e << 8) & 0x300f00full;
        e = (e | e << 4) & 0x30c30c3ull;
        e = (e | e << 2) & 0x9249249ull;
        return e;
    };
    return (spread(quantize_hit(x, min_x, max_x)) << 0) |
           (spread(quantize_hit(y, min_y, max_y)) << 1) |
           (spread(quantize_hit(z, min_z, max_z)) << 2);
}

// ShadingQueue::schedule, ShadingQueue.h:90-141 (None / MaterialAware / TextureAware),
// computeTypeOrder ShadingQueue.h:277-329, scheduleCostBenefitAware ShadingQueue.h:437-510.
//   mat_type[m], mat_tex[m]: material table; tex_mem[t]: Texture::memoryBytes();
//   mat_cost / tex_cost: CostTracker::relativeCost tables (NULL = neutral, every cost 1.0).
// Returns 0, or -1 for an unknown policy.
int32_t orc_shading_schedule(int32_t policy, int32_t n, const int32_t* mat, const int32_t* tex,
                             const float* tr, const float* tg, const float* tb, const float* px,
                             const float* py, const float* pz, int32_t n_mats,
                             const int32_t* mat_type, const int32_t* mat_tex, int32_t n_tex,
                             const uint64_t* tex_mem, const double* mat_cost,
                             const double* tex_cost, int32_t* sorted) {
    (void)n_mats;
    std::iota(sorted, sorted + n, 0);
    if (policy == WF_POLICY_NONE) return 0;
    auto rel_mat = [&](int32_t m) { return mat_cost ? mat_cost[m] : 1.0; };
    // an id beyond the table is an untracked texture (relativeCost 1.0); the reference indexes
    // its table with any id >= 0
    auto rel_tex = [&](int32_t t) { return tex_cost && t < n_tex ? tex_cost[t] : 1.0; };

    if (policy == WF_POLICY_COST_BENEFIT) {
        if (n == 0) return 0;
        std::map<int32_t, double> sum;
        for (int32_t i = 0; i < n; ++i) {
            V3 t(tr[i], tg[i], tb[i]);
            sum[mat[i]] += luminance(t);
        }
        std::map<int32_t, double> priority;
        for (auto& e : sum) {
            int32_t m = e.first;
            double shading_cost = rel_mat(m);
            int32_t tid = mat_tex[m];
            double texture_cost = (tid >= 0) ? rel_tex(tid) : 0.0;
            priority[m] = e.second / std::max(shading_cost + texture_cost, 1e-6);
        }
        float mnx = *std::min_element(px, px + n), mny = *std::min_element(py, py + n),
              mnz = *std::min_element(pz, pz + n);
        float mxx = *std::max_element(px, px + n), mxy = *std::max_element(py, py + n),
              mxz = *std::max_element(pz, pz + n);
        std::vector<uint64_t> morton(n);
        for (int32_t i = 0; i < n; ++i)
            morton[i] = orc_hit_morton(px[i], py[i], pz[i], mnx, mny, mnz, mxx, mxy, mxz);
        std::stable_sort(sorted, sorted + n, [&](int32_t a, int32_t b) {
            int32_t ta = mat_type[mat[a]], tb2 = mat_type[mat[b]];
            if (ta != tb2) return ta < tb2;
            double pa = priority[mat[a]], pb = priority[mat[b]];
            if (pa != pb) return pa > pb;
            return morton[a] < morton[b];
        });
        return 0;
    }
    if (policy != WF_POLICY_MATERIAL && policy != WF_POLICY_TEXTURE) return -1;

    auto tex_bytes = [&](int32_t t) -> uint64_t {
        return (t >= 0 && t < n_tex) ? tex_mem[t] : 0u;
    };
    // computeTypeOrder
    std::map<int32_t, int64_t> count_by_type;
    std::map<int32_t, uint64_t> memory_by_type;
    for (int32_t i = 0; i < n; ++i) {
        int32_t ty = mat_type[mat[i]];
        count_by_type[ty] += 1;
        memory_by_type[ty] += tex_bytes(tex[i]);
    }
    std::vector<int32_t> order;
    for (auto& e : count_by_type) order.push_back(e.first);
    if (policy == WF_POLICY_MATERIAL) {
        // ShadingQueue.h:306-312: ray count descending, then type index descending.
        std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            if (count_by_type[a] != count_by_type[b]) return count_by_type[a] > count_by_type[b];
            return a > b;
        });
    } else {
        // ShadingQueue.h:315-316: texture bytes descending. The reference leaves equal sums in
        // unordered_map iteration order; the canonical tie order here is type index descending.
        std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            if (memory_by_type[a] != memory_by_type[b]) return memory_by_type[a] > memory_by_type[b];
            return a > b;
        });
    }
    int32_t rank[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t r = 0; r < (int32_t)order.size(); ++r) rank[order[r]] = r;
    std::stable_sort(sorted, sorted + n, [&](int32_t a, int32_t b) {
        int32_t ra = rank[mat_type[mat[a]]], rb = rank[mat_type[mat[b]]];
        if (ra != rb) return ra < rb;
        if (policy == WF_POLICY_TEXTURE) return tex_bytes(tex[a]) > tex_bytes(tex[b]);
        return rel_mat(mat[a]) > rel_mat(mat[b]);
    });
    return 0;
}

// materialize (RayQueue.h:182-215, ShadingQueue.h:233-275) for one 32-bit column:
// dst[i] = src[idx[i]].
void orc_gather_u32(int32_t n, const uint32_t* src, const int32_t* idx, uint32_t* dst) {
    for (int32_t i = 0; i < n; ++i) dst[i] = src[idx[i]];
}

}  // extern "C"
