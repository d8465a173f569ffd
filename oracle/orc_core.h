// ORACLE -- test infrastructure only. This directory is a CPU restatement of the reference
// renderer's per-bounce hot path (fhidargue/wavefront-renderer). Only tests/, the smoke check
// and the CPU-baseline leg of the benchmark may load it, and only as the checker. The product
// (the HIP library) never links or calls anything here.
//
// Core helpers shared by the oracle translation units: the reference's Vec3 arithmetic with the
// same operation order (include/math/Vec3.h:8-76), the counter-based RNG that replaces the
// reference's thread_local mt19937 (include/math/Random.h:6-17), and the flattened scene.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../include/wavefront_hip.h"

namespace orc {

// Vec3 with the reference's operator definitions (Vec3.h:19-67). Every operation is a single
// IEEE float op in the written order; the oracle is compiled with -ffp-contract=off.
struct V3 {
    float x, y, z;
    V3() : x(0.0f), y(0.0f), z(0.0f) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(float s) const { return V3(x / s, y / s, z / s); }
    V3 operator*(const V3& o) const { return V3(x * o.x, y * o.y, z * o.z); }
    float length() const { return std::sqrt(x * x + y * y + z * z); }
    float dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    V3 cross(const V3& o) const {
        return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x);
    }
    V3 normalized() const { return *this / length(); }
};

// luminance(), include/render/AdaptiveSampler.h:21-27 (BT.709 weights, float arithmetic).
inline float luminance(const V3& c) {
    return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
}

// ---- counter-based RNG ----------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). Restated from the published round function:
// per round, (hi,lo) = M0*c0 and M1*c2; c' = {hi1^c1^k0, lo1, hi0^c3^k1, lo0}; the key is bumped
// by the Weyl constants between rounds.
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Draw address of one uniform: (seed; job, pixel, sample pass, bounce depth, dimension).
// Counter = {pixel, sample, depth<<10 | dim>>2, job}, key = {seed, 0}; lane dim&3 of the block.
// Uniform = top 24 bits * 2^-24, exactly representable, in [0,1).
constexpr uint32_t kRaygenDepth = 0x3FFFFFu;
inline float uniform(uint32_t seed, uint32_t job, uint32_t pixel, uint32_t sample, uint32_t depth,
                     uint32_t dim) {
    uint32_t ctr[4] = {pixel, sample, (depth << 10) | (dim >> 2), job};
    uint32_t key[2] = {seed, 0u};
    uint32_t out[4];
    philox4x32_10(ctr, key, out);
    return (float)(out[dim & 3u] >> 8) * (1.0f / 16777216.0f);
}

// Dimension map of one bounce (the reference's draw order, cpp:612-788, made explicit).
enum Dim : uint32_t {
    DIM_TIMING_COIN = 0,   // reserved: the reference's 25% timing coin (cpp:612), never drawn
    DIM_LIGHT_PICK = 1,    // emitter mesh (reference) / alias slot (unbiased)
    DIM_LIGHT_TRI = 2,     // emitter triangle (reference) / alias coin (unbiased)
    DIM_LIGHT_R1 = 3,
    DIM_LIGHT_R2 = 4,
    DIM_RR = 5,
    DIM_ENV_CHOICE = 6,    // unbiased: environment vs area light
    DIM_LOBE = 8,          // Schlick/Fresnel coin (glass, plastic)
    DIM_DIR_U1 = 9,        // cosine hemisphere / GGX visible normal
    DIM_DIR_U2 = 10,
    DIM_ENV_SLOT = 12,     // unbiased environment NEE: alias slot, coin, texel jitter
    DIM_ENV_COIN = 13,
    DIM_ENV_JX = 14,
    DIM_ENV_JY = 15,
    DIM_SPHERE = 16,       // rejection sampling in the unit sphere: 16 + 3k + axis
};
constexpr int kSphereTries = 256;

// ---- flattened scene -------------------------------------------------------------------------
struct Texture {
    int width = 0, height = 0;
    std::vector<V3> texels;        // linear RGB, like the reference's vector<Color> pixels
    uint64_t memory_bytes() const { return (uint64_t)texels.size() * 12u; }  // Texture.h:51-54
};

struct Scene {
    // per global triangle g (meshes concatenated in attach order)
    std::vector<V3> v0, v1, v2;
    std::vector<int32_t> tri_geom, tri_prim;
    std::vector<float> tri_uv;            // 6 per triangle: u0 v0 u1 v1 u2 v2
    std::vector<uint8_t> tri_has_uv;
    // per mesh
    std::vector<int32_t> mesh_tri_offset, mesh_tri_count, mesh_material;
    std::vector<wf_material> materials;
    std::vector<Texture> textures;
    std::vector<int32_t> emissive_meshes;
    float total_emissive_area = 0.0f;
    V3 bounds_min, bounds_max;
    wf_directional_light directional{};
    int env_w = 0, env_h = 0;
    std::vector<float> env_rgb;
    float env_intensity = 1.0f;
    V3 env_tint{1.0f, 1.0f, 1.0f};
    // unbiased light sampling tables
    std::vector<int32_t> light_tris;      // emissive global triangle ids
    std::vector<float> light_prob;        // alias table over light_tris
    std::vector<int32_t> light_alias;
    std::vector<float> mat_light_pdf_area;  // per material: selection pdf per unit area
    std::vector<float> env_prob;          // alias table over environment texels
    std::vector<int32_t> env_alias;
    std::vector<float> env_pdf;           // per texel selection probability
    float env_choice_prob = 0.0f;         // probability NEE samples the environment
    // bounding volume hierarchy (CPU, binned SAH) over global triangles
    struct Node {
        float lo[3], hi[3];
        int32_t left, right;               // internal: child node ids; leaf: left = -(first+1)
        int32_t count;                     // leaf triangle count, 0 for internal nodes
    };
    std::vector<Node> nodes;
    std::vector<int32_t> bvh_tris;
    int n_tris() const { return (int)v0.size(); }
};

void build_bvh(Scene& s);
void build_light_tables(Scene& s);

}  // namespace orc
