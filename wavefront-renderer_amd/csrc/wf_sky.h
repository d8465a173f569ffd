// What a ray that leaves the scene contributes: the reference's sky gradient (getSkyColor,
// src/render/WavefrontRenderer.cpp:879-889) or the environment map looked up at its direction
// (EnvironmentMap::sample, src/core/EnvironmentMap.cpp:30-48), weighted on the unbiased track by
// the power heuristic against the environment's light-sampling pdf (include/shading/MIS.h).
#pragma once

#include "wf_device.h"
#include "wf_fmath.h"
#include "wf_internal.h"

namespace wf {

constexpr float kPi = 3.14159265f;

// misBalance and the power heuristic (include/shading/MIS.h:6-31)
__device__ __forceinline__ float mis_balance(float a, float b) {
    const float den = a + b;
    return den <= 0.0f ? 0.0f : a / den;
}
__device__ __forceinline__ float mis_power(float a, float b) {
    const float a2 = a * a, b2 = b * b;
    const float den = a2 + b2;
    return den <= 0.0f ? 0.0f : a2 / den;
}

// EnvironmentMap::sample mapping (EnvironmentMap.cpp:30-48): nearest texel of the direction.
__device__ __forceinline__ int env_texel(const SceneDev& s, const V3& dir, float& lat_cos) {
    const V3 d = dir.normalized();
    const float u = 0.5f + fm::atan2(d.z, d.x) / (2.0f * 3.14159274f);
    const float v = 0.5f - fm::asin(smax(-1.0f, smin(1.0f, d.y))) / 3.14159274f;
    int x = (int)(u * s.env_w);
    int y = (int)(v * s.env_h);
    x = x < 0 ? 0 : (x > s.env_w - 1 ? s.env_w - 1 : x);
    y = y < 0 ? 0 : (y > s.env_h - 1 ? s.env_h - 1 : y);
    lat_cos = __builtin_sqrtf(smax(0.0f, 1.0f - d.y * d.y));
    return y * s.env_w + x;
}
__device__ __forceinline__ V3 env_radiance(const SceneDev& s, int texel_i) {
    const float4 r = s.env[texel_i];
    return V3(r.x, r.y, r.z) * s.env_intensity * V3(s.env_tint[0], s.env_tint[1], s.env_tint[2]);
}
__device__ __forceinline__ float env_pdf_sa(const SceneDev& s, int texel_i, float lc) {
    if (s.env_choice_prob <= 0.0f || lc <= 0.0f) return 0.0f;
    const float p = s.env_choice_prob * s.env_pdf[texel_i] * (float)s.env_w * (float)s.env_h;
    return p / (2.0f * kPi * kPi * lc);
}

// Adds the contribution of a missed ray (direction d, throughput T, accumulator index, depth and
// scatter pdf) to its accumulator (a plain read-modify-write: a path owns its accumulator).
__device__ __forceinline__ void add_miss(const SceneDev& s, const V3& d, const V3& T,
                                         int32_t pixel, int32_t depth, float pdf, int physics,
                                         float* __restrict__ accum) {
    V3 c;
    if (s.env_w > 0) {
        float lc;
        const int tx = env_texel(s, d, lc);
        float w = 1.0f;
        if (physics == WF_PHYSICS_UNBIASED && depth > 0 && pdf > 0.0f)
            w = mis_power(pdf, env_pdf_sa(s, tx, lc));
        c = physics == WF_PHYSICS_UNBIASED ? T * env_radiance(s, tx) * w : T * env_radiance(s, tx);
    } else {
        const float blend = 0.5f * (d.normalized().y + 1.0f);
        const V3 dark(0.02f, 0.02f, 0.05f), horizon(0.05f, 0.04f, 0.03f);
        c = T * (horizon * (1.0f - blend) + dark * blend);
    }
    float* acc = accum + (size_t)pixel * 3;
    acc[0] = acc[0] + c.x;
    acc[1] = acc[1] + c.y;
    acc[2] = acc[2] + c.z;
}

}  // namespace wf
