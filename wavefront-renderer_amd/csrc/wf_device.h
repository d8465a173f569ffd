// Device-side helpers shared by the kernels: the reference's Vec3 arithmetic (one IEEE op per
// operator, evaluated in the written order; the library is compiled with -ffp-contract=off so
// no multiply-add is fused), the Philox4x32-10 counter RNG and small wave utilities.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wf {

struct V3 {
    float x, y, z;
    __device__ __forceinline__ V3() : x(0.0f), y(0.0f), z(0.0f) {}
    __device__ __forceinline__ V3(float a, float b, float c) : x(a), y(b), z(c) {}
    __device__ __forceinline__ V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    __device__ __forceinline__ V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    __device__ __forceinline__ V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
    __device__ __forceinline__ V3 operator/(float s) const { return V3(x / s, y / s, z / s); }
    __device__ __forceinline__ V3 operator*(const V3& o) const { return V3(x * o.x, y * o.y, z * o.z); }
    __device__ __forceinline__ float dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    __device__ __forceinline__ V3 cross(const V3& o) const {
        return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x);
    }
    __device__ __forceinline__ float length() const { return __builtin_sqrtf(x * x + y * y + z * z); }
    __device__ __forceinline__ V3 normalized() const { return *this / length(); }
};

__device__ __forceinline__ V3 ld3(const float* x, const float* y, const float* z, int i) {
    return V3(x[i], y[i], z[i]);
}
__device__ __forceinline__ void st3(float* x, float* y, float* z, int i, const V3& v) {
    x[i] = v.x;
    y[i] = v.y;
    z[i] = v.z;
}
__device__ __forceinline__ V3 xyz(const float4& f) { return V3(f.x, f.y, f.z); }

// BT.709 luminance with the reference's float evaluation order.
__device__ __forceinline__ float luminance(const V3& c) {
    return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
}

// std::min/std::max semantics for non-NaN operands: min(a,b) = (b < a) ? b : a.
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

// ---- Philox4x32-10 counter-based generator -------------------------------------------------
// Counter {pixel, sample, depth<<10 | dim>>2, job}, key {seed, 0}; output lane dim&3;
// uniform = (x >> 8) * 2^-24 in [0,1).
struct Philox {
    uint32_t v[4];
    __device__ __forceinline__ Philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                      uint32_t k0, uint32_t k1) {
#pragma unroll
        for (int r = 0; r < 10; ++r) {
            uint32_t lo0 = 0xD2511F53u * c0;
            uint32_t hi0 = __umulhi(0xD2511F53u, c0);
            uint32_t lo1 = 0xCD9E8D57u * c2;
            uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
            uint32_t n0 = hi1 ^ c1 ^ k0;
            uint32_t n2 = hi0 ^ c3 ^ k1;
            c0 = n0;
            c1 = lo1;
            c2 = n2;
            c3 = lo0;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        v[0] = c0;
        v[1] = c1;
        v[2] = c2;
        v[3] = c3;
    }
};

constexpr uint32_t kRaygenDepth = 0x3FFFFFu;

struct PathRng {
    uint32_t seed, job, pixel, sample, depth;
    // one Philox block (4 dims) cached
    uint32_t cached_block = 0xFFFFFFFFu;
    uint32_t cache[4];
    __device__ __forceinline__ float operator()(uint32_t dim) {
        uint32_t blk = dim >> 2;
        if (blk != cached_block) {
            Philox p(pixel, sample, (depth << 10) | blk, job, seed, 0u);
            cache[0] = p.v[0];
            cache[1] = p.v[1];
            cache[2] = p.v[2];
            cache[3] = p.v[3];
            cached_block = blk;
        }
        const uint32_t lane = dim & 3u;
        uint32_t x = lane == 0 ? cache[0] : (lane == 1 ? cache[1] : (lane == 2 ? cache[2] : cache[3]));
        return (float)(x >> 8) * (1.0f / 16777216.0f);
    }
};

enum : uint32_t {
    D_COST_COIN = 0,   // shouldTimeThisRay (cpp:612): no effect on the path
    D_LIGHT_PICK = 1,
    D_LIGHT_TRI = 2,
    D_LIGHT_R1 = 3,
    D_LIGHT_R2 = 4,
    D_RR = 5,
    D_ENV_CHOICE = 6,
    D_LOBE = 8,
    D_DIR_U1 = 9,
    D_DIR_U2 = 10,
    D_ENV_SLOT = 12,
    D_ENV_COIN = 13,
    D_ENV_JX = 14,
    D_ENV_JY = 15,
    D_SPHERE = 16,
};
constexpr int kSphereTries = 256;

// ---- wave utilities (wave64) ----------------------------------------------------------------
__device__ __forceinline__ uint64_t lanemask_lt() {
    const unsigned lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64u - lane));
}

// Camera::getRay for pixel (x, y) with jitter (jx, jy) (cpp:401-457, Camera.h:35-40); cam = eye,
// top-left corner, horizontal, vertical.
__device__ __forceinline__ V3 primary_dir(const float* cam, int x, int y, int width, int height,
                                          float jx, float jy) {
    const float u = ((float)x + jx) / (float)(width - 1);
    const float v = 1.0f - ((float)y + jy) / (float)(height - 1);
    const V3 O(cam[0], cam[1], cam[2]), TL(cam[3], cam[4], cam[5]), H(cam[6], cam[7], cam[8]);
    const V3 Vv(cam[9], cam[10], cam[11]);
    return (TL + H * u + Vv * v - O).normalized();
}

// Unordered append of the lanes with `take` set at a device cursor, one global atomic per block.
// Every thread of the block must call it (same number of times); returns this lane's slot or -1.
// `scratch` is two words of LDS owned by the caller.
__device__ __forceinline__ int block_append(int32_t* cursor, bool take, uint32_t* scratch) {
    if (threadIdx.x == 0) scratch[0] = 0u;
    __syncthreads();
    const uint64_t b = __ballot(take);
    uint32_t off = 0;
    if (__lane_id() == 0 && b) off = atomicAdd(&scratch[0], (uint32_t)__popcll(b));
    off = __shfl(off, 0, 64);
    __syncthreads();
    if (threadIdx.x == 0) scratch[1] = scratch[0] ? (uint32_t)atomicAdd(cursor, (int32_t)scratch[0]) : 0u;
    __syncthreads();
    const uint32_t base = scratch[1];
    __syncthreads();   // scratch is reused by the next call
    return take ? (int)(base + off + (uint32_t)__popcll(b & lanemask_lt())) : -1;
}

// Appends the items of ITEMS rounds (round j of every thread of the block, then round j + 1) to
// a list at *cursor with one cursor atomic per block (device-scope atomics on one address
// serialise, so fewer and larger claims keep the cursor off the critical path). Slots run in
// round order, then wave order, then lane order; slot[j] = (T)-1 where take[j] is false. Every
// thread of the BLOCK-thread block must call it; lds holds at least ITEMS * BLOCK / 64 words
// owned by the caller. T: int32_t or unsigned long long cursors.
template <int ITEMS, int BLOCK, typename T>
__device__ __forceinline__ void block_append_rounds(T* cursor, const bool (&take)[ITEMS],
                                                    T (&slot)[ITEMS], uint32_t* lds) {
    constexpr int W = BLOCK / 64;
    constexpr int E = ITEMS * W;                 // (round, wave) counts
    constexpr int PER = (E + 63) / 64;           // counts per lane of the scanning wave
    __shared__ T s_base;
    const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
    uint64_t b[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        b[j] = __ballot(take[j]);
        if (lane == 0) lds[j * W + wave] = (uint32_t)__popcll(b[j]);
    }
    __syncthreads();
    if (wave == 0) {
        // exclusive scan of the E counts in (round, wave) order; lane l owns PER consecutive
        uint32_t c[PER], sum = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = lane * PER + q;
            c[q] = e < E ? lds[e] : 0u;
            sum += c[q];
        }
        uint32_t inc = sum;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        const uint32_t total = __shfl(inc, 63, 64);
        uint32_t run = inc - sum;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = lane * PER + q;
            if (e < E) lds[e] = run;
            run += c[q];
        }
        if (lane == 0) s_base = total ? atomicAdd(cursor, (T)total) : (T)0;
    }
    __syncthreads();
    const T base = s_base;
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
        slot[j] = take[j] ? base + (T)(lds[j * W + wave] + (uint32_t)__popcll(b[j] & lt)) : (T)-1;
    __syncthreads();   // lds and s_base are reused by the next call
}

// Appends c items per thread (any count) to a list at *cursor with one cursor atomic per block;
// returns the first slot of this thread's run (runs in thread order). Every thread of the
// BLOCK-thread block must call it; lds holds at least BLOCK / 64 words owned by the caller.
template <int BLOCK, typename T>
__device__ __forceinline__ T block_append_counts(T* cursor, uint32_t c, uint32_t* lds) {
    constexpr int W = BLOCK / 64;
    __shared__ T s_base;
    const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
    uint32_t inc = c;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < W; ++w) {
            const uint32_t x = lds[w];
            lds[w] = run;
            run += x;
        }
        s_base = run ? atomicAdd(cursor, (T)run) : (T)0;
    }
    __syncthreads();
    const T slot = s_base + (T)(lds[wave] + inc - c);
    __syncthreads();   // lds and s_base are reused by the next call
    return slot;
}

// Element count of a launch: the device counter when present, else the host value.
__device__ __forceinline__ int load_count(int n, const int32_t* __restrict__ dev) {
    return dev ? *dev : n;
}

}  // namespace wf
