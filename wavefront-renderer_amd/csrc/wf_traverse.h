// Ray traversal shared by the closest-hit, occlusion and shadow kernels: the reference's
// Moller-Trumbore test (include/geometry/Triangle.h:6-42) with the same float operations, a
// conservative slab test, and a stack walk over the four- or eight-wide BVH. The first kStackLds stack
// entries live in LDS (stack[entry * kTraceBlock + lane], so a wave's pushes and pops hit 64
// different banks); deeper entries, which few rays reach, spill to a per-lane private array.
#pragma once

#include <type_traits>

#include "wf_device.h"
#include "wf_internal.h"
#include "wf_rays.h"

namespace wf {

constexpr int kTraceBlock = 128;
// Largest traversal stack (entries per lane) a kernel is instantiated for; the scene build
// measures the exact bound and the launch picks the smallest instantiation >= it.
constexpr int kStackMax = 64;
// Stack entries per lane kept in LDS (8 KB per 128-lane block, so LDS never limits occupancy).
constexpr int kStackLds = 16;
template <int STACK>
constexpr int lds_entries() { return STACK < kStackLds ? STACK : kStackLds; }

// Calls f(std::integral_constant<int, S>) for the smallest stack size S in {16, 32, 48, 64} that
// holds `bound` entries.
template <class F>
inline void with_stack(int bound, F&& f) {
    if (bound <= 16) f(std::integral_constant<int, 16>{});
    else if (bound <= 32) f(std::integral_constant<int, 32>{});
    else if (bound <= 48) f(std::integral_constant<int, 48>{});
    else f(std::integral_constant<int, 64>{});
}

// Calls f(nodes, std::integral_constant<int, S>) with the scene's node array (four- or
// eight-wide) and the stack instantiation for its stack bound.
template <class F>
inline void with_bvh(const SceneDev& d, F&& f) {
    with_stack(d.stack_bound, [&](auto S) {
        if (d.width == 8) f((const Bvh8Node*)d.nodes8, S);
        else f((const Bvh4Node*)d.nodes4, S);
    });
}

__device__ __forceinline__ bool mt_test(const V3& o, const V3& d, const V3& v0, const V3& e1,
                                        const V3& e2, float tmin, float tmax, float& t, float& u,
                                        float& v) {
    const float EPSILON = 0.0000001f;
    V3 pvec = d.cross(e2);
    float det = e1.dot(pvec);
    if (det > -EPSILON && det < EPSILON) return false;
    float inv = 1.0f / det;
    V3 tvec = o - v0;
    float uu = inv * tvec.dot(pvec);
    if (uu < 0.0f || uu > 1.0f) return false;
    V3 qvec = tvec.cross(e1);
    float vv = inv * d.dot(qvec);
    if (vv < 0.0f || uu + vv > 1.0f) return false;
    float tt = inv * e2.dot(qvec);
    if (tt < tmin || tt > tmax) return false;
    t = tt;
    u = uu;
    v = vv;
    return true;
}

// 1/b correctly rounded, as the reference's `1.0f / det` (Triangle.h:20). The hardware
// reciprocal is faithful (error below one ulp); one fma correction y + y (1 - b y) of any faithful
// reciprocal is the correctly rounded one (Markstein), except when b's significand is all ones,
// which takes the division, as do results outside the normal range (checked for every
// significand and every faithful approximation on the host, and for every significand on the
// device against the division).
__device__ __forceinline__ float rcp_cr(float b) {
    const float y = __builtin_amdgcn_rcpf(b);
    float z = __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
    const float a = fabsf(b);
    if ((__float_as_uint(b) & 0x7FFFFFu) == 0x7FFFFFu || !(a >= 0x1p-125f && a <= 0x1p125f)) {
        // an empty volatile statement keeps the division inside this branch (the compiler would
        // otherwise evaluate it for every lane and select)
        float bb = b;
        asm volatile("" : "+v"(bb));
        z = 1.0f / bb;
    }
    return z;
}

__device__ __forceinline__ float safe_rcp(float x) {
    float a = fabsf(x) < 1e-20f ? copysignf(1e-20f, x) : x;
    return 1.0f / a;
}

// The float next to f towards +inf (f finite); towards -inf as -next_up(-f).
__device__ __forceinline__ float next_up(float f) {
    const uint32_t b = __float_as_uint(f);
    if ((b & 0x7FFFFFFFu) == 0u) return __uint_as_float(1u);
    return __uint_as_float((int32_t)b >= 0 ? b + 1u : b - 1u);
}

// Entry distance of the box [lo, hi] along the ray, clipped to [0, tmax]; true when the true ray
// meets the box inside [0, tmax]. The planes are cut with the rule of boxes4 below (one fma per
// plane with o inv rounded outward, exit distance times 1 + 2^-21). The plain (x - o) inv would
// drop a box whose face holds the origin when the direction has a zero component: 0 * inv = 0
// makes the exit distance 0 where the ray in fact runs inside the slab.
__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz,
                                     const V3& inv, const V3& nup, const V3& ndn, float tmax,
                                     float& tenter) {
    // inv is never zero (safe_rcp), so its sign says which plane comes first
    const bool fx = inv.x < 0.0f, fy = inv.y < 0.0f, fz = inv.z < 0.0f;
    const float a = fmaxf(fmaxf(__builtin_fmaf(fx ? hx : lx, inv.x, nup.x),
                                __builtin_fmaf(fy ? hy : ly, inv.y, nup.y)),
                          fmaxf(__builtin_fmaf(fz ? hz : lz, inv.z, nup.z), 0.0f));
    const float b = fminf(fminf(__builtin_fmaf(fx ? lx : hx, inv.x, ndn.x),
                                __builtin_fmaf(fy ? ly : hy, inv.y, ndn.y)),
                          fminf(__builtin_fmaf(fz ? lz : hz, inv.z, ndn.z), tmax));
    tenter = a;
    return a <= b * 1.00000048f;
}

// The four child boxes of a four-wide node: entry distance t0[k] (clipped to [0, tmax]) and
// whether the ray enters the box. Each axis is cut by the plane the ray meets first and the one
// it meets last: lo then hi for a non-negative inverse direction, hi then lo otherwise. A plane
// at x is met at (x - o) inv = x inv - o inv; with p = o inv rounded, the first plane's distance
// is fma(x, inv, -next_up(p)) and the last one's fma(x, inv, -next_down(p)) (nup, ndn hold
// -next_up(p) and -next_down(p)). next_up(p) >= o inv >= next_down(p) exactly and rounding is
// monotonic, so a computed entry distance is at most the rounded exact (x - o) inv and a computed
// exit distance at least the rounded exact one: one fma per plane, with no cancellation error.
// inv is the rounded 1/d (relative error u = 2^-24), so against the true ray o + t d an entry
// distance is at most (1 + u)^2 and an exit distance at least (1 - u)^2 times the true one; the
// exit side's factor 1 + 2^-21 >= (1 + u)^2 / (1 - u)^3 (its product rounded once more) makes
// the test accept every box that the true ray meets inside [0, tmax]. The node's rows are read
// at byte offsets from the (uniform) node array: node * 128 + row * 16 with rows lo x, y, z, hi
// x, y, z; an axis whose inverse direction is negative reads its hi row first. An empty slot
// holds the inverted infinite box, whose entry distance is +inf and exit distance -inf on every
// axis.
__device__ __forceinline__ void boxes4(const char* __restrict__ nb, uint32_t off, const V3& inv,
                                       const V3& nup, const V3& ndn, float tmax, float (&t0)[4],
                                       bool (&in)[4]) {
    // inv is never zero (safe_rcp), so its sign bit says which plane comes first
    const uint32_t fxo = (uint32_t)(__float_as_int(inv.x) >> 31) & 48u;
    const uint32_t fyo = (uint32_t)(__float_as_int(inv.y) >> 31) & 48u;
    const uint32_t fzo = (uint32_t)(__float_as_int(inv.z) >> 31) & 48u;
    // (the row constants stay outside the 32-bit sums so that they fold into the loads' offsets;
    // fxo ^ 48 = 48 - fxo for fxo in {0, 48})
    const float4 nx = *reinterpret_cast<const float4*>(nb + (size_t)(off + fxo));
    const float4 ny = *reinterpret_cast<const float4*>(nb + (size_t)(off + fyo) + 16);
    const float4 nz = *reinterpret_cast<const float4*>(nb + (size_t)(off + fzo) + 32);
    const float4 ex = *reinterpret_cast<const float4*>(nb + (size_t)(off + (fxo ^ 48u)));
    const float4 ey = *reinterpret_cast<const float4*>(nb + (size_t)(off + (fyo ^ 48u)) + 16);
    const float4 ez = *reinterpret_cast<const float4*>(nb + (size_t)(off + (fzo ^ 48u)) + 32);
    const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, nya[4] = {ny.x, ny.y, ny.z, ny.w};
    const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, exa[4] = {ex.x, ex.y, ex.z, ex.w};
    const float eya[4] = {ey.x, ey.y, ey.z, ey.w}, eza[4] = {ez.x, ez.y, ez.z, ez.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float a = fmaxf(fmaxf(__builtin_fmaf(nxa[k], inv.x, nup.x),
                                    __builtin_fmaf(nya[k], inv.y, nup.y)),
                              fmaxf(__builtin_fmaf(nza[k], inv.z, nup.z), 0.0f));
        const float b = fminf(fminf(__builtin_fmaf(exa[k], inv.x, ndn.x),
                                    __builtin_fmaf(eya[k], inv.y, ndn.y)),
                              fminf(__builtin_fmaf(eza[k], inv.z, ndn.z), tmax));
        t0[k] = a;
        in[k] = a <= b * 1.00000048f;
    }
}

struct TraceHit {
    float t;
    int32_t leaf;   // leaf (sorted) triangle index, -1 = none
    int32_t g;      // global triangle index
    float u, v;
    int32_t nodes = 0, tris = 0;   // visit counts (STATS instantiations only)
};

// Moller-Trumbore without early exits: every quantity is computed and `ok` is the conjunction of
// the negated early-return conditions of mt_test, so accepted triangles get the same t, u, v.
// Computing all of them keeps the triangle loads unconditional (one memory round trip for all
// candidate triangles of a node).
__device__ __forceinline__ bool mt_all(const V3& o, const V3& d, const V3& v0, const V3& e1,
                                       const V3& e2, float tmin, float tmax, float& t, float& u,
                                       float& v) {
    const float EPSILON = 0.0000001f;
    const V3 pvec = d.cross(e2);
    const float det = e1.dot(pvec);
    const float inv = rcp_cr(det);
    const V3 tvec = o - v0;
    u = inv * tvec.dot(pvec);
    const V3 qvec = tvec.cross(e1);
    v = inv * d.dot(qvec);
    t = inv * e2.dot(qvec);
    const bool det_ok = !(det > -EPSILON && det < EPSILON);
    const bool u_ok = !(u < 0.0f || u > 1.0f);
    const bool v_ok = !(v < 0.0f || u + v > 1.0f);
    const bool t_ok = !(t < tmin || t > tmax);
    return det_ok && u_ok && v_ok && t_ok;
}

// Candidate triangles a lane of the cooperative walk holds between test rounds, and the wave's
// LDS for that walk: 64 int4 job/result slots, then kStash int32 per lane.
constexpr int kStash = 6;
constexpr int kXchgPerWave = 64 + kStash * 64 / 4;
// a lane visits a node only with room for its four leaves, and the wave sums the stash counts
// from three ballot bits
static_assert(kStash >= 4 && kStash <= 7, "stash counts must fit three bits");

// Resumable walk over the four-wide BVH. At each node the four child boxes are tested; leaf
// children are tested at once, the hit inner children are visited nearest first (the others are
// pushed far to near). ANY stops at the first accepted triangle. The stack never holds more than
// the build's stack bound, which STACK covers: the first lds_entries<STACK>() entries live in the
// block's LDS column array (stack[entry * kTraceBlock + lane]), deeper ones in `spill`. All child
// boxes of a node, and all candidate triangles of a node, are loaded and tested together so that
// each costs a single memory round trip.
template <bool ANY, int STACK, bool STATS = false>
struct Walker {
    static constexpr int L = lds_entries<STACK>();
    static constexpr int X = STACK - L;
    V3 o, d, inv;
    V3 nup, ndn;   // -next_up(o inv), -next_down(o inv) per axis (the box test's plane offsets)
    float tnear, tfar;
    TraceHit h;
    int node, sp;
    int pc = 0;    // candidates in the stash, not tested yet (cooperative walk)

    __device__ __forceinline__ void start(const V3& ro, const V3& rd, float t0, float t1) {
        o = ro;
        d = rd;
        inv = V3(safe_rcp(rd.x), safe_rcp(rd.y), safe_rcp(rd.z));
        const V3 p = ro * inv;
        nup = V3(-next_up(p.x), -next_up(p.y), -next_up(p.z));
        ndn = V3(next_up(-p.x), next_up(-p.y), next_up(-p.z));
        tnear = t0;
        tfar = t1;
        h.t = t1;
        h.leaf = -1;
        h.g = 0x7FFFFFFF;
        h.u = h.v = 0.0f;
        h.nodes = h.tris = 0;
        node = 0;
        sp = 0;
        pc = 0;
    }

    // Processes one node; returns true when the walk is over (h holds the result).
    __device__ __forceinline__ bool step(const Bvh4Node* __restrict__ nodes,
                                         const float4* __restrict__ tri,
                                         int32_t* __restrict__ stack, volatile int32_t* spill) {
        const int lane = threadIdx.x;
        const float kInf = __int_as_float(0x7F800000);
        if (STATS) ++h.nodes;
        const char* nb = reinterpret_cast<const char*>(nodes);
        const uint32_t off = (uint32_t)node * (uint32_t)sizeof(Bvh4Node);
        const int4 ch = *reinterpret_cast<const int4*>(nb + (size_t)off + 96);
        int c[4] = {ch.x, ch.y, ch.z, ch.w};
        const float tcur = ANY ? tfar : h.t;
        float t0[4];
        bool box[4];
        boxes4(nb, off, inv, nup, ndn, tcur, t0, box);
        float dist[4];
        bool leaf_hit[4];
        bool any_leaf = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            leaf_hit[k] = box[k] && c[k] < 0;
            any_leaf = any_leaf || leaf_hit[k];
            dist[k] = (box[k] && c[k] >= 0) ? t0[k] : kInf;
        }
        if (any_leaf) {
            // every candidate triangle of this node in one round trip (non-candidates re-read
            // the first candidate and are ignored)
            const int first = leaf_hit[0] ? ~c[0] : (leaf_hit[1] ? ~c[1] : (leaf_hit[2] ? ~c[2] : ~c[3]));
            float4 A[4], B[4], C[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const size_t li = 3 * (size_t)(leaf_hit[k] ? ~c[k] : first);
                A[k] = tri[li];
                B[k] = tri[li + 1];
                C[k] = tri[li + 2];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float t, u, v;
                const bool ok = mt_all(o, d, xyz(A[k]), xyz(B[k]), xyz(C[k]), tnear, tfar, t, u, v);
                if (STATS && leaf_hit[k]) ++h.tris;
                if (!(leaf_hit[k] && ok)) continue;
                const int32_t g = __float_as_int(A[k].w);
                if (ANY) {
                    h.t = t;
                    h.leaf = ~c[k];
                    h.g = g;
                    return true;
                }
                if (h.leaf < 0 || t < h.t || (t == h.t && g < h.g)) {
                    h.t = t;
                    h.leaf = ~c[k];
                    h.g = g;
                    h.u = u;
                    h.v = v;
                }
            }
        }
        // order the inner children nearest first (misses and leaves sort last as +inf)
        auto cswap = [&](int x, int y) {
            if (dist[y] < dist[x]) {
                const float td = dist[x];
                dist[x] = dist[y];
                dist[y] = td;
                const int tc = c[x];
                c[x] = c[y];
                c[y] = tc;
            }
        };
        if (!ANY) {
            cswap(0, 1);
            cswap(2, 3);
            cswap(0, 2);
            cswap(1, 3);
            cswap(1, 2);
        } else {
            cswap(0, 1);
            cswap(0, 2);
            cswap(0, 3);
        }
#pragma unroll
        for (int k = 3; k >= 1; --k)
            if (dist[k] < kInf) {
                if (X == 0 || sp < L) stack[sp * kTraceBlock + lane] = c[k];
                else spill[sp - L] = c[k];
                ++sp;
            }
        if (dist[0] < kInf) {
            node = c[0];
            return false;
        }
        if (sp == 0) return true;
        --sp;
        node = (X == 0 || sp < L) ? stack[sp * kTraceBlock + lane] : spill[sp - L];
        return false;
    }

    // Wave-cooperative step over a four-wide node, called by all 64 lanes of the wave together
    // (active = this lane has a walk in progress). The box tests are per lane as in step(); the
    // hit inner children are ordered and pushed right away and the nearest becomes the lane's
    // next node, and the candidate triangles (leaf children whose box the ray enters) go to the
    // lane's stash; a lane whose stash could not take another node's four leaves pauses. The
    // stashed candidates of the whole wave are tested together, one (ray, triangle) pair per
    // lane and 64 pairs per pass, in a round that runs only when it is due: a full pass is
    // pending, no lane is walking, or the lanes not walking (paused, waiting for the results of
    // a finished walk, or idle) reach due_lanes (the refill threshold, so that finished lanes
    // can be refilled). Between rounds a lane keeps walking with the closest hit it had: the
    // boxes it enters are a superset of those it would enter with the tested candidates
    // applied, so the candidates tested are too, and the closest hit, the (t, global triangle
    // index) minimum over them, is the same as step()'s. A lane with no candidate in a round
    // helps test the others' (the owner's ray comes by lane shuffle). xchg: this wave's
    // kXchgPerWave LDS int4 (64 job and result slots, then the stash, kStash entries per lane
    // in lane columns). An any-hit walk ends at the round that accepts one of its candidates.
    __device__ __forceinline__ bool step_wave(bool active, const Bvh4Node* __restrict__ nodes,
                                              const float4* __restrict__ tri,
                                              int32_t* __restrict__ stack, volatile int32_t* spill,
                                              int4* __restrict__ xchg, int due_lanes) {
        const int lane = threadIdx.x;
        const int wl = (int)__lane_id();
        const float kInf = __int_as_float(0x7F800000);
        int32_t* stash = reinterpret_cast<int32_t*>(xchg + 64);
        if (active && node >= 0 && pc <= kStash - 4) {
            if (STATS) ++h.nodes;
            const char* nb = reinterpret_cast<const char*>(nodes);
            const uint32_t off = (uint32_t)node * (uint32_t)sizeof(Bvh4Node);
            const int4 ch = *reinterpret_cast<const int4*>(nb + (size_t)off + 96);
            int c[4] = {ch.x, ch.y, ch.z, ch.w};
            const float tcur = ANY ? tfar : h.t;
            float t0[4];
            bool box[4];
            boxes4(nb, off, inv, nup, ndn, tcur, t0, box);
            float dist[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (box[k] && c[k] < 0) stash[(pc++) * 64 + wl] = ~c[k];
                dist[k] = (box[k] && c[k] >= 0) ? t0[k] : kInf;
            }
            // order the inner children nearest first (misses and leaves sort last as +inf)
            auto cswap = [&](int x, int y) {
                if (dist[y] < dist[x]) {
                    const float td = dist[x];
                    dist[x] = dist[y];
                    dist[y] = td;
                    const int tc = c[x];
                    c[x] = c[y];
                    c[y] = tc;
                }
            };
            if (!ANY) {
                cswap(0, 1);
                cswap(2, 3);
                cswap(0, 2);
                cswap(1, 3);
                cswap(1, 2);
            } else {
                cswap(0, 1);
                cswap(0, 2);
                cswap(0, 3);
            }
#pragma unroll
            for (int k = 3; k >= 1; --k)
                if (dist[k] < kInf) {
                    if (X == 0 || sp < L) stack[sp * kTraceBlock + lane] = c[k];
                    else spill[sp - L] = c[k];
                    ++sp;
                }
            if (dist[0] < kInf) {
                node = c[0];
            } else if (sp > 0) {
                --sp;
                node = (X == 0 || sp < L) ? stack[sp * kTraceBlock + lane] : spill[sp - L];
            } else {
                node = -1;   // walk over once the stashed candidates are tested
            }
        }
        // pending candidates: this lane's go to slots [off, off + pc) of T in the wave
        const uint64_t lt = lanemask_lt();
        const uint64_t b0 = __ballot(pc & 1), b1 = __ballot(pc & 2), b2 = __ballot(pc & 4);
        const int T = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
        const bool walking = active && node >= 0 && pc <= kStash - 4;
        const int free_lanes = __popcll(__ballot(!walking));
        const bool due = T > 0 && (T >= 64 || __ballot(walking) == 0 || free_lanes >= due_lanes);
        if (due) {
            const int off = __popcll(b0 & lt) + 2 * __popcll(b1 & lt) + 4 * __popcll(b2 & lt);
            for (int base = 0; base < T; base += 64) {
                // stash entry k of this lane goes to slot off + k
                {
                    const int s0 = off - base;
#pragma unroll
                    for (int k = 0; k < kStash; ++k) {
                        const int sl = s0 + k;
                        if (k < pc && sl >= 0 && sl < 64)
                            xchg[sl] = make_int4(wl, stash[k * 64 + wl], 0, 0);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                const bool work = base + wl < T;
                const int4 job = work ? xchg[wl] : make_int4(wl, 0, 0, 0);
                // the owner's ray
                const V3 ro(__shfl(o.x, job.x, 64), __shfl(o.y, job.x, 64), __shfl(o.z, job.x, 64));
                const V3 rd(__shfl(d.x, job.x, 64), __shfl(d.y, job.x, 64), __shfl(d.z, job.x, 64));
                const float r0 = __shfl(tnear, job.x, 64), r1 = __shfl(tfar, job.x, 64);
                int4 res = make_int4(-1, 0, 0, 0);
                if (work) {
                    const size_t li = 3 * (size_t)job.y;
                    const float4 A = tri[li], B = tri[li + 1], C = tri[li + 2];
                    float t, u, v;
                    if (mt_all(ro, rd, xyz(A), xyz(B), xyz(C), r0, r1, t, u, v))
                        res = make_int4(job.y, __float_as_int(t), __float_as_int(u),
                                        __float_as_int(v));
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (work) xchg[wl] = res;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                // results in slot order (the closest-hit rule is a total order, so the order
                // only decides which accepted triangle an any-hit walk reports)
                const int s0 = off - base;
#pragma unroll
                for (int k = 0; k < kStash; ++k) {
                    const int slot = s0 + k;
                    if (!(k < pc && slot >= 0 && slot < 64)) continue;
                    if (STATS) ++h.tris;
                    const int4 r = xchg[slot];
                    if (r.x < 0) continue;
                    const float t = __int_as_float(r.y);
                    if (ANY) {
                        if (h.leaf < 0) {
                            h.t = t;
                            h.leaf = r.x;
                        }
                        continue;
                    }
                    bool take = h.leaf < 0 || t < h.t;
                    if (!take && t == h.t) {
                        // equal distances: the smaller global triangle index wins
                        take = __float_as_int(tri[3 * (size_t)r.x].w) <
                               __float_as_int(tri[3 * (size_t)h.leaf].w);
                    }
                    if (take) {
                        h.t = t;
                        h.leaf = r.x;
                        h.u = __int_as_float(r.z);
                        h.v = __int_as_float(r.w);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
            pc = 0;
        }
        if (!active) return false;
        if (ANY && h.leaf >= 0) return true;
        return node < 0 && pc == 0;
    }

    // Eight-wide nodes keep the per-lane step (their leaves are tested in groups of four).
    __device__ __forceinline__ bool step_wave(bool active, const Bvh8Node* __restrict__ nodes,
                                              const float4* __restrict__ tri,
                                              int32_t* __restrict__ stack, volatile int32_t* spill,
                                              int4* __restrict__, int) {
        return active ? step(nodes, tri, stack, spill) : false;
    }

    // Tests up to four leaf children of an eight-wide node (child words cw, hit flags lh) in one
    // round trip. Returns true when an ANY walk accepted a triangle.
    __device__ __forceinline__ bool leaves4(const float4* __restrict__ tri, int4 cw, bool l0,
                                            bool l1, bool l2, bool l3) {
        if (!(l0 || l1 || l2 || l3)) return false;
        const int cc[4] = {cw.x, cw.y, cw.z, cw.w};
        const bool lh[4] = {l0, l1, l2, l3};
        const int first = l0 ? ~cw.x : (l1 ? ~cw.y : (l2 ? ~cw.z : ~cw.w));
        float4 A[4], B[4], C[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const size_t li = 3 * (size_t)(lh[k] ? ~cc[k] : first);
            A[k] = tri[li];
            B[k] = tri[li + 1];
            C[k] = tri[li + 2];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float t, u, v;
            const bool ok = mt_all(o, d, xyz(A[k]), xyz(B[k]), xyz(C[k]), tnear, tfar, t, u, v);
            if (STATS && lh[k]) ++h.tris;
            if (!(lh[k] && ok)) continue;
            const int32_t g = __float_as_int(A[k].w);
            if (ANY) {
                h.t = t;
                h.leaf = ~cc[k];
                h.g = g;
                return true;
            }
            if (h.leaf < 0 || t < h.t || (t == h.t && g < h.g)) {
                h.t = t;
                h.leaf = ~cc[k];
                h.g = g;
                h.u = u;
                h.v = v;
            }
        }
        return false;
    }

    // Eight-wide quantised node: child boxes are decoded exactly as the build encoded them
    // (origin + byte * power-of-two scale, so the product is exact and the decoded box contains
    // the child), then walked like the four-wide node. The build stores leaf children first, so
    // nodes with at most four leaves test their triangles in one round trip.
    __device__ __forceinline__ bool step(const Bvh8Node* __restrict__ nodes,
                                         const float4* __restrict__ tri,
                                         int32_t* __restrict__ stack, volatile int32_t* spill) {
        const int lane = threadIdx.x;
        const float kInf = __int_as_float(0x7F800000);
        if (STATS) ++h.nodes;
        const uint4* np = reinterpret_cast<const uint4*>(nodes + node);
        const uint4 org = np[0], cl = np[1], chh = np[2];
        const uint4 qa = np[3], qb = np[4], qc = np[5];
        const float ox = __uint_as_float(org.x), oy = __uint_as_float(org.y);
        const float oz = __uint_as_float(org.z);
        const float sx = __uint_as_float((org.w & 0xFFu) << 23);
        const float sy = __uint_as_float(((org.w >> 8) & 0xFFu) << 23);
        const float sz = __uint_as_float(((org.w >> 16) & 0xFFu) << 23);
        int c[8] = {(int)cl.x, (int)cl.y, (int)cl.z, (int)cl.w,
                    (int)chh.x, (int)chh.y, (int)chh.z, (int)chh.w};
        const float tcur = ANY ? tfar : h.t;
        float dist[8];
        bool leaf_hit[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const bool hi4 = k >= 4;
            const uint32_t b = 8u * (uint32_t)(k & 3);
            const float lx = ox + (float)(((hi4 ? qa.y : qa.x) >> b) & 0xFFu) * sx;
            const float ly = oy + (float)(((hi4 ? qa.w : qa.z) >> b) & 0xFFu) * sy;
            const float lz = oz + (float)(((hi4 ? qb.y : qb.x) >> b) & 0xFFu) * sz;
            const float hx = ox + (float)(((hi4 ? qb.w : qb.z) >> b) & 0xFFu) * sx;
            const float hy = oy + (float)(((hi4 ? qc.y : qc.x) >> b) & 0xFFu) * sy;
            const float hz = oz + (float)(((hi4 ? qc.w : qc.z) >> b) & 0xFFu) * sz;
            float t0;
            const bool box = slab(lx, ly, lz, hx, hy, hz, inv, nup, ndn, tcur, t0) &&
                              c[k] != kEmptyChild;
            leaf_hit[k] = box && c[k] < 0;
            dist[k] = (box && c[k] >= 0) ? t0 : kInf;
        }
        if (leaves4(tri, make_int4(c[0], c[1], c[2], c[3]), leaf_hit[0], leaf_hit[1], leaf_hit[2],
                    leaf_hit[3]))
            return true;
        if (leaves4(tri, make_int4(c[4], c[5], c[6], c[7]), leaf_hit[4], leaf_hit[5], leaf_hit[6],
                    leaf_hit[7]))
            return true;
        auto cswap = [&](int x, int y) {
            if (dist[y] < dist[x]) {
                const float td = dist[x];
                dist[x] = dist[y];
                dist[y] = td;
                const int tc = c[x];
                c[x] = c[y];
                c[y] = tc;
            }
        };
        if (!ANY) {
            // 19-comparator sorting network for eight keys
            cswap(0, 2); cswap(1, 3); cswap(4, 6); cswap(5, 7);
            cswap(0, 4); cswap(1, 5); cswap(2, 6); cswap(3, 7);
            cswap(0, 1); cswap(2, 3); cswap(4, 5); cswap(6, 7);
            cswap(2, 4); cswap(3, 5);
            cswap(1, 4); cswap(3, 6);
            cswap(1, 2); cswap(3, 4); cswap(5, 6);
        } else {
#pragma unroll
            for (int k = 1; k < 8; ++k) cswap(0, k);
        }
#pragma unroll
        for (int k = 7; k >= 1; --k)
            if (dist[k] < kInf) {
                if (X == 0 || sp < L) stack[sp * kTraceBlock + lane] = c[k];
                else spill[sp - L] = c[k];
                ++sp;
            }
        if (dist[0] < kInf) {
            node = c[0];
            return false;
        }
        if (sp == 0) return true;
        --sp;
        node = (X == 0 || sp < L) ? stack[sp * kTraceBlock + lane] : spill[sp - L];
        return false;
    }
};

// Whole walk of one ray (callers without dynamic ray refill).
template <bool ANY, int STACK, bool STATS = false, class NODE>
__device__ __forceinline__ bool traverse(const NODE* __restrict__ nodes,
                                         const float4* __restrict__ tri, int n_tris,
                                         const V3& o, const V3& d, float tnear, float tfar,
                                         int32_t* __restrict__ stack, TraceHit& h) {
    Walker<ANY, STACK, STATS> w;
    w.start(o, d, tnear, tfar);
    if (n_tris > 0) {
        volatile int32_t spill[Walker<ANY, STACK, STATS>::X > 0 ? Walker<ANY, STACK, STATS>::X : 1];
        while (!w.step(nodes, tri, stack, spill)) {
        }
    }
    h = w.h;
    return h.leaf >= 0;
}

// Work distribution for the persistent traversal kernels. Device-scope atomics on one address
// serialise (about 11 ns each on MI355X, so a single claim counter caps a kernel at ~88M claims
// per second), so the item range [0, n) is split into kWorkSegments contiguous segments, each
// with its own claim counter on its own 128-byte line. A wave starts in segment (global wave
// index mod kWorkSegments), claims chunks there until the segment is exhausted, then moves to the
// next segment; once it has found kWorkSegments segments in a row exhausted (counters only grow,
// so then every segment is) it is done. Chunks are contiguous, so sorted rays stay coherent.
constexpr int kWorkSegments = 64;
constexpr int kWorkStride = 32;   // int32 per counter slot
constexpr int kWorkInts = kWorkSegments * kWorkStride;

struct WorkQueue {
    int32_t* ctr;   // kWorkInts, zero before the launch
    int n, seg_size, seg, dry;

    __device__ __forceinline__ void init(int32_t* counters, int count) {
        ctr = counters;
        n = count;
        seg_size = (count + kWorkSegments - 1) / kWorkSegments;
        seg = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % kWorkSegments;
        dry = 0;
    }
    __device__ __forceinline__ bool exhausted() const { return dry >= kWorkSegments; }
};

// Lanes whose item finished take the next unclaimed items, one atomic per wave, once at least
// refill_lanes lanes are idle (or every lane is). Returns the new item index or -1.
__device__ __forceinline__ int refill(WorkQueue& q, bool idle, int refill_lanes) {
    const uint64_t want = __ballot(idle && !q.exhausted());
    const int k = __popcll(want);
    if (k == 0 || (k < refill_lanes && __ballot(!idle) != 0)) return -1;
    const int leader = __ffsll((unsigned long long)want) - 1;
    int first = -1, end = 0;
    while (!q.exhausted()) {
        const int s0 = q.seg * q.seg_size;
        const int s1 = min(q.n, s0 + q.seg_size);
        if (s0 < s1) {
            int base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(&q.ctr[q.seg * kWorkStride], k);
            base = __shfl(base, leader, 64);
            if (s0 + base < s1) {
                first = s0 + base;
                end = s1;
                q.dry = 0;
                break;
            }
        }
        ++q.dry;
        q.seg = q.seg + 1 == kWorkSegments ? 0 : q.seg + 1;
    }
    if (!idle || first < 0) return -1;
    const int r = first + __popcll(want & lanemask_lt());
    return r < end ? r : -1;
}

// Closest hits of rays [0, n) by the calling persistent block (the body of the closest-hit
// kernels): each lane walks one ray at a time and takes the next from the work queue when done.
// stack / xchg: the block's LDS stack columns and candidate-exchange slots; spill: this lane's
// private stack overflow.
template <class NODE, int STACK, class R>
__device__ __forceinline__ void closest_walks(const NODE* __restrict__ nodes,
                                              const float4* __restrict__ tri, int n_tris,
                                              const R& rays, int n, float tnear, float tfar,
                                              int32_t* __restrict__ out_leaf,
                                              float* __restrict__ out_t, float* __restrict__ out_u,
                                              float* __restrict__ out_v,
                                              uint32_t* __restrict__ out_flag,
                                              int32_t* __restrict__ counter, int refill_lanes,
                                              int32_t* __restrict__ stack, int4* __restrict__ xchg,
                                              volatile int32_t* spill) {
    using W = Walker<false, STACK>;
    int4* wx = xchg + (threadIdx.x >> 6) * kXchgPerWave;
    W w;
    int ray = -1;
    WorkQueue wq;
    wq.init(counter, n);
    for (;;) {
        const int r = refill(wq, ray < 0, refill_lanes);
        if (r >= 0) {
            ray = r;
            V3 o, d;
            ld_od(rays, r, o, d);
            w.start(o, d, tnear, tfar);
        }
        if (__ballot(ray >= 0) == 0) {
            if (wq.exhausted()) break;
            continue;
        }
        // every lane of the wave takes part in the step (idle lanes test others' triangles)
        const bool done = w.step_wave(ray >= 0 && n_tris > 0, nodes, tri, stack, spill, wx,
                                      refill_lanes);
        if (ray >= 0 && (n_tris <= 0 || done)) {
            out_leaf[ray] = w.h.leaf;
            out_t[ray] = w.h.t;
            out_u[ray] = w.h.u;
            out_v[ray] = w.h.v;
            if (out_flag) out_flag[ray] = w.h.leaf >= 0 ? 1u : 0u;
            ray = -1;
        }
    }
}

// Persistent-grid size for traversal kernels: blocks_per_cu resident blocks on every CU
// (0 = the context's trace_blocks_per_cu).
inline unsigned persistent_grid(int n, const wf_context* ctx, int blocks_per_cu = 0) {
    const unsigned need = grid_for(n, kTraceBlock);
    const int bpc = blocks_per_cu > 0 ? blocks_per_cu : ctx->tuning.trace_blocks_per_cu;
    const unsigned fill = (unsigned)ctx->cu_count * (unsigned)bpc;
    return need < fill ? need : fill;
}

}  // namespace wf
