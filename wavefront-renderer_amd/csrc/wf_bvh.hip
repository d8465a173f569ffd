// Higher-quality accelerator build: PLOC clustering of the Morton-ordered leaves (Meister and
// Bittner, "Parallel Locally-Ordered Clustering for Bounding Volume Hierarchy Construction",
// IEEE TVCG 2018) and a surface-area-optimal collapse of the binary tree into wide nodes.
//
// PLOC, per iteration over the current clusters (in Morton order):
//   1. every cluster finds, within `radius` positions, the cluster whose union box has the
//      smallest surface area (ties: nearer position, then pairs starting at an even position,
//      then the lower position, so that equal boxes still pair up two by two);
//   2. mutual nearest neighbours merge into a new internal node kept at the lower position;
//   3. the surviving clusters are compacted in order.
// Internal nodes are numbered downwards from n - 2, so the last merge (the root) is node 0: the
// output arrays have the layout of the Karras build (left/right child: node index, or ~leaf).
//
// The collapse keeps a subset of the binary nodes as wide nodes (four or eight children). With one
// triangle per leaf every leaf sits in exactly one wide node whatever the choice, so the surface
// area heuristic reduces to the sum of the kept nodes' box areas (the expected number of node
// visits of a random ray). A bottom-up dynamic programme over the binary tree finds the subset of
// least total area (the slot-distribution recurrence of Ylitie, Karras and Laine, "Efficient
// Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs", HPG 2017, with the leaf term
// constant); the top-down pass then writes the wide nodes level by level, numbered in order after
// the previous level, carrying the exact traversal stack bound down the levels. Eight-wide nodes
// store their child boxes quantised outward to 8 bits on a per-node power-of-two grid.
#include <algorithm>

#include "wf_device.h"
#include "wf_internal.h"

namespace wf {

constexpr int kPlocBlock = 256;
constexpr int kPlocMaxRadius = 32;

__device__ __forceinline__ float half_area(const float4& l, const float4& h) {
    const float dx = h.x - l.x, dy = h.y - l.y, dz = h.z - l.z;
    return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ float union_area(const float4& la, const float4& ha, const float4& lb,
                                            const float4& hb) {
    const float4 l = make_float4(fminf(la.x, lb.x), fminf(la.y, lb.y), fminf(la.z, lb.z), 0.0f);
    const float4 h = make_float4(fmaxf(ha.x, hb.x), fmaxf(ha.y, hb.y), fmaxf(ha.z, hb.z), 0.0f);
    return half_area(l, h);
}

__global__ void k_ploc_init(int n, const float4* __restrict__ slo, const float4* __restrict__ shi,
                            int32_t* __restrict__ cref, float4* __restrict__ clo,
                            float4* __restrict__ chi, int32_t* __restrict__ lparent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cref[i] = ~i;
    clo[i] = slo[i];
    chi[i] = shi[i];
    lparent[i] = -1;
}

// Nearest neighbour of every cluster within the window (boxes of the block's range plus the
// halo staged in LDS).
__global__ __launch_bounds__(kPlocBlock) void k_ploc_nn(int m, int radius,
                                                        const float4* __restrict__ clo,
                                                        const float4* __restrict__ chi,
                                                        int32_t* __restrict__ nn) {
    __shared__ float4 tl[kPlocBlock + 2 * kPlocMaxRadius];
    __shared__ float4 th[kPlocBlock + 2 * kPlocMaxRadius];
    const int b0 = blockIdx.x * kPlocBlock - radius;
    const int span = kPlocBlock + 2 * radius;
    for (int k = threadIdx.x; k < span; k += kPlocBlock) {
        const int j = b0 + k;
        if (j >= 0 && j < m) {
            tl[k] = clo[j];
            th[k] = chi[j];
        }
    }
    __syncthreads();
    const int i = blockIdx.x * kPlocBlock + threadIdx.x;
    if (i >= m) return;
    const int li = i - b0;
    const float4 la = tl[li], ha = th[li];
    float best = __int_as_float(0x7F800000);
    int bj = -1, bdist = 0x7FFFFFFF, bpar = 2, bmin = 0x7FFFFFFF;
    const int lo = i - radius < 0 ? 0 : i - radius;
    const int hi = i + radius > m - 1 ? m - 1 : i + radius;
    for (int j = lo; j <= hi; ++j) {
        if (j == i) continue;
        const float d = union_area(la, ha, tl[j - b0], th[j - b0]);
        const int dist = j > i ? j - i : i - j;
        const int mn = j < i ? j : i;
        const int par = mn & 1;
        const bool better = d < best ||
                            (d == best && (dist < bdist ||
                                           (dist == bdist && (par < bpar ||
                                                              (par == bpar && mn < bmin)))));
        if (better) {
            best = d;
            bj = j;
            bdist = dist;
            bpar = par;
            bmin = mn;
        }
    }
    nn[i] = bj;
}

// Merge flags (lower position of a mutual pair) and survival flags.
__global__ void k_ploc_flags(int m, const int32_t* __restrict__ nn, uint32_t* __restrict__ merge,
                             uint32_t* __restrict__ keep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = j >= 0 && nn[j] == i;
    merge[i] = (mutual && i < j) ? 1u : 0u;
    keep[i] = (mutual && i > j) ? 0u : 1u;
}

// Creates the merged nodes and writes the surviving clusters compacted into the other buffer.
__global__ void k_ploc_write(int m, int n, const int32_t* __restrict__ created,
                             const int32_t* __restrict__ nn, const uint32_t* __restrict__ merge,
                             const uint32_t* __restrict__ mpos, const uint32_t* __restrict__ keep,
                             const uint32_t* __restrict__ kpos, const int32_t* __restrict__ cref,
                             const float4* __restrict__ clo, const float4* __restrict__ chi,
                             int32_t* __restrict__ oref, float4* __restrict__ olo,
                             float4* __restrict__ ohi, int32_t* __restrict__ left,
                             int32_t* __restrict__ right, int32_t* __restrict__ iparent,
                             int32_t* __restrict__ lparent, float4* __restrict__ nlo,
                             float4* __restrict__ nhi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || !keep[i]) return;
    const int k = (int)kpos[i];
    if (!merge[i]) {
        oref[k] = cref[i];
        olo[k] = clo[i];
        ohi[k] = chi[i];
        return;
    }
    const int j = nn[i];
    const int node = n - 2 - (*created + (int)mpos[i]);
    const int a = cref[i], b = cref[j];
    const float4 la = clo[i], ha = chi[i], lb = clo[j], hb = chi[j];
    const float4 L = make_float4(fminf(la.x, lb.x), fminf(la.y, lb.y), fminf(la.z, lb.z), 0.0f);
    const float4 H = make_float4(fmaxf(ha.x, hb.x), fmaxf(ha.y, hb.y), fmaxf(ha.z, hb.z), 0.0f);
    left[node] = a;
    right[node] = b;
    if (a < 0) lparent[~a] = node; else iparent[a] = node;
    if (b < 0) lparent[~b] = node; else iparent[b] = node;
    nlo[node] = L;
    nhi[node] = H;
    oref[k] = node;
    olo[k] = L;
    ohi[k] = H;
}

__global__ void k_ploc_commit(int32_t* created, const int32_t* merges) { *created += *merges; }

// ---- surface-area-optimal collapse ------------------------------------------------------------
// For a binary node m placed in j child slots of the wide node above it (1 <= j <= W):
//   G(m, j) = min over 1 <= k < j of F(left, k) + F(right, j - k)    (m opened: its two subtrees
//             share the j slots),
//   F(m, 1) = A(m) + G(m, W)                                          (m kept as a wide node),
//   F(m, j) = min(F(m, 1), G(m, j)) for j >= 2,
// and F(leaf, j) = 0 (a leaf takes one slot wherever it goes). F(m, j) is non-increasing in j.
// The choice word of m holds, for j = 2..W, the split k of G(m, j) (kSplitBits bits each) and,
// above them, whether m is kept (F(m, 1) <= G(m, j)).
template <int W>
struct CollapseChoice {
    static constexpr int kSplitBits = W == 4 ? 2 : 3;
    static constexpr int kKeepBase = kSplitBits * (W - 1);
    __device__ static int split(uint32_t w, int j) {
        return (int)((w >> (kSplitBits * (j - 2))) & ((1u << kSplitBits) - 1u));
    }
    __device__ static bool keep(uint32_t w, int j) { return (w >> (kKeepBase + j - 2)) & 1u; }
};

__device__ __forceinline__ float coherent_load_f(const float* p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ void coherent_store_f(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Bottom-up: every leaf climbs; the second arrival at a node has both subtrees' F and computes
// the node's (F(m, 1..W) at F[m * W + j - 1], its choice word in choice[m]).
template <int W>
__global__ void k_collapse_cost(int n, const int32_t* __restrict__ left,
                                const int32_t* __restrict__ right,
                                const int32_t* __restrict__ iparent,
                                const int32_t* __restrict__ lparent,
                                const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                float* F, uint32_t* __restrict__ choice, uint32_t* visits) {
    using CC = CollapseChoice<W>;
    const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= n) return;
    int p = lparent[leaf];
    while (p >= 0) {
        __threadfence();
        if (atomicAdd(&visits[p], 1u) == 0u) return;   // the sibling subtree is not done yet
        __threadfence();
        const int cl = left[p], cr = right[p];
        float fl[W], fr[W];
#pragma unroll
        for (int j = 0; j < W; ++j) {
            fl[j] = cl < 0 ? 0.0f : coherent_load_f(F + (size_t)cl * W + j);
            fr[j] = cr < 0 ? 0.0f : coherent_load_f(F + (size_t)cr * W + j);
        }
        float g[W + 1];
        uint32_t word = 0;
#pragma unroll
        for (int j = 2; j <= W; ++j) {
            float best = fl[0] + fr[j - 2];
            int bk = 1;
#pragma unroll
            for (int k = 2; k < j; ++k) {
                const float c = fl[k - 1] + fr[j - k - 1];
                if (c < best) {
                    best = c;
                    bk = k;
                }
            }
            g[j] = best;
            word |= (uint32_t)bk << (CC::kSplitBits * (j - 2));
        }
        const float f1 = half_area(nlo[p], nhi[p]) + g[W];
        coherent_store_f(F + (size_t)p * W, f1);
#pragma unroll
        for (int j = 2; j <= W; ++j) {
            const bool keep = f1 <= g[j];
            word |= (uint32_t)keep << (CC::kKeepBase + j - 2);
            coherent_store_f(F + (size_t)p * W + j - 1, keep ? f1 : g[j]);
        }
        choice[p] = word;
        p = iparent[p];
    }
}

// Chooses the (up to W) children of each queued binary node by expanding its W slots along the
// choice words; kint = inner children.
template <int W>
__global__ void k_collapse_pick(int count, const int32_t* __restrict__ qb,
                                const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                                const uint32_t* __restrict__ choice, int32_t* __restrict__ picked,
                                uint32_t* __restrict__ kint) {
    using CC = CollapseChoice<W>;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int b = qb[t];
    int c[W];
    // pending (subtree, slots); each expansion replaces one entry by two whose slots sum to its
    // own, and every entry holds at least one slot, so at most W are pending
    int sm[W], sj[W];
    int top = 0, m = 0;
    const int kb = CC::split(choice[b], W);
    sm[top] = right[b];
    sj[top++] = W - kb;
    sm[top] = left[b];
    sj[top++] = kb;
    while (top > 0) {
        --top;
        const int x = sm[top], j = sj[top];
        if (x < 0 || j == 1 || CC::keep(choice[x], j)) {
            c[m++] = x;
            continue;
        }
        const int k = CC::split(choice[x], j);
        sm[top] = right[x];
        sj[top++] = j - k;
        sm[top] = left[x];
        sj[top++] = k;
    }
    for (int q = m; q < W; ++q) c[q] = kEmptyChild;
    uint32_t k = 0;
    for (int q = 0; q < W; ++q) {
        k += (c[q] >= 0) ? 1u : 0u;
        picked[(size_t)t * W + q] = c[q];
    }
    kint[t] = k;
}

__device__ __forceinline__ float qscale(uint32_t ebits) { return __uint_as_float(ebits << 23); }
__device__ __forceinline__ float qdecode(float o, uint32_t q, float scale) {
    return o + (float)q * scale;
}

// Smallest power-of-two scale (biased exponent, >= 1) whose 255 steps from lo reach hi.
__device__ __forceinline__ uint32_t quant_exponent(float lo, float hi) {
    const float s = (hi - lo) / 255.0f;
    uint32_t e = (__float_as_uint(s) >> 23) & 0xFFu;
    if (s > qscale(e)) ++e;
    if (e < 1) e = 1;
    while (e < 254 && qdecode(lo, 255u, qscale(e)) < hi) ++e;
    return e;
}

// Outward 8-bit quantisation of [clo, chi] on the grid (o, scale).
__device__ __forceinline__ void quantise(float o, float scale, float clo, float chi, uint32_t& qlo,
                                         uint32_t& qhi) {
    float fl = floorf((clo - o) / scale), fh = ceilf((chi - o) / scale);
    int l = (int)fminf(fmaxf(fl, 0.0f), 255.0f), h = (int)fminf(fmaxf(fh, 0.0f), 255.0f);
    while (l > 0 && qdecode(o, (uint32_t)l, scale) > clo) --l;
    while (h < 255 && qdecode(o, (uint32_t)h, scale) < chi) ++h;
    qlo = (uint32_t)l;
    qhi = (uint32_t)h;
}

// Writes the wide nodes of a level and queues their inner children for the next level.
template <int W>
__global__ void k_collapse_write(int count, int next_base, const int32_t* __restrict__ qw,
                                 const int32_t* __restrict__ qd,
                                 const int32_t* __restrict__ picked,
                                 const uint32_t* __restrict__ kint,
                                 const uint32_t* __restrict__ koff,
                                 const float4* __restrict__ slo, const float4* __restrict__ shi,
                                 const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                 Bvh4Node* __restrict__ out4, Bvh8Node* __restrict__ out8,
                                 int32_t* __restrict__ nqb, int32_t* __restrict__ nqw,
                                 int32_t* __restrict__ nqd, int32_t* __restrict__ stack_bound) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int c[W];
    if (W == 8) {
        // leaf children first, then inner children, then empty slots (the eight-wide walker tests
        // the leaves of slots 0..3 and 4..7 in one round trip each)
        int m = 0;
        for (int pass = 0; pass < 3; ++pass)
            for (int q = 0; q < W; ++q) {
                const int v = picked[(size_t)t * W + q];
                const int cls = v == kEmptyChild ? 2 : (v < 0 ? 0 : 1);
                if (cls == pass) c[m++] = v;
            }
    } else {
        for (int q = 0; q < W; ++q) c[q] = picked[(size_t)t * W + q];
    }
    const int k = (int)kint[t];
    const int depth_in = qd[t];
    const int child_depth = depth_in + (k > 0 ? k - 1 : 0);
    atomicMax(stack_bound, child_depth);
    float4 lo[W], hi[W];
    int ch[W];
    int r = 0;
    for (int q = 0; q < W; ++q) {
        if (c[q] == kEmptyChild) {
            lo[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            hi[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            ch[q] = kEmptyChild;
        } else if (c[q] < 0) {
            lo[q] = slo[~c[q]];
            hi[q] = shi[~c[q]];
            ch[q] = c[q];
        } else {
            lo[q] = nlo[c[q]];
            hi[q] = nhi[c[q]];
            const int slot = (int)koff[t] + r++;
            ch[q] = next_base + slot;
            nqb[slot] = c[q];
            nqw[slot] = next_base + slot;
            nqd[slot] = child_depth;
        }
    }
    if (W == 4) {
        // an empty slot holds the inverted infinite box (lo = +inf, hi = -inf), which the
        // four-wide box test rejects for either direction sign without checking the child word
        const float inf = __int_as_float(0x7F800000);
        for (int q = 0; q < W; ++q)
            if (ch[q] == kEmptyChild) {
                lo[q] = make_float4(inf, inf, inf, 0.0f);
                hi[q] = make_float4(-inf, -inf, -inf, 0.0f);
            }
        Bvh4Node nd;
        nd.lox = make_float4(lo[0].x, lo[1].x, lo[2].x, lo[3].x);
        nd.loy = make_float4(lo[0].y, lo[1].y, lo[2].y, lo[3].y);
        nd.loz = make_float4(lo[0].z, lo[1].z, lo[2].z, lo[3].z);
        nd.hix = make_float4(hi[0].x, hi[1].x, hi[2].x, hi[3].x);
        nd.hiy = make_float4(hi[0].y, hi[1].y, hi[2].y, hi[3].y);
        nd.hiz = make_float4(hi[0].z, hi[1].z, hi[2].z, hi[3].z);
        nd.child = make_int4(ch[0], ch[1], ch[2], ch[3]);
        nd.pad = make_int4(0, 0, 0, 0);
        out4[qw[t]] = nd;
        return;
    }
    // node grid: union of the valid children
    float L[3] = {3.0e38f, 3.0e38f, 3.0e38f}, H[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int q = 0; q < W; ++q) {
        if (ch[q] == kEmptyChild) continue;
        L[0] = fminf(L[0], lo[q].x); L[1] = fminf(L[1], lo[q].y); L[2] = fminf(L[2], lo[q].z);
        H[0] = fmaxf(H[0], hi[q].x); H[1] = fmaxf(H[1], hi[q].y); H[2] = fmaxf(H[2], hi[q].z);
    }
    uint32_t e[3];
    for (int a = 0; a < 3; ++a) e[a] = quant_exponent(L[a], H[a]);
    uint32_t bytes[6][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};   // [bound][half]
    for (int q = 0; q < W; ++q) {
        if (ch[q] == kEmptyChild) continue;
        const float cl[3] = {lo[q].x, lo[q].y, lo[q].z}, chh[3] = {hi[q].x, hi[q].y, hi[q].z};
        for (int a = 0; a < 3; ++a) {
            uint32_t ql, qh;
            quantise(L[a], qscale(e[a]), cl[a], chh[a], ql, qh);
            bytes[a][q >> 2] |= ql << (8 * (q & 3));
            bytes[3 + a][q >> 2] |= qh << (8 * (q & 3));
        }
    }
    Bvh8Node nd;
    nd.origin = make_float4(L[0], L[1], L[2], __uint_as_float(e[0] | (e[1] << 8) | (e[2] << 16)));
    nd.child_lo = make_int4(ch[0], ch[1], ch[2], ch[3]);
    nd.child_hi = make_int4(ch[4 % W], ch[5 % W], ch[6 % W], ch[7 % W]);
    nd.qa = make_uint4(bytes[0][0], bytes[0][1], bytes[1][0], bytes[1][1]);
    nd.qb = make_uint4(bytes[2][0], bytes[2][1], bytes[3][0], bytes[3][1]);
    nd.qc = make_uint4(bytes[4][0], bytes[4][1], bytes[5][0], bytes[5][1]);
    nd.pad[0] = make_int4(0, 0, 0, 0);
    nd.pad[1] = make_int4(0, 0, 0, 0);
    out8[qw[t]] = nd;
}

__global__ void k_set_root(int32_t* qb, int32_t* qw, int32_t* qd, int32_t* iparent) {
    qb[0] = 0;
    qw[0] = 0;
    qd[0] = 0;
    iparent[0] = -1;
}

wf_status ploc_build(wf_context* ctx, int n, int radius, const float4* slo, const float4* shi,
                     int32_t* left, int32_t* right, int32_t* iparent, int32_t* lparent,
                     float4* nlo, float4* nhi) {
    radius = std::max(1, std::min(radius, kPlocMaxRadius));
    void *pr[2], *pl[2], *ph[2], *pnn, *pm, *pmp, *pk, *pkp, *pcnt;
    WF_TRY(workspace(ctx, "ploc_ref0", sizeof(int32_t) * n, &pr[0]));
    WF_TRY(workspace(ctx, "ploc_ref1", sizeof(int32_t) * n, &pr[1]));
    WF_TRY(workspace(ctx, "ploc_lo0", sizeof(float4) * n, &pl[0]));
    WF_TRY(workspace(ctx, "ploc_lo1", sizeof(float4) * n, &pl[1]));
    WF_TRY(workspace(ctx, "ploc_hi0", sizeof(float4) * n, &ph[0]));
    WF_TRY(workspace(ctx, "ploc_hi1", sizeof(float4) * n, &ph[1]));
    WF_TRY(workspace(ctx, "ploc_nn", sizeof(int32_t) * n, &pnn));
    WF_TRY(workspace(ctx, "ploc_merge", sizeof(uint32_t) * n, &pm));
    WF_TRY(workspace(ctx, "ploc_mpos", sizeof(uint32_t) * n, &pmp));
    WF_TRY(workspace(ctx, "ploc_keep", sizeof(uint32_t) * n, &pk));
    WF_TRY(workspace(ctx, "ploc_kpos", sizeof(uint32_t) * n, &pkp));
    WF_TRY(workspace(ctx, "ploc_counts", sizeof(int32_t) * 4, &pcnt));
    int32_t* cnt = (int32_t*)pcnt;   // [0] created, [1] merges, [2] survivors
    WF_TRY(wf_fill_u32(ctx, pcnt, 0u, 4));
    WF_TRY(wf_fill_u32(ctx, iparent, 0xFFFFFFFFu, (size_t)n));
    k_ploc_init<<<grid_for(n, 256), 256, 0, ctx->stream>>>(n, slo, shi, (int32_t*)pr[0],
                                                           (float4*)pl[0], (float4*)ph[0], lparent);
    WF_HIP(hipGetLastError());
    int m = n, cur = 0, iters = 0;
    // every iteration merges at least the globally closest pair; the pairing rule halves runs of
    // equal boxes, so the count is small in practice: the cap only guards against pathologies
    const int max_iters = 4096;
    while (m > 1) {
        if (++iters > max_iters)
            return set_error(WF_ERR_UNSUPPORTED, "PLOC did not converge in %d iterations", max_iters);
        const int nxt = cur ^ 1;
        k_ploc_nn<<<grid_for(m, kPlocBlock), kPlocBlock, 0, ctx->stream>>>(
            m, radius, (float4*)pl[cur], (float4*)ph[cur], (int32_t*)pnn);
        WF_HIP(hipGetLastError());
        k_ploc_flags<<<grid_for(m, 256), 256, 0, ctx->stream>>>(m, (int32_t*)pnn, (uint32_t*)pm,
                                                                (uint32_t*)pk);
        WF_HIP(hipGetLastError());
        WF_TRY(exclusive_scan_u32(ctx, (uint32_t*)pm, (uint32_t*)pmp, host_count(m), cnt + 1,
                                  "bvh_build"));
        WF_TRY(exclusive_scan_u32(ctx, (uint32_t*)pk, (uint32_t*)pkp, host_count(m), cnt + 2,
                                  "bvh_build"));
        k_ploc_write<<<grid_for(m, 256), 256, 0, ctx->stream>>>(
            m, n, cnt, (int32_t*)pnn, (uint32_t*)pm, (uint32_t*)pmp, (uint32_t*)pk,
            (uint32_t*)pkp, (int32_t*)pr[cur], (float4*)pl[cur], (float4*)ph[cur],
            (int32_t*)pr[nxt], (float4*)pl[nxt], (float4*)ph[nxt], left, right, iparent, lparent,
            nlo, nhi);
        WF_HIP(hipGetLastError());
        k_ploc_commit<<<1, 1, 0, ctx->stream>>>(cnt, cnt + 1);
        WF_HIP(hipGetLastError());
        int32_t survivors = 0;
        WF_TRY(readback(ctx, &survivors, cnt + 2, sizeof(int32_t)));
        if (survivors >= m) return set_error(WF_ERR_HIP, "PLOC made no progress");
        m = survivors;
        cur = nxt;
    }
    int32_t created = 0;
    WF_TRY(readback(ctx, &created, cnt, sizeof(int32_t)));
    if (created != n - 1) return set_error(WF_ERR_HIP, "PLOC created %d nodes for %d leaves", created, n);
    return WF_OK;
}

template <int W>
static wf_status collapse_levels(wf_context* ctx, int n, const int32_t* left, const int32_t* right,
                                 int32_t* iparent, const int32_t* lparent, const float4* slo,
                                 const float4* shi, const float4* nlo, const float4* nhi,
                                 Bvh4Node* out4, Bvh8Node* out8, int32_t* n_nodes,
                                 int32_t* stack_bound) {
    const int ni = n - 1;
    void *pf, *pch, *pvis;
    WF_TRY(workspace(ctx, "col_cost", sizeof(float) * W * (size_t)ni, &pf));
    WF_TRY(workspace(ctx, "col_choice", sizeof(uint32_t) * ni, &pch));
    WF_TRY(workspace(ctx, "col_visits", sizeof(uint32_t) * ni, &pvis));
    WF_TRY(wf_fill_u32(ctx, pvis, 0u, (size_t)ni));
    k_collapse_cost<W><<<grid_for(n, 256), 256, 0, ctx->stream>>>(
        n, left, right, iparent, lparent, nlo, nhi, (float*)pf, (uint32_t*)pch, (uint32_t*)pvis);
    WF_HIP(hipGetLastError());
    void *pqb[2], *pqw[2], *pqd[2], *ppick, *pk, *pkoff, *pcnt;
    for (int k = 0; k < 2; ++k) {
        char name[32];
        snprintf(name, sizeof(name), "col_qb%d", k);
        WF_TRY(workspace(ctx, name, sizeof(int32_t) * ni, &pqb[k]));
        snprintf(name, sizeof(name), "col_qw%d", k);
        WF_TRY(workspace(ctx, name, sizeof(int32_t) * ni, &pqw[k]));
        snprintf(name, sizeof(name), "col_qd%d", k);
        WF_TRY(workspace(ctx, name, sizeof(int32_t) * ni, &pqd[k]));
    }
    WF_TRY(workspace(ctx, "col_pick", sizeof(int32_t) * W * (size_t)ni, &ppick));
    WF_TRY(workspace(ctx, "col_kint", sizeof(uint32_t) * ni, &pk));
    WF_TRY(workspace(ctx, "col_koff", sizeof(uint32_t) * ni, &pkoff));
    WF_TRY(workspace(ctx, "col_counts", sizeof(int32_t) * 4, &pcnt));
    int32_t* cnt = (int32_t*)pcnt;   // [0] next level size, [1] stack bound
    WF_TRY(wf_fill_u32(ctx, pcnt, 0u, 4));
    k_set_root<<<1, 1, 0, ctx->stream>>>((int32_t*)pqb[0], (int32_t*)pqw[0], (int32_t*)pqd[0],
                                         iparent);
    WF_HIP(hipGetLastError());
    int count = 1, total = 1, cur = 0;
    while (count > 0) {
        const int nxt = cur ^ 1;
        k_collapse_pick<W><<<grid_for(count, 256), 256, 0, ctx->stream>>>(
            count, (int32_t*)pqb[cur], left, right, (const uint32_t*)pch, (int32_t*)ppick,
            (uint32_t*)pk);
        WF_HIP(hipGetLastError());
        WF_TRY(exclusive_scan_u32(ctx, (uint32_t*)pk, (uint32_t*)pkoff, host_count(count), cnt,
                                  "bvh_build"));
        k_collapse_write<W><<<grid_for(count, 256), 256, 0, ctx->stream>>>(
            count, total, (int32_t*)pqw[cur], (int32_t*)pqd[cur], (int32_t*)ppick, (uint32_t*)pk,
            (uint32_t*)pkoff, slo, shi, nlo, nhi, out4, out8, (int32_t*)pqb[nxt],
            (int32_t*)pqw[nxt], (int32_t*)pqd[nxt], cnt + 1);
        WF_HIP(hipGetLastError());
        int32_t next = 0;
        WF_TRY(readback(ctx, &next, cnt, sizeof(int32_t)));
        total += next;
        if (total > ni) return set_error(WF_ERR_HIP, "wide collapse overflow");
        count = next;
        cur = nxt;
    }
    int32_t bound = 0;
    WF_TRY(readback(ctx, &bound, cnt + 1, sizeof(int32_t)));
    *n_nodes = total;
    *stack_bound = bound;
    return WF_OK;
}

wf_status collapse_wide(wf_context* ctx, int n, const int32_t* left, const int32_t* right,
                        int32_t* iparent, const int32_t* lparent, const float4* slo,
                        const float4* shi, const float4* nlo, const float4* nhi, Bvh4Node* out4,
                        Bvh8Node* out8, int32_t* n_nodes, int32_t* stack_bound) {
    if (out8)
        return collapse_levels<8>(ctx, n, left, right, iparent, lparent, slo, shi, nlo, nhi,
                                  nullptr, out8, n_nodes, stack_bound);
    return collapse_levels<4>(ctx, n, left, right, iparent, lparent, slo, shi, nlo, nhi, out4,
                              nullptr, n_nodes, stack_bound);
}

}  // namespace wf
