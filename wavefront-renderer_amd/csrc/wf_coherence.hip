// Coherence metrics of shading batches, the statistics renderScene keeps beside its timers
// (src/render/WavefrontRenderer.cpp:67-92 computeAverageRunLength, 187-246 the weighted batch
// sums and cacheLineHomogeneity):
//   run length  = (entries with a texture) / (runs of equal IDs among those entries, in shading
//                 order), for material IDs and for texture IDs; 0 when no entry has a texture;
//   homogeneity = (16-entry lines of the ID array holding one ID) / (lines), for material and
//                 texture IDs over all entries; 0 for an empty queue.
// Every batch adds value * batch size to a running sum; the render reports sum / total size.
// All counts are integers and the doubles are formed with the same operations as the
// reference's, so the values are exact.
#include "wf_device.h"
#include "wf_internal.h"

namespace wf {

constexpr int kCohBlock = 256;
constexpr int kLineIds = 16;   // 64-byte cache line of int32 IDs (cpp:207)

// counters: [0] entries with a texture (scan total), [1] material runs, [2] texture runs,
// [3] one-ID material lines, [4] one-ID texture lines
enum { K_FILTERED = 0, K_MAT_RUNS, K_TEX_RUNS, K_MAT_LINES, K_TEX_LINES, K_COUNT };

__global__ void k_coh_flags(const int32_t* __restrict__ tex, int n_host,
                            const int32_t* __restrict__ nd, uint32_t* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < load_count(n_host, nd)) flags[i] = tex[i] >= 0 ? 1u : 0u;
}

// the IDs of the entries with a texture, in queue order
__global__ void k_coh_compact(const int32_t* __restrict__ mat, const int32_t* __restrict__ tex,
                              const uint32_t* __restrict__ pos, int n_host,
                              const int32_t* __restrict__ nd, int32_t* __restrict__ cm,
                              int32_t* __restrict__ ct) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= load_count(n_host, nd) || tex[i] < 0) return;
    cm[pos[i]] = mat[i];
    ct[pos[i]] = tex[i];
}

// adds the wave's count of set flags to *ctr (one atomic per wave)
__device__ __forceinline__ void wave_count(bool flag, uint32_t* ctr) {
    const uint32_t c = (uint32_t)__popcll(__ballot(flag));
    if (__lane_id() == 0 && c) atomicAdd(ctr, c);
}

// a run starts at every compacted entry whose ID differs from the previous one's
// (computeAverageRunLength's prevID walk, cpp:76-86)
__global__ void k_coh_runs(const int32_t* __restrict__ cm, const int32_t* __restrict__ ct,
                           int cap, uint32_t* __restrict__ counters) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = (int)counters[K_FILTERED];
    const bool in = k < cap && k < total;
    wave_count(in && (k == 0 || cm[k] != cm[k - 1]), &counters[K_MAT_RUNS]);
    wave_count(in && (k == 0 || ct[k] != ct[k - 1]), &counters[K_TEX_RUNS]);
}

// cacheLineHomogeneity (cpp:210-234): line L covers entries [16 L, min(16 L + 16, n))
__global__ void k_coh_lines(const int32_t* __restrict__ mat, const int32_t* __restrict__ tex,
                            int n_host, const int32_t* __restrict__ nd,
                            uint32_t* __restrict__ counters) {
    const int n = load_count(n_host, nd);
    const int line = blockIdx.x * blockDim.x + threadIdx.x;
    const int start = line * kLineIds;
    bool mono_m = false, mono_t = false;
    if (start < n) {
        const int end = min(start + kLineIds, n);
        mono_m = mono_t = true;
        for (int j = start + 1; j < end; ++j) {
            mono_m = mono_m && mat[j] == mat[start];
            mono_t = mono_t && tex[j] == tex[start];
        }
    }
    wave_count(mono_m, &counters[K_MAT_LINES]);
    wave_count(mono_t, &counters[K_TEX_LINES]);
}

// acc: [0..3] value * batch size summed over batches (material / texture run length, material
// / texture homogeneity), [4] summed batch sizes, [5..8] this batch's four values
__global__ void k_coh_finish(int n_host, const int32_t* __restrict__ nd,
                             const uint32_t* __restrict__ counters, int lines,
                             double* __restrict__ acc) {
    const int n = load_count(n_host, nd);
    if (n <= 0) return;   // empty batches are not counted (cpp:188)
    const uint32_t filtered = counters[K_FILTERED];
    const double mrl = filtered ? (double)filtered / (double)counters[K_MAT_RUNS] : 0.0;
    const double trl = filtered ? (double)filtered / (double)counters[K_TEX_RUNS] : 0.0;
    double mh = 0.0, th = 0.0;
    if (lines) {
        const int total_lines = (n + kLineIds - 1) / kLineIds;
        mh = (double)counters[K_MAT_LINES] / (double)total_lines;
        th = (double)counters[K_TEX_LINES] / (double)total_lines;
    }
    acc[0] += mrl * (double)n;
    acc[1] += trl * (double)n;
    acc[2] += mh * (double)n;
    acc[3] += th * (double)n;
    acc[4] += (double)n;
    acc[5] = mrl;
    acc[6] = trl;
    acc[7] = mh;
    acc[8] = th;
}

wf_status coherence_accumulate(wf_context* ctx, const int32_t* mat, const int32_t* tex, Count n,
                               bool lines, double* acc) {
    if (n.n <= 0) return WF_OK;
    const size_t nn = (size_t)n.n;
    void *pflags, *ppos, *pcm, *pct, *pcnt;
    WF_TRY(workspace(ctx, "coh_flags", nn * 4, &pflags));
    WF_TRY(workspace(ctx, "coh_pos", nn * 4, &ppos));
    WF_TRY(workspace(ctx, "coh_mat", nn * 4, &pcm));
    WF_TRY(workspace(ctx, "coh_tex", nn * 4, &pct));
    WF_TRY(workspace(ctx, "coh_counters", sizeof(uint32_t) * K_COUNT, &pcnt));
    uint32_t* counters = (uint32_t*)pcnt;
    WF_TRY(wf_fill_u32(ctx, counters, 0u, K_COUNT));
    Timed tm(ctx, "coherence", 16.0, n);
    k_coh_flags<<<grid_for(n.n, kCohBlock), kCohBlock, 0, ctx->stream>>>(tex, n.n, n.dev,
                                                                         (uint32_t*)pflags);
    WF_HIP(hipGetLastError());
    WF_TRY(exclusive_scan_u32(ctx, (uint32_t*)pflags, (uint32_t*)ppos, n,
                              (int32_t*)&counters[K_FILTERED], "coherence"));
    k_coh_compact<<<grid_for(n.n, kCohBlock), kCohBlock, 0, ctx->stream>>>(
        mat, tex, (uint32_t*)ppos, n.n, n.dev, (int32_t*)pcm, (int32_t*)pct);
    WF_HIP(hipGetLastError());
    k_coh_runs<<<grid_for(n.n, kCohBlock), kCohBlock, 0, ctx->stream>>>(
        (int32_t*)pcm, (int32_t*)pct, n.n, counters);
    WF_HIP(hipGetLastError());
    if (lines) {
        k_coh_lines<<<grid_for((n.n + kLineIds - 1) / kLineIds, kCohBlock), kCohBlock, 0,
                      ctx->stream>>>(mat, tex, n.n, n.dev, counters);
        WF_HIP(hipGetLastError());
    }
    k_coh_finish<<<1, 1, 0, ctx->stream>>>(n.n, n.dev, counters, lines ? 1 : 0, acc);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

wf_status coherence_read(wf_context* ctx, const double* acc, wf_render_stats* st) {
    double h[9];
    WF_TRY(readback(ctx, h, acc, sizeof(h)));
    // the render's averages: weighted sums over the total batch size (cpp:336-362)
    if (h[4] > 0.0) {
        st->material_run_length = h[0] / h[4];
        st->texture_run_length = h[1] / h[4];
        st->material_homogeneity = h[2] / h[4];
        st->texture_homogeneity = h[3] / h[4];
    }
    return WF_OK;
}

}  // namespace wf

using namespace wf;

extern "C" wf_status wf_coherence_stats(wf_context* ctx, const int32_t* material_ids,
                                        const int32_t* texture_ids, int32_t n,
                                        int32_t cache_lines, double out[4]) {
    clear_error();
    WF_REQUIRE(ctx && out && (n == 0 || (material_ids && texture_ids)),
               "wf_coherence_stats: NULL argument");
    WF_REQUIRE(n >= 0, "wf_coherence_stats: negative count");
    for (int k = 0; k < 4; ++k) out[k] = 0.0;
    if (n == 0) return WF_OK;
    void* pacc = nullptr;
    WF_TRY(workspace(ctx, "coh_acc", sizeof(double) * 9, &pacc));
    WF_TRY(wf_fill_u32(ctx, pacc, 0u, 18));
    WF_TRY(coherence_accumulate(ctx, material_ids, texture_ids, host_count(n), cache_lines != 0,
                                (double*)pacc));
    double h[9];
    WF_TRY(readback(ctx, h, pacc, sizeof(h)));
    for (int k = 0; k < 4; ++k) out[k] = h[5 + k];
    return WF_OK;
}
