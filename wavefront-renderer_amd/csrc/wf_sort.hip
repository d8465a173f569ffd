// Device-wide primitives: stable LSD radix sort of (u64 key, i32 value) pairs or of packed u64
// words (key above the element's index), and an exclusive prefix sum over u32 counts. Both are
// deterministic: ranks inside a tile are computed with wave-level ballots in tile order, never
// with atomics whose order could vary.
//
// Radix sort (6-, 8- or 11-bit digits by tuning; 256 threads = 4 waves of 64; tiles of 2048 pairs
// or 4096 packed words): one histogram sweep gives every pass's digit totals; each pass is then a
// single kernel (k_sweep_pass) in which each wave ranks its keys round by round: lanes holding
// the same digit are matched with one ballot per digit bit, the rank is the popcount of matching
// lower lanes plus the wave's running count for that digit; waves are offset by an exclusive
// prefix over waves (so equal digits keep input order), and tiles by the published counts of all
// earlier tiles (decoupled look-back). The tile is staged in LDS in output order and written out
// with consecutive threads on consecutive output slots.
#include <algorithm>

#include "wf_device.h"
#include "wf_internal.h"

namespace wf {

constexpr int kSortBlock = 256;
// keys per thread of a pass: (key, value) pairs stage 12 bytes per key in LDS, packed words 8,
// so packed tiles are twice as long (half the tiles, half the tile claims and look-back chains)
template <int PAYLOAD>
constexpr int sort_items() { return PAYLOAD == 0 ? 8 : 16; }
template <int PAYLOAD>
constexpr int sort_tile() { return kSortBlock * sort_items<PAYLOAD>(); }
constexpr int kSortWaves = kSortBlock / 64;

// ---- one-sweep passes ------------------------------------------------------------------------
// Digit counts of every pass come from one histogram sweep over the keys; each pass is then one
// kernel: a block takes the next tile id from a counter, ranks its tile as above, publishes
// its per-digit counts, and resolves the counts of all earlier tiles by looking back through the
// published states (a state is the tile's own count, or the inclusive prefix once known).
// Every earlier tile id was taken by a block that is already running and publishes its count
// before looking back itself, so every wait ends; the spin is still bounded, and an exhausted
// bound is reported through `fault` instead of hanging.
constexpr uint32_t kStateCount = 1u << 30;    // flag: tile count published
constexpr uint32_t kStatePrefix = 2u << 30;   // flag: inclusive prefix published
constexpr uint32_t kStateMask = (1u << 30) - 1u;
constexpr int kMaxPasses = 11;
constexpr int kSpinLimit = 1 << 22;
constexpr int kLookBack = 8;   // predecessor states read per look-back round trip

// Digit width of the one-sweep passes: BITS in {6, 8, 11} (R = 2^BITS digits). A thread owns
// DPT consecutive digits in the per-digit steps.
template <int BITS>
struct Digits {
    static constexpr int R = 1 << BITS;
    static constexpr int DPT = R > kSortBlock ? R / kSortBlock : 1;
    static constexpr int MaxPasses = (64 + BITS - 1) / BITS;
};

// Pass p ranks bits [lo_bit + BITS p, ...): BITS wide, the last pass up to BITS + 1 (the rest
// of the key). Its digit totals go to hist[p * stride ..].
template <int BITS>
__global__ __launch_bounds__(kSortBlock) void k_sweep_hist(const uint64_t* __restrict__ keys,
                                                           int n_host,
                                                           const int32_t* __restrict__ nd,
                                                           int lo_bit, int hi_bit, int passes,
                                                           int stride,
                                                           uint32_t* __restrict__ hist) {
    const int n = load_count(n_host, nd);
    // (room for a 9-bit last pass of 8-bit digits)
    __shared__ uint32_t cnt[Digits<BITS>::MaxPasses * (BITS == 8 ? 2 : 1) * Digits<BITS>::R];
    for (int i = threadIdx.x; i < passes * stride; i += kSortBlock) cnt[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int span = hi_bit - lo_bit + 1;
    const uint64_t span_mask = span >= 64 ? ~0ull : (1ull << span) - 1ull;
    // Runs of equal keys are common (every primary ray of a camera shares one origin, up to its
    // direction octant), and lanes adding to one LDS counter serialise: up to kHistLeaders times,
    // the first lane not yet counted gathers the lanes whose sorted bits equal its own and counts
    // them at once; the lanes left over count themselves. Trip counts are wave-uniform (ballots).
    constexpr int kHistLeaders = 1;
    for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n;
         base += gridDim.x * blockDim.x) {
        const int i = base + lane;
        const bool valid = i < n;
        const uint64_t k = valid ? keys[i] : 0ull;
        const uint64_t bits = (k >> lo_bit) & span_mask;
        uint64_t left = __ballot(valid);   // lanes not counted yet
        uint32_t c = 1u;                   // what this lane adds when it counts
        bool counted = false;
        for (int r = 0; r < kHistLeaders && left; ++r) {
            const int leader = __builtin_ctzll(left);
            const uint32_t llo = __shfl((uint32_t)bits, leader, 64);
            const uint32_t lhi = __shfl((uint32_t)(bits >> 32), leader, 64);
            const bool same = !counted && valid && bits == (((uint64_t)lhi << 32) | llo);
            const uint64_t peers = __ballot(same);
            if (lane == leader) c = (uint32_t)__popcll(peers);
            else if (same) counted = true;
            left &= ~peers;
        }
        if (valid && !counted) {
            for (int p = 0; p < passes; ++p) {
                // the same digit the pass ranks by
                const int shift = lo_bit + BITS * p;
                const int nbits = p == passes - 1 ? hi_bit + 1 - shift : BITS;
                atomicAdd(&cnt[p * stride + ((uint32_t)(k >> shift) & ((1u << nbits) - 1u))], c);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * stride; i += kSortBlock)
        if (cnt[i]) atomicAdd(&hist[i], cnt[i]);
}

// Exclusive prefix over the 256 threads of a block of one value per thread (wave shuffles + LDS).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t v = x;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) part[wave] = v;
    __syncthreads();
    uint32_t off = 0;
    for (int w = 0; w < wave; ++w) off += part[w];
    return off + v - x;
}

// hist[p][d] -> exclusive prefix over digits (the pass's output base of digit d); block b
// handles pass first + b, whose digits start at hist + (first + b) * stride.
template <int BITS>
__global__ __launch_bounds__(kSortBlock) void k_sweep_bases(uint32_t* hist, int first,
                                                            int stride) {
    constexpr int R = Digits<BITS>::R, DPT = Digits<BITS>::DPT;
    __shared__ uint32_t part[kSortWaves];
    uint32_t* h = hist + (size_t)(first + blockIdx.x) * stride;
    const int base = threadIdx.x * DPT;
    uint32_t x[DPT], sum = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        x[k] = base + k < R ? h[base + k] : 0u;
        sum += x[k];
    }
    uint32_t run = block_exclusive_scan(sum, part);
#pragma unroll
    for (int k = 0; k < DPT; ++k)
        if (base + k < R) {
            h[base + k] = run;
            run += x[k];
        }
}

__device__ __forceinline__ uint32_t state_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void state_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// What a pass moves: (key, value) pairs, or packed words (key above, element index in the low
// idx_bits bits) written as words, or, in the last pass of a packed sort, as their indices.
enum SweepPayload { kPairs = 0, kPacked = 1, kPackedIndex = 2 };

template <int BITS, int PAYLOAD>
__global__ __launch_bounds__(kSortBlock) void k_sweep_pass(
    const uint64_t* __restrict__ keys_in, const int32_t* __restrict__ vals_in,
    uint64_t* __restrict__ keys_out, int32_t* __restrict__ vals_out, int n_host,
    const int32_t* __restrict__ nd, int shift, uint32_t mask, int nbits,
    const uint32_t* __restrict__ bases, uint32_t* state, int32_t* tile_counter,
    int32_t* __restrict__ fault, uint32_t idx_mask) {
    constexpr int R = Digits<BITS>::R, DPT = Digits<BITS>::DPT;
    constexpr bool kVals = PAYLOAD == kPairs;
    constexpr int kSortItems = sort_items<PAYLOAD>(), kSortTile = sort_tile<PAYLOAD>();
    // the per-wave digit counts and, once every slot is known, the staged tile share one buffer
    constexpr int kCntBytes = kSortWaves * R * 4, kStageBytes = kSortTile * (kVals ? 12 : 8);
    __shared__ __align__(16) unsigned char lds[kCntBytes > kStageBytes ? kCntBytes : kStageBytes];
    uint32_t(*wave_cnt)[R] = reinterpret_cast<uint32_t(*)[R]>(lds);
    uint64_t* stage_k = reinterpret_cast<uint64_t*>(lds);
    int32_t* stage_v = reinterpret_cast<int32_t*>(lds + kSortTile * 8);
    __shared__ uint32_t digit_base[R];   // global position of tile slot 0 of each digit's run
    __shared__ uint32_t scan_part[kSortWaves];
    __shared__ int tile_s;
    const int n = load_count(n_host, nd);
    // launches over a device count cover the capacity: exactly the blocks that have a tile claim
    // one (an idle block would still pay a serialised counter atomic)
    if ((int)blockIdx.x * kSortTile >= n) return;
    if (threadIdx.x == 0) tile_s = atomicAdd(tile_counter, 1);
    __syncthreads();
    const int tile = tile_s;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kSortWaves * R; i += kSortBlock) (&wave_cnt[0][0])[i] = 0;
    __syncthreads();

    const int chunk = tile * kSortTile + wave * 64 * kSortItems;
    uint64_t k[kSortItems];
    int32_t v[kSortItems];
    uint32_t d[kSortItems];
    uint32_t r[kSortItems];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int i = chunk + it * 64 + lane;
        const bool valid = i < n;
        k[it] = valid ? keys_in[i] : 0ull;
        if (kVals) v[it] = valid ? (vals_in ? vals_in[i] : i) : 0;
        const uint32_t dig = valid ? ((uint32_t)(k[it] >> shift) & mask) : 0u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < nbits; ++b) {
            const bool bit = (dig >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t before = wave_cnt[wave][dig];
        const uint32_t rank_in = (uint32_t)__popcll(peers & lt);
        r[it] = before + rank_in;
        d[it] = dig;
        if (valid && rank_in == 0) wave_cnt[wave][dig] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // per digit (thread t owns digits t*DPT ..): wave offsets, tile counts, the tile's digit-run
    // starts, publish, look back
    const int dbase = threadIdx.x * DPT;
    uint32_t tc[DPT], tsum = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
        const int dg = dbase + q;
        uint32_t c = 0;
        if (dg < R) {
#pragma unroll
            for (int w = 0; w < kSortWaves; ++w) {
                const uint32_t x = wave_cnt[w][dg];
                wave_cnt[w][dg] = c;
                c += x;
            }
        }
        tc[q] = c;
        tsum += c;
    }
    uint32_t run = block_exclusive_scan(tsum, scan_part);
    uint32_t* my = state + (size_t)tile * R;
#pragma unroll
    for (int q = 0; q < DPT; ++q)
        if (dbase + q < R) state_store(my + dbase + q, (tile == 0 ? kStatePrefix : kStateCount) | tc[q]);
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
        const int dg = dbase + q;
        if (dg >= R) break;
        uint32_t excl = 0;
        if (tile > 0) {
            // look back kLookBack predecessors per round trip: sum published counts from the
            // nearest tile down to the first inclusive prefix; an unpublished state ends the
            // window and is loaded again (tile 0 always holds a prefix, so the walk ends)
            int j = tile - 1;
            int spins = 0;
            bool found = false;
            while (!found) {
                uint32_t st[kLookBack];
#pragma unroll
                for (int w = 0; w < kLookBack; ++w)
                    st[w] = j - w >= 0 ? state_load(state + (size_t)(j - w) * R + dg) : 0u;
                int used = 0;
#pragma unroll
                for (int w = 0; w < kLookBack; ++w) {
                    if (found || used < w || j - w < 0) continue;
                    if (st[w] & kStatePrefix) {
                        excl += st[w] & kStateMask;
                        found = true;
                    } else if (st[w] & kStateCount) {
                        excl += st[w] & kStateMask;
                        ++used;
                    }
                }
                j -= used;
                if (!found && used == 0) {
                    if (++spins > kSpinLimit) {
                        atomicOr(fault, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);   // leave the issue slots to working waves
                }
            }
            state_store(my + dg, kStatePrefix | (excl + tc[q]));
        }
        digit_base[dg] = bases[dg] + excl - run;
        // wave offsets become tile slots: slot of an element = run start + wave offset + rank
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) wave_cnt[w][dg] += run;
        run += tc[q];
    }
    __syncthreads();
    // stage the tile in LDS in output order, then write it out with consecutive threads on
    // consecutive slots (each digit's run is contiguous in the output)
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) r[it] += wave_cnt[wave][d[it]];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int i = chunk + it * 64 + lane;
        if (i < n) {
            stage_k[r[it]] = k[it];
            if (kVals) stage_v[r[it]] = v[it];
        }
    }
    __syncthreads();
    const int valid = min(kSortTile, n - tile * kSortTile);
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int j = it * kSortBlock + (int)threadIdx.x;
        if (j < valid) {
            const uint64_t key = stage_k[j];
            const uint32_t dig = (uint32_t)(key >> shift) & mask;
            const uint32_t pos = digit_base[dig] + (uint32_t)j;
            if (PAYLOAD == kPackedIndex) {
                vals_out[pos] = (int32_t)((uint32_t)key & idx_mask);
            } else {
                keys_out[pos] = key;
                if (kVals) vals_out[pos] = stage_v[j];
            }
        }
    }
}

// ---- exclusive scan ----------------------------------------------------------------------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* lds,
                                                         uint32_t* block_total) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) lds[wave] = v;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / 64; ++w) {
        uint32_t t = lds[w];
        if (w < wave) off += t;
        tot += t;
    }
    __syncthreads();
    *block_total = tot;
    return off + v - x;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                            int n_host,
                                                            const int32_t* __restrict__ nd,
                                                            uint32_t* __restrict__ sums) {
    const int n = load_count(n_host, nd);
    __shared__ uint32_t lds[kScanBlock / 64];
    const int base = blockIdx.x * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int i = base + k * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    block_exclusive_scan(s, lds, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(const uint32_t* in, uint32_t* out,
                                                          int n_host,
                                                          const int32_t* __restrict__ nd,
                                                          const uint32_t* __restrict__ offsets,
                                                          int32_t* __restrict__ total) {
    const int n = load_count(n_host, nd);
    if (total && n <= 0 && blockIdx.x == 0 && threadIdx.x == 0) *total = 0;
    if (blockIdx.x * kScanTile >= n) return;
    __shared__ uint32_t tile[kScanTile];
    __shared__ uint32_t lds[kScanBlock / 64];
    const int base = blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int i = base + k * kScanBlock + threadIdx.x;
        tile[k * kScanBlock + threadIdx.x] = i < n ? in[i] : 0u;
    }
    __syncthreads();
    uint32_t x[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        x[k] = tile[threadIdx.x * kScanItems + k];
        s += x[k];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, lds, &tot) + (offsets ? offsets[blockIdx.x] : 0u);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        tile[threadIdx.x * kScanItems + k] = run;
        run += x[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int i = base + k * kScanBlock + threadIdx.x;
        if (i < n) {
            uint32_t e = tile[k * kScanBlock + threadIdx.x];
            if (total && i == n - 1) *total = (int32_t)(e + in[i]);
            out[i] = e;
        }
    }
}

__global__ void k_iota_dev(int32_t* dst, int n_host, const int32_t* __restrict__ nd) {
    const int n = load_count(n_host, nd);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = i;
}

__global__ void k_copy_keys(const uint64_t* __restrict__ a, uint64_t* __restrict__ b, int n_host,
                            const int32_t* __restrict__ nd) {
    const int n = load_count(n_host, nd);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_copy_vals(const int32_t* __restrict__ a, int32_t* __restrict__ b, int n_host,
                            const int32_t* __restrict__ nd) {
    const int n = load_count(n_host, nd);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_sub_count(int32_t* dst, const int32_t* a, const int32_t* b) { *dst = *a - *b; }

__global__ void k_iota(int32_t* dst, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = i;
}

__global__ void k_fill_u32(uint32_t* dst, uint32_t value, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = value;
}

__global__ void k_set_i32(int32_t* dst, int32_t value) { *dst = value; }

// Streaming copy: one 16-byte element per thread over a grid that covers the buffer (measured on
// MI355X at 1 GiB: 6.2 TB/s, against 4.3-4.9 TB/s for grid-stride loops of 2048-16384 blocks with
// one to eight elements per thread, nontemporal stores or not, and 5.2 TB/s for hipMemcpy).
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ src,
                                                     float4* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// AND and OR of all keys: bits where they differ are the only ones a sort needs to look at.
__global__ void k_key_and_or(const uint64_t* __restrict__ keys, int n,
                             unsigned long long* __restrict__ out) {
    unsigned long long a = ~0ull, o = 0ull;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned long long k = keys[i];
        a &= k;
        o |= k;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        a &= __shfl_xor(a, off, 64);
        o |= __shfl_xor(o, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAnd(&out[0], a);
        atomicOr(&out[1], o);
    }
}

static wf_status scan_level(wf_context* ctx, const uint32_t* in, uint32_t* out, Count n,
                            int32_t* total, int level, const char* tag) {
    const int tiles = (n.n + kScanTile - 1) / kScanTile;
    if (tiles <= 1) {
        Timed t(ctx, tag, 8.0, n);
        k_scan_down<<<1, kScanBlock, 0, ctx->stream>>>(in, out, n.n, n.dev, nullptr, total);
        WF_HIP(hipGetLastError());
        return WF_OK;
    }
    char name[32];
    snprintf(name, sizeof(name), "scan_sums_%d", level);
    void* p = nullptr;
    WF_TRY(workspace(ctx, name, sizeof(uint32_t) * (size_t)tiles, &p));
    uint32_t* sums = (uint32_t*)p;
    {
        Timed t(ctx, tag, 4.0, n);
        k_scan_reduce<<<tiles, kScanBlock, 0, ctx->stream>>>(in, n.n, n.dev, sums);
        WF_HIP(hipGetLastError());
    }
    WF_TRY(scan_level(ctx, sums, sums, host_count(tiles), nullptr, level + 1, tag));
    {
        Timed t(ctx, tag, 8.0, n);
        k_scan_down<<<tiles, kScanBlock, 0, ctx->stream>>>(in, out, n.n, n.dev, sums, total);
        WF_HIP(hipGetLastError());
    }
    return WF_OK;
}

wf_status exclusive_scan_u32(wf_context* ctx, const uint32_t* in, uint32_t* out, Count n,
                             int32_t* total_dev, const char* tag) {
    if (n.n <= 0) {
        if (total_dev) {
            k_set_i32<<<1, 1, 0, ctx->stream>>>(total_dev, 0);
            WF_HIP(hipGetLastError());
        }
        return WF_OK;
    }
    return scan_level(ctx, in, out, n, total_dev, 0, tag);
}

wf_status iota_i32(wf_context* ctx, int32_t* dst, int32_t n) {
    if (n <= 0) return WF_OK;
    k_iota<<<grid_for(n, 256), 256, 0, ctx->stream>>>(dst, n);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

wf_status iota_count(wf_context* ctx, int32_t* dst, Count n) {
    if (n.n <= 0) return WF_OK;
    k_iota_dev<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(dst, n.n, n.dev);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

wf_status sub_count(wf_context* ctx, int32_t* dst, const int32_t* a, const int32_t* b) {
    k_sub_count<<<1, 1, 0, ctx->stream>>>(dst, a, b);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

// Identity "sort": values 0..n-1 (or vals_in), keys copied.
static wf_status sort_identity(wf_context* ctx, const uint64_t* keys_in, const int32_t* vals_in,
                               uint64_t* keys_out, int32_t* vals_out, Count n) {
    if (vals_in)
        k_copy_vals<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(vals_in, vals_out, n.n, n.dev);
    else
        k_iota_dev<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(vals_out, n.n, n.dev);
    WF_HIP(hipGetLastError());
    if (keys_in && keys_out) {
        k_copy_keys<<<grid_for(n.n, 256), 256, 0, ctx->stream>>>(keys_in, keys_out, n.n, n.dev);
        WF_HIP(hipGetLastError());
    }
    return WF_OK;
}

// One pass of BITS-wide digits with the payload of the sort (pairs, packed words, or the index
// pass that ends a packed sort).
template <int BITS>
static void launch_pass(wf_context* ctx, bool packed, int tiles, const uint64_t* ksrc,
                        const int32_t* vsrc, uint64_t* kdst, int32_t* vdst, int32_t* vals_out,
                        Count n, int shift, uint32_t mask, int nbits, const uint32_t* hp,
                        uint32_t* sp, int32_t* counter, uint32_t idx_mask, bool last) {
    if (!packed)
        k_sweep_pass<BITS, kPairs><<<tiles, kSortBlock, 0, ctx->stream>>>(
            ksrc, vsrc, kdst, vdst, n.n, n.dev, shift, mask, nbits, hp, sp, counter, ctx->fault,
            0u);
    else if (!last)
        k_sweep_pass<BITS, kPacked><<<tiles, kSortBlock, 0, ctx->stream>>>(
            ksrc, nullptr, kdst, nullptr, n.n, n.dev, shift, mask, nbits, hp, sp, counter,
            ctx->fault, idx_mask);
    else
        k_sweep_pass<BITS, kPackedIndex><<<tiles, kSortBlock, 0, ctx->stream>>>(
            ksrc, nullptr, nullptr, vals_out, n.n, n.dev, shift, mask, nbits, hp, sp, counter,
            ctx->fault, idx_mask);
}

// Passes of BITS-wide digits over key bits [lo_bit, hi_bit]. Pairs (idx_bits 0): the last pass
// writes keys_out and vals_out. Packed words (idx_bits > 0, vals_in NULL): the passes move the
// words alone, keys_out is the second ping-pong buffer, and the last pass writes the low idx_bits
// bits of each word to vals_out.
template <int BITS>
static wf_status sweep_sort(wf_context* ctx, const uint64_t* keys_in, const int32_t* vals_in,
                            uint64_t* keys_out, int32_t* vals_out, Count n, int lo_bit, int hi_bit,
                            const char* tag, int idx_bits = 0, const SortArena* arena = nullptr) {
    constexpr int R = Digits<BITS>::R;
    const int span = hi_bit - lo_bit + 1;
    const bool packed = idx_bits > 0;
    const int tile_keys = packed ? sort_tile<kPacked>() : sort_tile<kPairs>();
    const int tiles = (n.n + tile_keys - 1) / tile_keys;
    int passes = (span + BITS - 1) / BITS;
    // a one-bit last pass (the 33-bit ray key over 8-bit digits) is folded into the pass before
    // it, which then ranks a 9-bit digit: one pass fewer
    const bool wide_last = BITS == 8 && passes >= 2 && span - BITS * (passes - 1) == 1;
    if (wide_last) --passes;
    const int stride = wide_last ? 2 * R : R;   // digit slots per pass in hist and state
    if (passes > Digits<BITS>::MaxPasses || passes > kMaxPasses)
        return set_error(WF_ERR_INVALID_ARGUMENT, "sort key too wide");
    void *pk = nullptr, *pv = nullptr, *pm = nullptr;
    if (arena) {
        if (!packed) return set_error(WF_ERR_INVALID_ARGUMENT, "fixed sort buffers hold packed words");
        pk = arena->keys_alt;
    } else {
        WF_TRY(workspace(ctx, "sort_keys_alt", sizeof(uint64_t) * (size_t)n.n, &pk));
        if (!packed) WF_TRY(workspace(ctx, "sort_vals_alt", sizeof(int32_t) * (size_t)n.n, &pv));
    }
    // one zero-filled block: a tile counter per pass (an exhausted look-back is reported through
    // ctx->fault), the digit totals of every pass, the look-back states of every pass
    // (each part starts on a 256-byte boundary: a tile's state row must not share cache lines
    // with another part)
    auto up = [](size_t words) { return (words + 63) & ~(size_t)63; };
    const size_t off_hist = up(kMaxPasses), off_state = off_hist + up((size_t)stride * passes);
    const size_t meta = off_state + (size_t)stride * tiles * passes;
    if (arena) {
        if (meta > arena->meta_words)
            return set_error(WF_ERR_INVALID_ARGUMENT, "sort of %d words needs %zu meta words, the "
                             "fixed buffers hold %zu", n.n, meta, arena->meta_words);
        pm = arena->meta;
    } else {
        WF_TRY(workspace(ctx, "sort_meta", sizeof(uint32_t) * meta, &pm));
    }
    int32_t* counters = (int32_t*)pm;
    uint32_t* hist = (uint32_t*)pm + off_hist;
    uint32_t* state = (uint32_t*)pm + off_state;
    WF_TRY(wf_fill_u32(ctx, pm, 0u, meta));
    {
        Timed t(ctx, tag, 8.0, n);
        const unsigned blocks = std::min<unsigned>(grid_for(n.n, kSortBlock), 1024u);
        k_sweep_hist<BITS><<<blocks, kSortBlock, 0, ctx->stream>>>(keys_in, n.n, n.dev, lo_bit,
                                                                   hi_bit, passes, stride, hist);
        WF_HIP(hipGetLastError());
        const int narrow = wide_last ? passes - 1 : passes;
        if (narrow > 0) {
            k_sweep_bases<BITS><<<narrow, kSortBlock, 0, ctx->stream>>>(hist, 0, stride);
            WF_HIP(hipGetLastError());
        }
        if (wide_last) {
            k_sweep_bases<(BITS == 8 ? 9 : BITS)><<<1, kSortBlock, 0, ctx->stream>>>(
                hist, passes - 1, stride);
            WF_HIP(hipGetLastError());
        }
    }
    // pairs ping-pong so that the last pass writes keys_out/vals_out; packed words alternate
    // between the two key buffers and the last pass writes vals_out
    const uint64_t* ksrc = keys_in;
    const int32_t* vsrc = vals_in;
    const uint32_t idx_mask = packed ? (uint32_t)((1ull << idx_bits) - 1ull) : 0u;
    for (int p = 0; p < passes; ++p) {
        const int shift = lo_bit + p * BITS;
        const bool last = p == passes - 1;
        const int nbits = last ? hi_bit + 1 - shift : BITS;
        const uint32_t mask = (1u << nbits) - 1u;
        const bool to_out = packed ? (p % 2) == 1 : ((passes - 1 - p) % 2) == 0;
        uint64_t* kdst = to_out ? keys_out : (uint64_t*)pk;
        int32_t* vdst = packed ? vals_out : (to_out ? vals_out : (int32_t*)pv);
        const uint32_t* hp = hist + (size_t)stride * p;
        uint32_t* sp = state + (size_t)stride * tiles * p;
        {
            Timed t(ctx, tag, packed ? (last ? 12.0 : 16.0) : 24.0, n);
            if (last && wide_last)
                launch_pass<(BITS == 8 ? 9 : BITS)>(ctx, packed, tiles, ksrc, vsrc, kdst, vdst,
                                                    vals_out, n, shift, mask, nbits, hp, sp,
                                                    counters + p, idx_mask, true);
            else
                launch_pass<BITS>(ctx, packed, tiles, ksrc, vsrc, kdst, vdst, vals_out, n, shift,
                                  mask, nbits, hp, sp, counters + p, idx_mask, last);
            WF_HIP(hipGetLastError());
        }
        ksrc = kdst;
        vsrc = vdst;
    }
    return WF_OK;
}

wf_status radix_sort_packed(wf_context* ctx, const uint64_t* words, uint64_t* scratch,
                            int32_t* order, Count n, int32_t idx_bits, int32_t key_bits,
                            const char* tag, int32_t key_lo) {
    if (n.n <= 0) return WF_OK;
    if (idx_bits < 1 || idx_bits > 31 || key_bits < 0 || idx_bits + key_bits > 64)
        return set_error(WF_ERR_INVALID_ARGUMENT, "radix_sort_packed: %d key bits above %d "
                         "index bits do not fit a 64-bit word", key_bits, idx_bits);
    if ((int64_t)n.n > (1ll << idx_bits))
        return set_error(WF_ERR_INVALID_ARGUMENT, "radix_sort_packed: %d elements need more than "
                         "%d index bits", n.n, idx_bits);
    if (key_lo < 0 || key_lo > key_bits)
        return set_error(WF_ERR_INVALID_ARGUMENT, "radix_sort_packed: low key bit %d of %d", key_lo,
                         key_bits);
    if (key_bits == key_lo) return iota_count(ctx, order, n);
    int lo_bit = idx_bits + key_lo, hi_bit = idx_bits + key_bits - 1;
    if (!n.dev) {
        // as for pairs: sort only the key bits where the words differ
        void* pr = nullptr;
        WF_TRY(workspace(ctx, "sort_and_or", 16, &pr));
        unsigned long long init[2] = {~0ull, 0ull};
        WF_HIP(hipMemcpyAsync(pr, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
        Timed t(ctx, tag, 8.0 * n.n);
        const int blocks = (int)std::min<int64_t>(grid_for(n.n, 256), 1024);
        k_key_and_or<<<blocks, 256, 0, ctx->stream>>>(words, n.n, (unsigned long long*)pr);
        WF_HIP(hipGetLastError());
        unsigned long long r[2];
        WF_TRY(readback(ctx, r, pr, sizeof(r)));
        unsigned long long diff = (r[0] ^ r[1]) >> idx_bits;
        if (key_bits < 64) diff &= (1ull << key_bits) - 1ull;
        diff &= ~((1ull << key_lo) - 1ull);
        if (diff == 0) return iota_count(ctx, order, n);
        lo_bit = idx_bits + __builtin_ctzll(diff);
        hi_bit = idx_bits + 63 - __builtin_clzll(diff);
    }
    switch (ctx->tuning.radix_bits) {
        case 6: return sweep_sort<6>(ctx, words, nullptr, scratch, order, n, lo_bit, hi_bit, tag, idx_bits);
        case 11: return sweep_sort<11>(ctx, words, nullptr, scratch, order, n, lo_bit, hi_bit, tag, idx_bits);
        default: return sweep_sort<8>(ctx, words, nullptr, scratch, order, n, lo_bit, hi_bit, tag, idx_bits);
    }
}

wf_status radix_sort_packed_fixed(wf_context* ctx, uint64_t* words, int32_t* order, Count n,
                                  int32_t idx_bits, int32_t key_bits, const SortArena& arena,
                                  const char* tag) {
    if (n.n <= 0) return WF_OK;
    if (idx_bits < 1 || idx_bits > 31 || key_bits < 1 || idx_bits + key_bits > 64)
        return set_error(WF_ERR_INVALID_ARGUMENT, "radix_sort_packed_fixed: %d key bits above %d "
                         "index bits do not fit a 64-bit word", key_bits, idx_bits);
    // every key bit is sorted (no readback of the varying range): the same stable permutation
    return sweep_sort<8>(ctx, words, nullptr, words, order, n, idx_bits, idx_bits + key_bits - 1,
                         tag, idx_bits, &arena);
}

wf_status radix_sort_pairs(wf_context* ctx, const uint64_t* keys_in, const int32_t* vals_in,
                           uint64_t* keys_out, int32_t* vals_out, Count n, int32_t key_bits,
                           const char* tag) {
    if (n.n <= 0) return WF_OK;
    if (key_bits <= 0) return sort_identity(ctx, keys_in, vals_in, keys_out, vals_out, n);
    int lo_bit = 0, hi_bit = key_bits - 1;
    if (!n.dev) {
        // Only the bit range where keys differ can change the order: find it and sort just that
        // range (stable, so the result is the same permutation as sorting all key_bits).
        void* pr = nullptr;
        WF_TRY(workspace(ctx, "sort_and_or", 16, &pr));
        unsigned long long init[2] = {~0ull, 0ull};
        WF_HIP(hipMemcpyAsync(pr, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
        Timed t(ctx, tag, 8.0 * n.n);
        const int blocks = (int)std::min<int64_t>(grid_for(n.n, 256), 1024);
        k_key_and_or<<<blocks, 256, 0, ctx->stream>>>(keys_in, n.n, (unsigned long long*)pr);
        WF_HIP(hipGetLastError());
        unsigned long long r[2];
        WF_TRY(readback(ctx, r, pr, sizeof(r)));
        unsigned long long diff = (r[0] ^ r[1]);
        if (key_bits < 64) diff &= (1ull << key_bits) - 1ull;
        if (diff == 0) return sort_identity(ctx, keys_in, vals_in, keys_out, vals_out, n);
        lo_bit = __builtin_ctzll(diff);
        hi_bit = 63 - __builtin_clzll(diff);
    }
    switch (ctx->tuning.radix_bits) {
        case 6: return sweep_sort<6>(ctx, keys_in, vals_in, keys_out, vals_out, n, lo_bit, hi_bit, tag);
        case 11: return sweep_sort<11>(ctx, keys_in, vals_in, keys_out, vals_out, n, lo_bit, hi_bit, tag);
        default: return sweep_sort<8>(ctx, keys_in, vals_in, keys_out, vals_out, n, lo_bit, hi_bit, tag);
    }
}

}  // namespace wf

extern "C" wf_status wf_stream_copy(wf_context* ctx, void* dst, const void* src, size_t bytes) {
    WF_REQUIRE(ctx, "wf_stream_copy: NULL context");
    WF_REQUIRE(bytes % 16 == 0, "wf_stream_copy: size %zu is not a multiple of 16", bytes);
    WF_REQUIRE(bytes == 0 || (dst && src), "wf_stream_copy: NULL buffer");
    if (bytes == 0) return WF_OK;
    const size_t n = bytes / 16;
    WF_REQUIRE((n + 255) / 256 <= 0x7FFFFFFFu, "wf_stream_copy: %zu bytes is too large", bytes);
    const unsigned grid = (unsigned)((n + 255) / 256);
    wf::Timed tm(ctx, "stream_copy", 2.0 * (double)bytes);
    wf::k_stream_copy<<<grid, 256, 0, ctx->stream>>>((const float4*)src, (float4*)dst, n);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

extern "C" wf_status wf_fill_u32(wf_context* ctx, void* dst, uint32_t value, size_t count) {
    WF_REQUIRE(ctx, "wf_fill_u32: NULL context");
    if (count == 0) return WF_OK;
    wf::k_fill_u32<<<(unsigned)((count + 255) / 256), 256, 0, ctx->stream>>>((uint32_t*)dst, value,
                                                                           count);
    WF_HIP(hipGetLastError());
    return WF_OK;
}
