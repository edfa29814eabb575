// Stable LSD radix sort of (uint32 key, uint32 value) pairs for gfx950.
//
// Replaces thrust::stable_sort_by_key in bvh.cuh:452-454.  One sweep per
// 8-bit digit, after one histogram pass:
//   k_sort_init   clears the per-pass digit totals and tile counters;
//   k_hist_all    one read of the keys (and segment ids) counts the digits
//                 of every pass at once, and clears the look-back words;
//   k_onesweep    per pass: a workgroup takes the next tile (4096 or 8192 pairs) in
//                 launch order, ranks its keys with eight wave64 ballots per
//                 64-key row (the lanes holding the same digit form a peer
//                 mask) and per-wave running counts in LDS, publishes the
//                 tile's digit counts, and finds the counts of all tiles
//                 before it by decoupled look-back (Merrill and Garland,
//                 "Single-pass Parallel Prefix Scan with Decoupled
//                 Look-back", 2016): a predecessor's word holds either its
//                 own count (aggregate) or the count of every tile up to it
//                 (inclusive), flag and count in one 32-bit word.  The digit
//                 base (an exclusive scan of the pass's totals) plus that
//                 prefix plus the in-tile rank is each pair's destination.
//                 The pairs go through LDS in sorted tile order first, so
//                 each run of equal digits leaves as consecutive stores
//                 rather than one scattered store per lane.
// Ranks follow tile order (wave, row, lane), so equal keys keep their input
// order: the sort is stable, which the reference relies on (objects with
// equal Morton codes stay in index order).  Tiles are numbered by an atomic
// counter as workgroups start, so every tile a workgroup waits on has
// started before it; the wait is bounded all the same.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.hpp"

namespace snch {

namespace {

constexpr int ST = 512;          // threads per tile
// keys per thread: tiles of 4096 keys below LARGE_N, 8192 from there on.
// Larger tiles spread the per-tile fixed costs (tile number, digit base,
// look-back) over twice the keys and leave longer runs of equal digits per
// store; below ~1.5M keys they would leave CUs without a tile.  Measured per
// sort of 30-bit keys: 8.4M 304 -> 255 us, 4.2M 195 -> 158, 2.1M 126 -> 109,
// 1M 87 against 91 with the larger tiles.
constexpr int SI_SMALL = 8, SI_LARGE = 16;
constexpr uint32_t LARGE_N = 3u << 19;
constexpr int TILE_MIN = ST * SI_SMALL; // sizes the look-back words
constexpr int RADIX = 256;
constexpr int WAVES = ST / 64;
constexpr int MAX_PASSES = 8;    // up to 4 key digits and 4 segment digits
constexpr uint32_t LB_AGG = 1u << 30, LB_INC = 2u << 30, LB_COUNT = (1u << 30) - 1u;
constexpr uint32_t LB_SPIN = 1u << 22; // bound on polls of one look-back word
constexpr uint32_t LB_WIN = 8;         // look-back words read at once
constexpr uint32_t HIST_BLOCKS = 512;   // k_hist_all workgroups (2 per CU); each loops over tiles

__global__ void k_sort_init(uint32_t* __restrict__ ghist, uint32_t* __restrict__ tile_ctr)
{
    for (uint32_t i = threadIdx.x; i < MAX_PASSES * RADIX; i += blockDim.x)
        ghist[i] = 0u;
    if (threadIdx.x < MAX_PASSES)
        tile_ctr[threadIdx.x] = 0u;
}

// digit counts of every pass (pass p < key_passes: key byte p; then byte
// p - key_passes of seg[value]); clears this tile's look-back words.  Each
// thread counts SI consecutive keys and adds a run of equal digits with one
// LDS atomic; when every lane of the wave ends on the same digit (Morton
// codes of neighbouring primitives mostly share their high bytes), one lane
// adds the wave's total.  Same-address LDS atomics from many lanes run one
// lane at a time, and the LDS is shared by the CU's four SIMDs.
template <int SI>
__global__ __launch_bounds__(ST) void k_hist_all(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                 const uint32_t* __restrict__ seg, uint32_t n, int key_passes,
                                                 int seg_passes, uint32_t ntiles, uint32_t* __restrict__ ghist,
                                                 uint32_t* __restrict__ status)
{
    constexpr int TILE = ST * SI;
    __shared__ uint32_t h[MAX_PASSES][RADIX];
    const uint32_t tid = threadIdx.x;
    const int passes = key_passes + seg_passes;
    for (uint32_t i = tid; i < MAX_PASSES * RADIX; i += ST)
        h[i / RADIX][i % RADIX] = 0u;
    __syncthreads();
    // a grid-stride loop over the tiles: one flush into the global totals
    // per workgroup (the totals are few words hit by every workgroup)
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t base = tile * TILE + tid * SI;
        uint32_t kk[SI], sg[SI];
        uint32_t nv = 0; // valid keys of this thread (a prefix of its run)
        if (base + SI <= n && (reinterpret_cast<uintptr_t>(keys) & 15u) == 0u) {
#pragma unroll
            for (int c = 0; c < SI / 4; ++c) {
                const uint4 a = *reinterpret_cast<const uint4*>(keys + base + 4 * c);
                kk[4 * c] = a.x;
                kk[4 * c + 1] = a.y;
                kk[4 * c + 2] = a.z;
                kk[4 * c + 3] = a.w;
            }
            nv = SI;
        } else {
#pragma unroll
            for (int i = 0; i < SI; ++i)
                kk[i] = base + i < n ? keys[base + i] : 0u;
            nv = base < n ? min((uint32_t)SI, n - base) : 0u;
        }
#pragma unroll
        for (int i = 0; i < SI; ++i)
            sg[i] = seg_passes > 0 && (uint32_t)i < nv ? seg[vals[base + i]] : 0u;
        static_assert(SI % 4 == 0, "16-byte loads per thread");
        for (int p = 0; p < passes; ++p) {
            const bool on_key = p < key_passes;
            const int sh = 8 * (on_key ? p : p - key_passes);
            uint32_t cur = ((on_key ? kk[0] : sg[0]) >> sh) & 0xFFu, cnt = 0;
#pragma unroll
            for (int i = 0; i < SI; ++i) {
                const uint32_t d = ((on_key ? kk[i] : sg[i]) >> sh) & 0xFFu;
                if ((uint32_t)i < nv) {
                    if (d != cur) {
                        atomicAdd(&h[p][cur], cnt);
                        cur = d;
                        cnt = 0;
                    }
                    ++cnt;
                }
            }
            // the last run of each lane: one atomic for the whole wave when
            // every lane with keys ends on the same digit
            const uint64_t live = __ballot(cnt != 0u);
            const uint32_t first = __shfl(cur, live ? (uint32_t)__ffsll((unsigned long long)live) - 1u : 0u, 64);
            if (live && __ballot(cnt != 0u && cur == first) == live) {
                uint32_t sum = cnt;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1)
                    sum += __shfl_xor(sum, o, 64);
                if ((threadIdx.x & 63u) == (uint32_t)__ffsll((unsigned long long)live) - 1u)
                    atomicAdd(&h[p][first], sum);
            } else if (cnt != 0u) {
                atomicAdd(&h[p][cur], cnt);
            }
        }
        if (tid < RADIX)
            for (int p = 0; p < passes; ++p)
                status[((size_t)p * ntiles + tile) * RADIX + tid] = 0u;
    }
    __syncthreads();
    if (tid < RADIX)
        for (int p = 0; p < passes; ++p)
            if (h[p][tid] != 0u)
                atomicAdd(&ghist[p * RADIX + tid], h[p][tid]);
}

// one pass.  SEG: the digit comes from seg[value] instead of the key
template <bool SEG, int SI>
__global__ __launch_bounds__(ST) void k_onesweep(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                  const uint32_t* __restrict__ seg, uint32_t* __restrict__ kout,
                                                  uint32_t* __restrict__ vout, uint32_t n, int shift,
                                                  const uint32_t* __restrict__ ghist, uint32_t* status,
                                                  uint32_t* tile_ctr, uint32_t* err)
{
    constexpr int TILE = ST * SI;
    __shared__ uint32_t wcnt[WAVES][RADIX];
    __shared__ uint32_t toff[RADIX];
    __shared__ uint32_t wsum[WAVES];
    __shared__ uint32_t tile_s;
    __shared__ uint32_t gdelta[RADIX];
    __shared__ uint32_t sk[TILE], svl[TILE];
    __shared__ uint8_t sd[SEG ? TILE : 1]; // the digit (from seg[value]); keys carry their own
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0)
        tile_s = atomicAdd(tile_ctr, 1u);
    for (uint32_t i = tid; i < WAVES * RADIX; i += ST)
        wcnt[i / RADIX][i % RADIX] = 0u;
    // digit base: exclusive scan of the pass's digit totals (threads < RADIX)
    {
        const uint32_t x = tid < RADIX ? ghist[tid] : 0u;
        uint32_t incl = x;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= (uint32_t)o)
                incl += y;
        }
        if (lane == 63)
            wsum[w] = incl;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t q = 0; q < w; ++q)
            before += wsum[q];
        if (tid < RADIX)
            toff[tid] = before + incl - x;
    }
    const uint32_t tile = tile_s;
    const uint32_t base = tile * TILE + w * 64 * SI;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t k[SI], v[SI], rank[SI];
#pragma unroll
    for (int i = 0; i < SI; ++i) {
        uint32_t pos = base + i * 64 + lane;
        bool valid = pos < n;
        k[i] = valid ? kin[pos] : 0u;
        v[i] = valid ? vin[pos] : 0u;
    }
    uint32_t dsrc[SI];
#pragma unroll
    for (int i = 0; i < SI; ++i)
        dsrc[i] = SEG ? (base + i * 64 + lane < n ? seg[v[i]] : 0u) : k[i];
#pragma unroll
    for (int i = 0; i < SI; ++i) {
        bool valid = base + i * 64 + lane < n;
        uint32_t d = (dsrc[i] >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t r = __popcll(peers & lt);
        uint32_t before = wcnt[w][d];
        // the lowest peer publishes the new running count; every lane of the
        // wave has read the old value because LDS operations of one wave
        // complete in program order
        if (valid && r == 0)
            wcnt[w][d] = before + (uint32_t)__popcll(peers);
        rank[i] = before + r;
    }
    __syncthreads();
    // thread tid = digit: this tile's count, published at once, and its
    // start within the tile (a scan over the digits)
    uint32_t c[WAVES], own = 0, lincl = 0;
    if (tid < RADIX) {
#pragma unroll
        for (int q = 0; q < WAVES; ++q) {
            c[q] = wcnt[q][tid];
            own += c[q];
        }
        __hip_atomic_store(status + (size_t)tile * RADIX + tid, (tile == 0 ? LB_INC : LB_AGG) | own,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        lincl = own;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(lincl, o, 64);
            if (lane >= (uint32_t)o)
                lincl += y;
        }
        if (lane == 63)
            wsum[w] = lincl;
    }
    __syncthreads();
    if (tid < RADIX) {
        uint32_t lstart = lincl - own;
        for (uint32_t q = 0; q < w; ++q)
            lstart += wsum[q];
        uint32_t* word = status + (size_t)tile * RADIX + tid;
        // the count of every tile before this one (decoupled look-back):
        // predecessors are read LB_WIN at a time (independent loads), nearest
        // first, until one holds an inclusive count
        uint32_t excl = 0;
        bool done = tile == 0;
        for (uint32_t pt = tile; !done;) {
            uint32_t sv[LB_WIN];
#pragma unroll
            for (uint32_t u = 0; u < LB_WIN; ++u)
                sv[u] = u < pt ? __hip_atomic_load(status + (size_t)(pt - 1 - u) * RADIX + tid, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)
                               : LB_INC;
#pragma unroll
            for (uint32_t u = 0; u < LB_WIN; ++u) {
                if (done)
                    continue;
                uint32_t x = sv[u];
                for (uint32_t spin = 0; (x & ~LB_COUNT) == 0u; ++spin) {
                    if (spin >= LB_SPIN) { // never reached when every started tile finishes
                        atomicOr(err, 1u);
                        x = LB_INC;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    x = __hip_atomic_load(status + (size_t)(pt - 1 - u) * RADIX + tid, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                }
                excl += x & LB_COUNT; // 0 for the LB_INC stand-in below tile 0
                done = (x & LB_INC) != 0u;
            }
            pt = pt > LB_WIN ? pt - LB_WIN : 0u;
            done = done || pt == 0u;
        }
        if (tile > 0)
            __hip_atomic_store(word, LB_INC | (excl + own), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // tile position of wave q's first key of this digit; a key's
        // destination is its tile position + gdelta[digit]
        uint32_t run = lstart;
#pragma unroll
        for (int q = 0; q < WAVES; ++q) {
            wcnt[q][tid] = run;
            run += c[q];
        }
        gdelta[tid] = toff[tid] + excl - lstart;
    }
    __syncthreads();
    // keys in tile order through LDS, then out in runs of equal digits
#pragma unroll
    for (int i = 0; i < SI; ++i) {
        if (base + i * 64 + lane < n) {
            const uint32_t d = (dsrc[i] >> shift) & 0xFFu;
            const uint32_t lp = min(wcnt[w][d] + rank[i], (uint32_t)TILE - 1u); // < TILE when ranks agree
            sk[lp] = k[i];
            svl[lp] = v[i];
            if (SEG)
                sd[lp] = (uint8_t)d;
        }
    }
    __syncthreads();
    const uint32_t tile_n = min((uint32_t)TILE, n - tile * TILE);
#pragma unroll
    for (int i = 0; i < SI; ++i) {
        const uint32_t j = i * ST + tid;
        if (j < tile_n) {
            const uint32_t dst = j + gdelta[SEG ? (uint32_t)sd[j] : (sk[j] >> shift) & 0xFFu];
            if (dst < n) { // always true when histogram and ranks agree
                kout[dst] = sk[j];
                vout[dst] = svl[j];
            }
        }
    }
}

} // namespace

size_t radix_sort_scratch_bytes(uint32_t n, int passes)
{
    const size_t ntiles = (n + TILE_MIN - 1) / TILE_MIN;
    return sizeof(uint32_t) * ((size_t)MAX_PASSES * RADIX + 2 * MAX_PASSES + (size_t)passes * ntiles * RADIX);
}

size_t radix_sort_clear_words() { return (size_t)MAX_PASSES * RADIX + 2 * MAX_PASSES; }

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                            int bits, DeviceBuffer& hist, hipStream_t st, bool* result_in_alt, const uint32_t* seg,
                            int seg_bits, bool pre_cleared)
{
    *result_in_alt = false;
    if (n <= 1)
        return hipSuccess;
    const int key_passes = (bits + 7) / 8, seg_passes = seg ? (seg_bits + 7) / 8 : 0;
    if (key_passes > 4 || seg_passes > 4 || n > LB_COUNT)
        return hipErrorInvalidValue;
    if (key_passes + seg_passes == 0)
        return hipSuccess;
    hipError_t e = hist.ensure(radix_sort_scratch_bytes(n, key_passes + seg_passes));
    if (e != hipSuccess)
        return e;
    return radix_sort_pairs_in(keys, vals, keys_alt, vals_alt, n, bits, hist.as<uint32_t>(), st, result_in_alt,
                               traversal_error_word() + 1, seg, seg_bits, pre_cleared);
}

hipError_t radix_sort_pairs_in(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                               int bits, uint32_t* hist, hipStream_t st, bool* result_in_alt, uint32_t* err,
                               const uint32_t* seg, int seg_bits, bool pre_cleared)
{
    *result_in_alt = false;
    if (n <= 1)
        return hipSuccess;
    const int key_passes = (bits + 7) / 8, seg_passes = seg ? (seg_bits + 7) / 8 : 0;
    const int passes = key_passes + seg_passes;
    if (key_passes > 4 || seg_passes > 4 || n > LB_COUNT || !hist || !err)
        return hipErrorInvalidValue;
    if (passes == 0)
        return hipSuccess;
    const bool large = n >= LARGE_N;
    const uint32_t tile_keys = ST * (large ? SI_LARGE : SI_SMALL);
    const uint32_t ntiles = (n + tile_keys - 1) / tile_keys;
    uint32_t* ghist = hist;
    uint32_t* tile_ctr = ghist + MAX_PASSES * RADIX;
    uint32_t* status = tile_ctr + 2 * MAX_PASSES;
    if (!pre_cleared)
        hipLaunchKernelGGL(k_sort_init, dim3(1), dim3(256), 0, st, ghist, tile_ctr);
    if (large)
        hipLaunchKernelGGL(k_hist_all<SI_LARGE>, dim3(std::min(ntiles, HIST_BLOCKS)), dim3(ST), 0, st, keys, vals,
                           seg, n, key_passes, seg_passes, ntiles, ghist, status);
    else
        hipLaunchKernelGGL(k_hist_all<SI_SMALL>, dim3(std::min(ntiles, HIST_BLOCKS)), dim3(ST), 0, st, keys, vals,
                           seg, n, key_passes, seg_passes, ntiles, ghist, status);
    uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
    for (int pass = 0; pass < passes; ++pass) {
        const bool sp = pass >= key_passes;
        const int shift = 8 * (sp ? pass - key_passes : pass);
        uint32_t* pstat = status + (size_t)pass * ntiles * RADIX;
#define SNCH_ONESWEEP(SEG_, SI_)                                                                                 \
    hipLaunchKernelGGL((k_onesweep<SEG_, SI_>), dim3(ntiles), dim3(ST), 0, st, ki, vi, seg, ko, vo, n, shift,      \
                       ghist + pass * RADIX, pstat, tile_ctr + pass, err)
        if (sp && large)
            SNCH_ONESWEEP(true, SI_LARGE);
        else if (sp)
            SNCH_ONESWEEP(true, SI_SMALL);
        else if (large)
            SNCH_ONESWEEP(false, SI_LARGE);
        else
            SNCH_ONESWEEP(false, SI_SMALL);
#undef SNCH_ONESWEEP
        uint32_t* t = ki;
        ki = ko;
        ko = t;
        t = vi;
        vi = vo;
        vo = t;
        *result_in_alt = !*result_in_alt;
    }
    return hipGetLastError();
}

} // namespace snch
