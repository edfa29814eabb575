// Stable LSD radix sort of (uint32 key, uint32 value) pairs for gfx950.
//
// Replaces thrust::stable_sort_by_key in bvh.cuh:452-454.  Each 8-bit pass is
// reduce-then-scan:
//   upsweep    one tile (256 threads x 16 keys) -> 256-bin digit histogram,
//              stored digit-major so one exclusive scan gives every tile's
//              global write offset per digit;
//   scan       single-block exclusive scan of the digit-major histograms;
//   downsweep  each wave ranks its 64-key rows with eight wave64 ballots (the
//              lanes holding the same digit form a peer mask), keeps per-wave
//              running digit counts in LDS, adds the per-wave prefix and the
//              tile offset, and scatters.
// Ranks follow tile order (wave, row, lane), so equal keys keep their input
// order: the sort is stable, which the reference relies on (objects with
// equal Morton codes stay in index order).
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace snch {

namespace {

constexpr int ST = 256;          // threads per tile
constexpr int SI = 16;           // keys per thread
constexpr int TILE = ST * SI;    // 4096 keys
constexpr int RADIX = 256;
constexpr int WAVES = ST / 64;

// SEG: the digit comes from seg[value] (a segment id per value) instead of
// the key, which sorts stably by segment after the key passes
template <bool SEG>
__global__ __launch_bounds__(ST) void k_upsweep(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                 const uint32_t* __restrict__ seg, uint32_t n, int shift,
                                                 uint32_t ntiles, uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[RADIX];
    const uint32_t tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * TILE;
#pragma unroll 4
    for (int i = 0; i < SI; ++i) {
        uint32_t pos = base + i * ST + tid;
        if (pos < n)
            atomicAdd(&h[((SEG ? seg[vals[pos]] : keys[pos]) >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[tid * ntiles + blockIdx.x] = h[tid];
}

// exclusive scan of a[0..len) (len a multiple of 4) by one block of 1024
// threads: coalesced 4096-element tiles with a running carry
__global__ __launch_bounds__(1024) void k_scan_exclusive(uint32_t* __restrict__ a, uint32_t len)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_s;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < len; base += 4096) {
        const uint32_t idx = base + 4 * tid;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (idx < len)
            v = *reinterpret_cast<const uint4*>(a + idx);
        const uint32_t s = v.x + v.y + v.z + v.w;
        uint32_t incl = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= (uint32_t)o)
                incl += y;
        }
        if (lane == 63)
            wsum[w] = incl;
        __syncthreads();
        if (w == 0) {
            const uint32_t x0 = lane < 16 ? wsum[lane] : 0u;
            uint32_t x = x0;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const uint32_t y = __shfl_up(x, o, 64);
                if (lane >= (uint32_t)o)
                    x += y;
            }
            if (lane < 16)
                wsum[lane] = x - x0;
        }
        __syncthreads();
        const uint32_t ex = carry + wsum[w] + incl - s;
        if (idx < len)
            *reinterpret_cast<uint4*>(a + idx) = make_uint4(ex, ex + v.x, ex + v.x + v.y, ex + v.x + v.y + v.z);
        if (tid == 1023)
            carry_s = ex + s;
        __syncthreads();
        carry = carry_s;
        __syncthreads();
    }
}

// multi-block exclusive scan for long histograms: per-chunk sums, a scan of
// the sums, then each chunk scanned with its offset (chunks of 4096)
__device__ __forceinline__ uint32_t block_sum_1024(uint32_t x, uint32_t* wsum)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        x += __shfl_xor(x, o, 64);
    if (lane == 0)
        wsum[w] = x;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        t += wsum[k];
    return t;
}

__global__ __launch_bounds__(1024) void k_chunk_sums(const uint32_t* __restrict__ a, uint32_t len,
                                                     uint32_t* __restrict__ sums)
{
    __shared__ uint32_t wsum[16];
    const uint32_t idx = blockIdx.x * 4096 + 4 * threadIdx.x;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (idx < len)
        v = *reinterpret_cast<const uint4*>(a + idx);
    const uint32_t t = block_sum_1024(v.x + v.y + v.z + v.w, wsum);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = t;
}

__global__ __launch_bounds__(1024) void k_chunk_scan(uint32_t* __restrict__ a, uint32_t len,
                                                     const uint32_t* __restrict__ offs)
{
    __shared__ uint32_t wsum[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t idx = blockIdx.x * 4096 + 4 * tid;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (idx < len)
        v = *reinterpret_cast<const uint4*>(a + idx);
    const uint32_t s = v.x + v.y + v.z + v.w;
    uint32_t incl = s;
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
    for (uint32_t k = 0; k < w; ++k)
        before += wsum[k];
    const uint32_t ex = offs[blockIdx.x] + before + incl - s;
    if (idx < len)
        *reinterpret_cast<uint4*>(a + idx) = make_uint4(ex, ex + v.x, ex + v.x + v.y, ex + v.x + v.y + v.z);
}

template <bool SEG>
__global__ __launch_bounds__(ST) void k_downsweep(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                   const uint32_t* __restrict__ seg, uint32_t* __restrict__ kout,
                                                   uint32_t* __restrict__ vout, uint32_t n, int shift,
                                                   uint32_t ntiles, const uint32_t* __restrict__ offs)
{
    __shared__ uint32_t wcnt[WAVES][RADIX];
    __shared__ uint32_t toff[RADIX];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int k = 0; k < WAVES; ++k)
        wcnt[k][tid] = 0;
    toff[tid] = offs[tid * ntiles + blockIdx.x];
    __syncthreads();

    const uint32_t base = blockIdx.x * TILE + w * 64 * SI;
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
    {
        uint32_t c[WAVES];
#pragma unroll
        for (int q = 0; q < WAVES; ++q)
            c[q] = wcnt[q][tid];
        uint32_t run = toff[tid];
#pragma unroll
        for (int q = 0; q < WAVES; ++q) {
            wcnt[q][tid] = run;
            run += c[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SI; ++i) {
        bool valid = base + i * 64 + lane < n;
        if (valid) {
            uint32_t d = (dsrc[i] >> shift) & 0xFFu;
            uint32_t dst = wcnt[w][d] + rank[i];
            if (dst < n) { // always true when histogram and ranks agree
                kout[dst] = k[i];
                vout[dst] = v[i];
            }
        }
    }
}

} // namespace

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                            int bits, DeviceBuffer& hist, hipStream_t st, bool* result_in_alt, const uint32_t* seg,
                            int seg_bits)
{
    *result_in_alt = false;
    if (n <= 1)
        return hipSuccess;
    const uint32_t ntiles = (n + TILE - 1) / TILE;
    hipError_t e = hist.ensure(sizeof(uint32_t) * (size_t)RADIX * ntiles);
    if (e != hipSuccess)
        return e;
    const uint32_t hlen = RADIX * ntiles, nchunks = (hlen + 4095) / 4096;
    const bool chunked = nchunks > 4;
    if (chunked) {
        e = hist.ensure(sizeof(uint32_t) * ((size_t)hlen + ((nchunks + 3) & ~3u)));
        if (e != hipSuccess)
            return e;
    }
    uint32_t* h = hist.as<uint32_t>();
    uint32_t* sums = h + hlen;
    uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
    const int key_passes = (bits + 7) / 8, seg_passes = seg ? (seg_bits + 7) / 8 : 0;
    for (int pass = 0; pass < key_passes + seg_passes; ++pass) {
        const bool sp = pass >= key_passes;
        const int shift = 8 * (sp ? pass - key_passes : pass);
        if (sp)
            hipLaunchKernelGGL(k_upsweep<true>, dim3(ntiles), dim3(ST), 0, st, ki, vi, seg, n, shift, ntiles, h);
        else
            hipLaunchKernelGGL(k_upsweep<false>, dim3(ntiles), dim3(ST), 0, st, ki, vi, seg, n, shift, ntiles, h);
        if (chunked) {
            hipLaunchKernelGGL(k_chunk_sums, dim3(nchunks), dim3(1024), 0, st, h, hlen, sums);
            hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, st, sums, (nchunks + 3) & ~3u);
            hipLaunchKernelGGL(k_chunk_scan, dim3(nchunks), dim3(1024), 0, st, h, hlen, sums);
        } else {
            hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, st, h, hlen);
        }
        if (sp)
            hipLaunchKernelGGL(k_downsweep<true>, dim3(ntiles), dim3(ST), 0, st, ki, vi, seg, ko, vo, n, shift, ntiles,
                               h);
        else
            hipLaunchKernelGGL(k_downsweep<false>, dim3(ntiles), dim3(ST), 0, st, ki, vi, seg, ko, vo, n, shift,
                               ntiles, h);
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
