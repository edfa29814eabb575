// Silhouette-edge table of triangle scenes on the GPU (replaces
// scene<3>::compute_silhouettes and assign_edge_indices, scene.cuh:1135-1204).
//
// Edge id = order of first appearance over positions p = 3 i + j (triangle i,
// local edge j = (v_j, v_{j+1})); slots (opp0, lo, hi, opp1) with lo < hi; each
// edge is owned by the first triangle that lists it, so a triangle's owned
// edges are a run of consecutive ids starting at estart[i].
//
//   k_edge_keys    (hi, lo) vertex pair of every position
//   radix sort     stable by (lo, hi, p): sort.hip with hi as the key and lo
//                  as the segment id
//   k_edge_heads   first position of each run of equal pairs (in sorted order
//                  that is its smallest p, the sort being stable)
//   k_edge_groups  the head of a run marks its p as a first appearance and
//                  tells every member its head
//   scan           exclusive sum of the first-appearance marks in p order:
//                  the id of each new edge, and estart[i] at p = 3 i
//   k_edge_slots   vertices of each position's edge; two positions that
//                  write the same slot side resolve to the later p, as the
//                  reference's sequential loop does
// Vertex indices are scene-global (meshes of a batch have disjoint vertex
// ranges, so their edges never merge); ids are scene-global too, each mesh's
// edges a contiguous range.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.hpp"

namespace snch {

namespace {

constexpr uint32_t EB = 256;
constexpr uint32_t SCAN_T = 1024, SCAN_V = 4, SCAN_TILE = SCAN_T * SCAN_V;

__global__ __launch_bounds__(EB) void k_edge_keys(const int32_t* __restrict__ F, uint32_t n_pos,
                                                  uint32_t* __restrict__ hi, uint32_t* __restrict__ lo,
                                                  uint32_t* __restrict__ pos)
{
    const uint32_t p = blockIdx.x * EB + threadIdx.x;
    if (p >= n_pos)
        return;
    const uint32_t i = p / 3, j = p % 3;
    uint32_t a = (uint32_t)F[3 * i + j], b = (uint32_t)F[3 * i + (j + 1) % 3];
    if (a > b) {
        const uint32_t t = a;
        a = b;
        b = t;
    }
    hi[p] = b;
    lo[p] = a;
    pos[p] = p;
}

// head[k] = 1 where sorted position k starts a run of equal (lo, hi)
__device__ __forceinline__ uint2 edge_pair(const int32_t* __restrict__ F, uint32_t p)
{
    const uint32_t i = p / 3, j = p % 3;
    const uint32_t a = (uint32_t)F[3 * i + j], b = (uint32_t)F[3 * i + (j + 1) % 3];
    return a < b ? make_uint2(a, b) : make_uint2(b, a);
}

__global__ __launch_bounds__(EB) void k_edge_heads(const uint32_t* __restrict__ sp, const int32_t* __restrict__ F,
                                                   uint32_t n_pos, uint32_t* __restrict__ head)
{
    const uint32_t k = blockIdx.x * EB + threadIdx.x;
    if (k >= n_pos)
        return;
    uint32_t h = 1u;
    if (k > 0) {
        const uint2 a = edge_pair(F, sp[k]), b = edge_pair(F, sp[k - 1]);
        h = (a.x != b.x || a.y != b.y) ? 1u : 0u;
    }
    head[k] = h;
}

// each run head: first[p] = 1 for its own p, head_of[p'] = p for every member
__global__ __launch_bounds__(EB) void k_edge_groups(const uint32_t* __restrict__ sp,
                                                    const uint32_t* __restrict__ head, uint32_t n_pos,
                                                    uint32_t* __restrict__ first, uint32_t* __restrict__ head_of)
{
    const uint32_t k = blockIdx.x * EB + threadIdx.x;
    if (k >= n_pos || !head[k])
        return;
    const uint32_t p = sp[k];
    first[p] = 1u;
    head_of[p] = p;
    for (uint32_t m = k + 1; m < n_pos && !head[m]; ++m) {
        const uint32_t q = sp[m];
        first[q] = 0u;
        head_of[q] = p;
    }
}

// exclusive sum scan of a[0, n) into out (reduce, scan of the tile sums in
// one workgroup, then each tile with its offset); out[n] = total
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t x, uint32_t* wsum)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
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
    __syncthreads();
    return before + incl;
}

__global__ __launch_bounds__(SCAN_T) void k_tile_sums(const uint32_t* __restrict__ a, uint32_t n,
                                                      uint32_t* __restrict__ sums)
{
    __shared__ uint32_t wsum[SCAN_T / 64];
    const uint32_t base = blockIdx.x * SCAN_TILE + SCAN_V * threadIdx.x;
    uint32_t s = 0;
#pragma unroll
    for (uint32_t v = 0; v < SCAN_V; ++v)
        s += base + v < n ? a[base + v] : 0u;
    const uint32_t incl = block_incl_scan(s, wsum);
    if (threadIdx.x == SCAN_T - 1)
        sums[blockIdx.x] = incl;
}

// one workgroup: exclusive scan of the tile sums in place, in chunks of 1024
__global__ __launch_bounds__(SCAN_T) void k_scan_sums(uint32_t* __restrict__ sums, uint32_t n_tiles)
{
    __shared__ uint32_t wsum[SCAN_T / 64];
    __shared__ uint32_t carry_s;
    uint32_t carry = 0;
    for (uint32_t b = 0; b < n_tiles; b += SCAN_T) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t x = i < n_tiles ? sums[i] : 0u;
        const uint32_t incl = block_incl_scan(x, wsum);
        if (i < n_tiles)
            sums[i] = carry + incl - x;
        if (threadIdx.x == SCAN_T - 1)
            carry_s = carry + incl;
        __syncthreads();
        carry = carry_s;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        sums[n_tiles] = carry;
}

__global__ __launch_bounds__(SCAN_T) void k_tile_scan(const uint32_t* __restrict__ a, uint32_t n,
                                                      const uint32_t* __restrict__ sums, uint32_t n_tiles,
                                                      uint32_t* __restrict__ out)
{
    __shared__ uint32_t wsum[SCAN_T / 64];
    const uint32_t base = blockIdx.x * SCAN_TILE + SCAN_V * threadIdx.x;
    uint32_t x[SCAN_V], s = 0;
#pragma unroll
    for (uint32_t v = 0; v < SCAN_V; ++v) {
        x[v] = base + v < n ? a[base + v] : 0u;
        s += x[v];
    }
    uint32_t run = sums[blockIdx.x] + block_incl_scan(s, wsum) - s;
#pragma unroll
    for (uint32_t v = 0; v < SCAN_V; ++v) {
        if (base + v < n)
            out[base + v] = run;
        run += x[v];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        out[n] = sums[n_tiles];
}

// estart[i] = id[3 i]; the slot sides of each position (the reference's loop
// over (i, j): opp0 of an edge met with lo -> hi along the triangle, opp1
// otherwise); win[2 e + side] = 1 + the last position writing that side
__global__ __launch_bounds__(EB) void k_edge_winners(const int32_t* __restrict__ F, uint32_t n_pos,
                                                     const uint32_t* __restrict__ id,
                                                     const uint32_t* __restrict__ head_of,
                                                     uint32_t* __restrict__ estart, uint32_t* __restrict__ win)
{
    const uint32_t p = blockIdx.x * EB + threadIdx.x;
    if (p >= n_pos)
        return;
    const uint32_t i = p / 3, j = p % 3;
    if (j == 0)
        estart[i] = id[p];
    const uint32_t e = id[head_of[p]];
    const int32_t J = F[3 * i + j], K = F[3 * i + (j + 1) % 3];
    const uint32_t side = J > K ? 1u : 0u;
    atomicMax(win + 2 * e + side, p + 1);
}

__global__ __launch_bounds__(EB) void k_edge_slots(const int32_t* __restrict__ F, uint32_t n_pos,
                                                   const uint32_t* __restrict__ id,
                                                   const uint32_t* __restrict__ head_of,
                                                   const uint32_t* __restrict__ win, int32_t* __restrict__ slots)
{
    const uint32_t p = blockIdx.x * EB + threadIdx.x;
    if (p >= n_pos)
        return;
    const uint32_t i = p / 3, j = p % 3;
    const uint32_t e = id[head_of[p]];
    const int32_t I = F[3 * i + (j + 2) % 3], J = F[3 * i + j], K = F[3 * i + (j + 1) % 3];
    const uint32_t side = J > K ? 1u : 0u;
    if (p + 1 == win[2 * e + side]) {
        // opp0 or opp1 from the side's last writer; lo / hi are the same for
        // every position of the edge
        slots[4 * (size_t)e + (side ? 3 : 0)] = I;
        slots[4 * (size_t)e + 1] = min(J, K);
        slots[4 * (size_t)e + 2] = max(J, K);
    }
}

__global__ __launch_bounds__(EB) void k_fill_edges(int4* __restrict__ slots, uint32_t* __restrict__ win, uint32_t nE)
{
    const uint32_t e = blockIdx.x * EB + threadIdx.x;
    if (e >= nE)
        return;
    slots[e] = make_int4(-1, -1, -1, -1);
    win[2 * e] = 0u;
    win[2 * e + 1] = 0u;
}

inline uint32_t nblk(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

} // namespace

hipError_t edge_table_device(snch_scene* s, hipStream_t st)
{
    const uint32_t nF = s->nF, n_pos = 3 * nF, M = s->n_meshes;
    int vbits = 1;
    while (vbits < 32 && ((s->nV - 1) >> vbits) != 0)
        ++vbits;
    DeviceBuffer hi, lo, pos, hi_alt, pos_alt, head, first, head_of, id, sums, win, hist;
    hipError_t e;
#define SNCH_TRY(x)                  \
    do {                             \
        if ((e = (x)) != hipSuccess) \
            return e;                \
    } while (0)
    const uint32_t n_tiles = nblk(n_pos, SCAN_TILE);
    SNCH_TRY(hi.ensure(4 * (size_t)n_pos));
    SNCH_TRY(lo.ensure(4 * (size_t)n_pos));
    SNCH_TRY(pos.ensure(4 * (size_t)n_pos));
    SNCH_TRY(hi_alt.ensure(4 * (size_t)n_pos));
    SNCH_TRY(pos_alt.ensure(4 * (size_t)n_pos));
    SNCH_TRY(head.ensure(4 * (size_t)n_pos));
    SNCH_TRY(first.ensure(4 * (size_t)n_pos));
    SNCH_TRY(head_of.ensure(4 * (size_t)n_pos));
    SNCH_TRY(id.ensure(4 * ((size_t)n_pos + 1)));
    SNCH_TRY(sums.ensure(4 * ((size_t)n_tiles + 1)));
    const int32_t* F = s->dF.as<int32_t>();
    hipLaunchKernelGGL(k_edge_keys, dim3(nblk(n_pos, EB)), dim3(EB), 0, st, F, n_pos, hi.as<uint32_t>(),
                       lo.as<uint32_t>(), pos.as<uint32_t>());
    bool alt = false;
    SNCH_TRY(radix_sort_pairs(hi.as<uint32_t>(), pos.as<uint32_t>(), hi_alt.as<uint32_t>(), pos_alt.as<uint32_t>(),
                              n_pos, vbits, hist, st, &alt, lo.as<uint32_t>(), vbits));
    const uint32_t* sp = alt ? pos_alt.as<uint32_t>() : pos.as<uint32_t>();
    hipLaunchKernelGGL(k_edge_heads, dim3(nblk(n_pos, EB)), dim3(EB), 0, st, sp, F, n_pos, head.as<uint32_t>());
    hipLaunchKernelGGL(k_edge_groups, dim3(nblk(n_pos, EB)), dim3(EB), 0, st, sp, head.as<uint32_t>(), n_pos,
                       first.as<uint32_t>(), head_of.as<uint32_t>());
    hipLaunchKernelGGL(k_tile_sums, dim3(n_tiles), dim3(SCAN_T), 0, st, first.as<uint32_t>(), n_pos,
                       sums.as<uint32_t>());
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_T), 0, st, sums.as<uint32_t>(), n_tiles);
    hipLaunchKernelGGL(k_tile_scan, dim3(n_tiles), dim3(SCAN_T), 0, st, first.as<uint32_t>(), n_pos,
                       sums.as<uint32_t>(), n_tiles, id.as<uint32_t>());
    // per-mesh edge ranges: the scan at each mesh's first position
    std::vector<uint32_t> idx_h(M + 1);
    for (uint32_t m = 0; m <= M; ++m) {
        const size_t at = m < M ? 3 * (size_t)s->mesh_off[m] : (size_t)n_pos;
        SNCH_TRY(hipMemcpyAsync(&idx_h[m], id.as<uint32_t>() + at, 4, hipMemcpyDeviceToHost, st));
    }
    SNCH_TRY(hipStreamSynchronize(st));
    const uint32_t nE = idx_h[M];
    s->nE = nE;
    s->mesh_eoff.assign(idx_h.begin(), idx_h.end());
    s->mesh_toff = s->mesh_eoff; // in 3D every edge has exactly one owner
    SNCH_TRY(s->dSlots.ensure(sizeof(int4) * std::max<size_t>(nE, 1)));
    SNCH_TRY(s->dEstart.ensure(sizeof(uint32_t) * ((size_t)nF + 1)));
    SNCH_TRY(win.ensure(8 * std::max<size_t>(nE, 1)));
    hipLaunchKernelGGL(k_fill_edges, dim3(nblk(std::max(nE, 1u), EB)), dim3(EB), 0, st, s->dSlots.as<int4>(),
                       win.as<uint32_t>(), nE);
    hipLaunchKernelGGL(k_edge_winners, dim3(nblk(n_pos, EB)), dim3(EB), 0, st, F, n_pos, id.as<uint32_t>(),
                       head_of.as<uint32_t>(), s->dEstart.as<uint32_t>(), win.as<uint32_t>());
    SNCH_TRY(hipMemcpyAsync(s->dEstart.as<uint32_t>() + nF, &nE, 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_edge_slots, dim3(nblk(n_pos, EB)), dim3(EB), 0, st, F, n_pos, id.as<uint32_t>(),
                       head_of.as<uint32_t>(), win.as<uint32_t>(), s->dSlots.as<int32_t>());
    SNCH_TRY(hipGetLastError());
    SNCH_TRY(hipStreamSynchronize(st));
#undef SNCH_TRY
    return hipSuccess;
}

} // namespace snch
