// LBVH construction for triangle scenes on gfx950 (replaces bvh::construct,
// bvh.cuh:380-613, and the per-object getters of scene.cuh:870-961).
//
//   k_init_scene_box  scene-box accumulators to (+inf, -inf), list and sort
//                     counters to zero
//   k_leafgen         one thread per triangle: leaf AABB (vertex extent
//                     +- FLT_EPSILON), leaf normal cone from the owned
//                     silhouette edges (the edges' normals computed one
//                     edge per lane across the wave), the packed
//                     silhouette-edge records,
//                     and a wave64 + LDS reduction of the scene AABB into
//                     one partial box per workgroup
//   k_box_final       the partial boxes into the scene AABB (ordered-int
//                     atomics, one per workgroup and component)
//   k_morton          30-bit Morton code of each leaf-box centroid
//   radix sort        stable (code, object index) sort, sort.hip
//   k_karras          Karras (2012) radix-tree emission.  Keys are compared
//                     as the 64-bit (code << 32 | index): with unique codes
//                     this is exactly the reference's 32-bit tree, and with
//                     duplicate codes it is the reference's 64-bit fallback
//                     (bvh.cuh:459-476), so no uniqueness pass or host
//                     round trip is needed.
//   k_refit           bottom-up AABB + cone merge over subtrees of at most
//                     1024 leaves in LDS (the parent box centroid feeds the
//                     cone merge, as in bvh.cuh:598); writes the
//                     reference-layout node data and the packed traversal
//                     records
//   k_refit_mid       the same above them, up to subtrees of 16384 leaves,
//                     one workgroup per subtree
//   k_refit_top       the nodes above those, one workgroup
//
// A scene may hold several meshes: every kernel runs over all of them at
// once.  Each mesh gets its own scene box, Morton codes in that box, and its
// own tree (the sort orders by (mesh, code, index)); a mesh's tree is
// bit-identical to the tree of the same mesh built alone.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.hpp"

namespace snch {

namespace {

constexpr int BT = 256;

__device__ __forceinline__ int f2ord(float f)
{
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float ord2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__device__ __forceinline__ v3 ld3(const float4* V, int i)
{
    float4 a = V[i];
    return mk(a.x, a.y, a.z);
}

// mesh of primitive i (pmesh is null for a single mesh)
__device__ __forceinline__ uint32_t mesh_of_prim(const uint32_t* __restrict__ pmesh, uint32_t i)
{
    return pmesh ? pmesh[i] : 0u;
}

// last mesh whose key (node base or internal-record base) is <= x
template <bool BY_NODE>
__device__ __forceinline__ uint32_t mesh_search(const MeshDesc* __restrict__ d, uint32_t M, uint32_t x)
{
    uint32_t lo = 0, hi = M; // answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t key = BY_NODE ? d[mid].nb : d[mid].off - mid;
        if (key <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// scene-box accumulators (+inf, -inf) of every mesh, 8 ints per mesh
// (and clears the list counters k_morton and k_karras append to, and the
// radix sort's digit totals and tile counters)
__global__ void k_init_scene_box(int* bits, uint32_t M, uint32_t* list_a, uint32_t* list_b, uint32_t* list_c,
                                 uint32_t* list_d, uint32_t* sort_words, uint32_t n_sort_words)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0) // the radix sort's digit totals and tile counters
        for (uint32_t k = threadIdx.x; k < n_sort_words; k += blockDim.x)
            sort_words[k] = 0u;
    if (t == 0) {
        list_a[0] = 0u;
        list_a[1] = 0u; // k_refit's next cut
        *list_b = 0u;
        *list_c = 0u;
        *list_d = 0u;
    }
    if (t >= 8 * M)
        return;
    const uint32_t k = t & 7u;
    bits[t] = k < 3 ? f2ord(F_INF) : f2ord(-F_INF);
}

// block reduction of a scene box: wave butterfly, then, when the whole block
// holds one mesh (always for a single mesh), the block's box in its partial
// record part[8 b, 8 b + 8) (box slots 0-5, then the mesh) for k_box_final;
// else atomics per wave or per lane and a record marked INVALID.  Thousands
// of workgroups updating the same six words one after the other would
// serialise on them.
template <int NV>
__device__ __forceinline__ void reduce_scene_box(float (&v)[NV], uint32_t mesh, bool valid, int* scene_bits,
                                                 const int* slot_of, float* __restrict__ part)
{
    __shared__ float red[BT / 64][NV];
    __shared__ uint32_t redm[BT / 64];
    __shared__ int redu[BT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t m0 = __shfl(mesh, 0, 64), m63 = __shfl(mesh, 63, 64);
    const bool uni = m0 == m63;
    if (uni) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
            for (int k = 0; k < NV; ++k)
                v[k] = k < NV / 2 ? fmn(v[k], __shfl_xor(v[k], o, 64)) : fmx(v[k], __shfl_xor(v[k], o, 64));
        }
    } else if (valid) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if (k < NV / 2)
                atomicMin(scene_bits + 8 * mesh + slot_of[k], f2ord(v[k]));
            else
                atomicMax(scene_bits + 8 * mesh + slot_of[k], f2ord(v[k]));
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k)
            red[w][k] = v[k];
        redm[w] = m0;
        redu[w] = uni ? 1 : 0;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const int k = threadIdx.x;
        float* rec = part + 8 * blockIdx.x;
        bool same = true;
        for (int q = 0; q < BT / 64; ++q)
            same = same && redu[q] && redm[q] == redm[0];
        if (k == 7)
            rec[6] = __uint_as_float(same ? redm[0] : INVALID);
        if (NV == 4 && k == 6) { // slots of z in 2D
            rec[2] = F_INF;
            rec[5] = -F_INF;
        }
        if (k < NV) {
            if (same) {
                float acc = red[0][k];
                for (int q = 1; q < BT / 64; ++q)
                    acc = k < NV / 2 ? fmn(acc, red[q][k]) : fmx(acc, red[q][k]);
                rec[slot_of[k]] = acc;
            } else {
                for (int q = 0; q < BT / 64; ++q) {
                    if (!redu[q])
                        continue;
                    if (k < NV / 2)
                        atomicMin(scene_bits + 8 * redm[q] + slot_of[k], f2ord(red[q][k]));
                    else
                        atomicMax(scene_bits + 8 * redm[q] + slot_of[k], f2ord(red[q][k]));
                }
            }
        }
    }
}

// the partial boxes of reduce_scene_box into the mesh accumulators: one
// record per thread, a wave butterfly when the wave's records hold one mesh,
// then one atomic per component per workgroup when all its waves do
constexpr int BF = 1024;
constexpr uint32_t WAVE_EMPTY = 0xFFFFFFFEu; // no record of a mesh in the wave
__global__ __launch_bounds__(BF) void k_box_final(const float* __restrict__ part, uint32_t n_part,
                                                  int* __restrict__ scene_bits)
{
    __shared__ float red[BF / 64][6];
    __shared__ uint32_t redm[BF / 64]; // the wave's mesh, WAVE_EMPTY, or INVALID (mixed: already applied)
    const uint32_t i = blockIdx.x * BF + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t mesh = INVALID;
    float v[6] = {F_INF, F_INF, F_INF, -F_INF, -F_INF, -F_INF};
    if (i < n_part) {
        mesh = __float_as_uint(part[8 * i + 6]);
        if (mesh != INVALID) {
#pragma unroll
            for (int k = 0; k < 6; ++k)
                v[k] = part[8 * i + k];
        }
    }
    const uint64_t live = __ballot(mesh != INVALID);
    const uint32_t m0 = live ? __shfl(mesh, __ffsll((long long)live) - 1, 64) : WAVE_EMPTY;
    const bool uni = live == 0ull || __ballot(mesh != INVALID && mesh != m0) == 0ull;
    if (uni) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
            for (int k = 0; k < 6; ++k)
                v[k] = k < 3 ? fmn(v[k], __shfl_xor(v[k], o, 64)) : fmx(v[k], __shfl_xor(v[k], o, 64));
        }
    } else if (mesh != INVALID) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            if (k < 3)
                atomicMin(scene_bits + 8 * mesh + k, f2ord(v[k]));
            else
                atomicMax(scene_bits + 8 * mesh + k, f2ord(v[k]));
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k)
            red[w][k] = v[k];
        redm[w] = uni ? m0 : INVALID;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t mb = WAVE_EMPTY;
        bool same = true;
        for (int q = 0; q < BF / 64; ++q) {
            const uint32_t mq = redm[q];
            if (mq == WAVE_EMPTY)
                continue;
            if (mb == WAVE_EMPTY)
                mb = mq;
            same = same && mq == mb && mq != INVALID;
        }
        if (same && mb != WAVE_EMPTY) {
            float acc = k < 3 ? F_INF : -F_INF;
            for (int q = 0; q < BF / 64; ++q)
                if (redm[q] != WAVE_EMPTY)
                    acc = k < 3 ? fmn(acc, red[q][k]) : fmx(acc, red[q][k]);
            if (k < 3)
                atomicMin(scene_bits + 8 * mb + k, f2ord(acc));
            else
                atomicMax(scene_bits + 8 * mb + k, f2ord(acc));
        } else if (!same) {
            for (int q = 0; q < BF / 64; ++q) {
                const uint32_t mq = redm[q];
                if (mq == WAVE_EMPTY || mq == INVALID)
                    continue;
                if (k < 3)
                    atomicMin(scene_bits + 8 * mq + k, f2ord(red[q][k]));
                else
                    atomicMax(scene_bits + 8 * mq + k, f2ord(red[q][k]));
            }
        }
    }
}

// Per-wave LDS of k_leafgen's flattened edge work: one round of 64 edge
// slots (edge id, owner lane) with the owner's box centroid and mesh edge
// offset, and per slot the results the owner folds into its cone: unit edge
// normal, unit face normals, the two radius terms and the face bits.
struct LeafEdgeLds {
    uint32_t edge[64];
    uint32_t owner[64];
    float cx[64], cy[64], cz[64];
    uint32_t eoff[64];
    float n[3][64], n0[3][64], n1[3][64];
    float ra[64], rb[64];
    uint32_t faces[64];
};

// Leaf box and leaf cone of each triangle (scene.cuh:870-961) and the packed
// silhouette-edge records.  A triangle owns 0..3 edges; rather than looping
// over them per lane (three divergent iterations per wave), the wave lists
// every owned edge as a run of slots and, in rounds of 64 slots, each lane
// computes the normals and radius terms of one edge and writes its record.
// Each owner then folds its edges in their original order, so the cone's
// sums and maxima are formed exactly as in the per-triangle loop.
__global__ __launch_bounds__(BT) void k_leafgen(const float4* __restrict__ V, const int32_t* __restrict__ F,
                                                const int4* __restrict__ slots, const uint32_t* __restrict__ estart,
                                                uint32_t nF, int conservative, float4* __restrict__ leafrec,
                                                float4* __restrict__ tedge,
                                                int* __restrict__ scene_bits, const uint32_t* __restrict__ pmesh,
                                                const MeshDesc* __restrict__ desc, float* __restrict__ part)
{
    __shared__ LeafEdgeLds lds[BT / 64];
    LeafEdgeLds& L = lds[threadIdx.x >> 6];
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t i = blockIdx.x * BT + threadIdx.x;
    const uint32_t mesh = mesh_of_prim(pmesh, i < nF ? i : nF - 1);
    const uint32_t eoff = desc[mesh].eoff;
    box3 b;
    b.lo = mk(F_INF, F_INF, F_INF);
    b.hi = mk(-F_INF, -F_INF, -F_INF);
    uint32_t e0 = 0, my = 0;
    v3 c = mk(0.0f, 0.0f, 0.0f);
    if (i < nF) {
        const v3 p0 = ld3(V, F[3 * i]), p1 = ld3(V, F[3 * i + 1]), p2 = ld3(V, F[3 * i + 2]);
        b = box_of_point(p0);
        box_expand(b, p1);
        box_expand(b, p2);
        c = box_centroid(b);
        e0 = estart[i];
        const uint32_t e1 = estart[i + 1];
        my = e1 > e0 ? min(e1 - e0, 3u) : 0u;
    }
    L.cx[lane] = c.x;
    L.cy[lane] = c.y;
    L.cz[lane] = c.z;
    L.eoff[lane] = eoff;
    const uint64_t b0 = __ballot(my & 1u), b1 = __ballot(my & 2u);
    const uint32_t off = (uint32_t)__popcll(b0 & lt) + 2u * (uint32_t)__popcll(b1 & lt);
    const uint32_t total = (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1);
    // leaf cone accumulators (scene.cuh:887-961)
    cone3 r;
    r.axis = mk(0.0f, 0.0f, 0.0f);
    r.half_angle = F_PI;
    r.radius = 0.0f;
    bool all_two = true;
    v3 nn[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        nn[k][0] = nn[k][1] = mk(0.0f, 0.0f, 0.0f);
    for (uint32_t base = 0; base < total; base += 64) {
#pragma unroll
        for (uint32_t k = 0; k < 3; ++k) {
            const uint32_t sl = off + k;
            if (k < my && sl >= base && sl < base + 64) {
                L.edge[sl - base] = e0 + k;
                L.owner[sl - base] = lane;
            }
        }
        // one edge per lane (LDS operations of one wave complete in order)
        if (lane < min(64u, total - base)) {
            const uint32_t e = L.edge[lane], o = L.owner[lane];
            const v3 co = mk(L.cx[o], L.cy[o], L.cz[o]);
            const int4 s = slots[e];
            const bool has0 = s.w != -1, has1 = s.x != -1;
            // the edge's four vertices, each read once; slots are (opp0, lo, hi,
            // opp1), face 0 is (pa, pb, opp0) and face 1 is (pb, pa, opp1), and a
            // face normal is the cross product of its first two edges from its
            // first vertex (scene.cuh:747-772)
            const v3 pa = ld3(V, s.y), pb = ld3(V, s.z);
            const v3 o0 = has0 ? ld3(V, s.w) : pa, o1 = has1 ? ld3(V, s.x) : pb;
            v3 raw0 = mk(0.0f, 0.0f, 0.0f), raw1 = mk(0.0f, 0.0f, 0.0f);
            v3 n = mk(0.0f, 0.0f, 0.0f);
            if (has0) {
                raw0 = cross3(mk(pb.x - pa.x, pb.y - pa.y, pb.z - pa.z), mk(o0.x - pa.x, o0.y - pa.y, o0.z - pa.z));
                n = mk(n.x + raw0.x, n.y + raw0.y, n.z + raw0.z);
            }
            if (has1) {
                raw1 = cross3(mk(pa.x - pb.x, pa.y - pb.y, pa.z - pb.z), mk(o1.x - pb.x, o1.y - pb.y, o1.z - pb.z));
                n = mk(n.x + raw1.x, n.y + raw1.y, n.z + raw1.z);
            }
            n = normalize3(n);
            float ra, rb = 0.0f;
            if (conservative) {
                ra = len3(mk(pa.x - co.x, pa.y - co.y, pa.z - co.z));
                rb = len3(mk(pb.x - co.x, pb.y - co.y, pb.z - co.z));
            } else {
                const v3 m = mk((pa.x + pb.x) / 2.0f, (pa.y + pb.y) / 2.0f, (pa.z + pb.z) / 2.0f);
                ra = len3(mk(m.x - co.x, m.y - co.y, m.z - co.z));
            }
            const v3 u0 = has0 ? normalize3(raw0) : mk(0.0f, 0.0f, 0.0f);
            const v3 u1 = has1 ? normalize3(raw1) : mk(0.0f, 0.0f, 0.0f);
            const uint32_t faces = (has0 ? 1u : 0u) | (has1 ? 2u : 0u);
            // packed silhouette-edge record (each edge has exactly one owner)
            tedge[4 * e + 0] = make_float4(pa.x, pa.y, pa.z, __uint_as_float(e - L.eoff[o]));
            tedge[4 * e + 1] = make_float4(pb.x, pb.y, pb.z, __uint_as_float(faces));
            tedge[4 * e + 2] = make_float4(u0.x, u0.y, u0.z, 0.0f);
            tedge[4 * e + 3] = make_float4(u1.x, u1.y, u1.z, 0.0f);
            L.n[0][lane] = n.x;
            L.n[1][lane] = n.y;
            L.n[2][lane] = n.z;
            L.n0[0][lane] = u0.x;
            L.n0[1][lane] = u0.y;
            L.n0[2][lane] = u0.z;
            L.n1[0][lane] = u1.x;
            L.n1[1][lane] = u1.y;
            L.n1[2][lane] = u1.z;
            L.ra[lane] = ra;
            L.rb[lane] = rb;
            L.faces[lane] = faces;
        }
        // each owner folds its edges of this round, in edge order
#pragma unroll
        for (uint32_t k = 0; k < 3; ++k) {
            const uint32_t sl = off + k;
            if (k < my && sl >= base && sl < base + 64) {
                const uint32_t q = sl - base;
                r.axis.x += L.n[0][q];
                r.axis.y += L.n[1][q];
                r.axis.z += L.n[2][q];
                r.radius = smax(r.radius, L.ra[q]);
                if (conservative)
                    r.radius = smax(r.radius, L.rb[q]);
                all_two = all_two && L.faces[q] == 3u;
                nn[k][0] = mk(L.n0[0][q], L.n0[1][q], L.n0[2][q]);
                nn[k][1] = mk(L.n1[0][q], L.n1[1][q], L.n1[2][q]);
            }
        }
    }
    if (i < nF) {
        if (my == 0) {
            r.half_angle = -F_PI;
        } else if (!all_two) {
            r.half_angle = F_PI;
        } else {
            const float an = len3(r.axis);
            if (an > F_EPS) {
                r.axis.x /= an;
                r.axis.y /= an;
                r.axis.z /= an;
                // the largest angle to a face normal: dm::acosf_ never
                // increases over [-1, 1] (checked for every float), so the
                // maximum of the angles is the angle of the smallest cosine,
                // bit for bit, for one arccos instead of up to six
                float cmin = 1.0f;
#pragma unroll
                for (uint32_t k = 0; k < 3; ++k) {
                    if (k >= my)
                        break;
                    for (int jj = 0; jj < 2; ++jj)
                        cmin = smin(cmin, clamp11(dot3(r.axis, nn[k][jj])));
                }
                r.half_angle = smax(0.0f, dm::acosf_(cmin));
            }
        }
        // the leaf record: box and cone in three float4 (no padding lanes)
        leafrec[3 * i] = make_float4(b.lo.x, b.lo.y, b.lo.z, b.hi.x);
        leafrec[3 * i + 1] = make_float4(b.hi.y, b.hi.z, r.axis.x, r.axis.y);
        leafrec[3 * i + 2] = make_float4(r.axis.z, r.half_angle, r.radius, 0.0f);
    }
    // scene AABB of the primitive's mesh
    float v[6] = {b.lo.x, b.lo.y, b.lo.z, b.hi.x, b.hi.y, b.hi.z};
    const int slot_of[6] = {0, 1, 2, 3, 4, 5};
    reduce_scene_box<6>(v, mesh, i < nF, scene_bits, slot_of, part);
}

// scene<2>: leaf box (scene.cuh:420-433) and leaf cone from the owned
// silhouette vertices (scene.cuh:435-502); slots are per owned silhouette
// (prev, v, next, -1).  Also writes the packed silhouette-vertex records.
__global__ __launch_bounds__(BT) void k_leafgen2d(const float4* __restrict__ V, const int32_t* __restrict__ S,
                                                  const int4* __restrict__ slots, const uint32_t* __restrict__ estart,
                                                  uint32_t nS, float4* __restrict__ leafrec,
                                                  float4* __restrict__ tedge,
                                                  int* __restrict__ scene_bits, const uint32_t* __restrict__ pmesh,
                                                  float* __restrict__ part)
{
    const uint32_t i = blockIdx.x * BT + threadIdx.x;
    const uint32_t mesh = mesh_of_prim(pmesh, i < nS ? i : nS - 1);
    box3 b;
    b.lo = mk(F_INF, F_INF, 0.0f);
    b.hi = mk(-F_INF, -F_INF, 0.0f);
    if (i < nS) {
        const v3 p0 = ld3(V, S[2 * i]), p1 = ld3(V, S[2 * i + 1]);
        b = seg_box(p0, p1);
        const v3 c = box_centroid(b);
        float ax = 0.0f, ay = 0.0f, ha = F_PI, rad = 0.0f;
        bool any = false, all_two = true;
        const uint32_t e0 = estart[i], e1 = estart[i + 1];
        float nn[2][2][2];
        for (uint32_t k = 0; k < 2; ++k) {
            if (e0 + k >= e1)
                break;
            const uint32_t e = e0 + k;
            const int4 s = slots[e];
            const bool has0 = s.z != -1, has1 = s.x != -1; // scene.cuh:322-325
            const v3 pv = ld3(V, s.y);
            float r0x = 0.0f, r0y = 0.0f, r1x = 0.0f, r1y = 0.0f, nx = 0.0f, ny = 0.0f;
            if (has0) { // normal(0): segment v -> next (scene.cuh:326-336)
                const v3 q = ld3(V, s.z);
                r0x = q.y - pv.y;
                r0y = -(q.x - pv.x);
                nx = nx + r0x;
                ny = ny + r0y;
            }
            if (has1) { // normal(1): segment prev -> v
                const v3 q = ld3(V, s.x);
                r1x = pv.y - q.y;
                r1y = -(pv.x - q.x);
                nx = nx + r1x;
                ny = ny + r1y;
            }
            float l = len2(nx, ny);
            nx = nx / l;
            ny = ny / l;
            ax += nx;
            ay += ny;
            rad = smax(rad, len2(pv.x - c.x, pv.y - c.y));
            all_two = all_two && has0 && has1;
            any = true;
            l = len2(r0x, r0y);
            nn[k][0][0] = has0 ? r0x / l : 0.0f;
            nn[k][0][1] = has0 ? r0y / l : 0.0f;
            l = len2(r1x, r1y);
            nn[k][1][0] = has1 ? r1x / l : 0.0f;
            nn[k][1][1] = has1 ? r1y / l : 0.0f;
            tedge[4 * e + 0] = make_float4(pv.x, pv.y, 0.0f, __int_as_float(s.y));
            tedge[4 * e + 1] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float((has0 ? 1u : 0u) | (has1 ? 2u : 0u)));
            tedge[4 * e + 2] = make_float4(nn[k][0][0], nn[k][0][1], 0.0f, 0.0f);
            tedge[4 * e + 3] = make_float4(nn[k][1][0], nn[k][1][1], 0.0f, 0.0f);
        }
        if (!any) {
            ha = -F_PI;
        } else if (!all_two) {
            ha = F_PI;
        } else {
            const float an = len2(ax, ay);
            if (an > F_EPS) {
                ax /= an;
                ay /= an;
                // one arccos of the smallest cosine (dm::acosf_ never
                // increases over [-1, 1]: the same bits as the largest angle)
                float cmin = 1.0f;
                for (uint32_t k = 0; k < 2; ++k) {
                    if (e0 + k >= e1)
                        break;
                    for (int jj = 0; jj < 2; ++jj)
                        cmin = smin(cmin, clamp11(dot2(ax, ay, nn[k][jj][0], nn[k][jj][1])));
                }
                ha = smax(0.0f, dm::acosf_(cmin));
            }
        }
        leafrec[3 * i] = make_float4(b.lo.x, b.lo.y, 0.0f, b.hi.x);
        leafrec[3 * i + 1] = make_float4(b.hi.y, 0.0f, ax, ay);
        leafrec[3 * i + 2] = make_float4(0.0f, ha, rad, 0.0f);
    }
    float v[4] = {b.lo.x, b.lo.y, b.hi.x, b.hi.y};
    const int slot_of[4] = {0, 1, 3, 4};
    reduce_scene_box<4>(v, mesh, i < nS, scene_bits, slot_of, part);
}

// bvh.cuh:269-277 and morton_code.cuh:85-92
// thread m < M: a mesh of one primitive has its leaf as root, which is its
// own refit subtree ("cut"); k_init_scene_box cleared the cut counter
__device__ __forceinline__ void mesh_of_one_root(const MeshDesc* __restrict__ desc, uint32_t M,
                                                 uint32_t* __restrict__ parent, uint32_t* __restrict__ cuts,
                                                 uint32_t cut_cap, uint32_t m)
{
    if (m >= M || desc[m].n != 1)
        return;
    parent[desc[m].nb] = INVALID;
    const uint32_t k = atomicAdd(cuts + cut_cap, 1u);
    if (k < cut_cap)
        cuts[k] = desc[m].nb;
}

__global__ __launch_bounds__(BT) void k_morton2d(const float4* __restrict__ leafrec, const int* __restrict__ bits_all,
                                                 uint32_t n, uint32_t* __restrict__ codes, uint32_t* __restrict__ idx,
                                                 const uint32_t* __restrict__ pmesh, const MeshDesc* __restrict__ desc,
                                                 uint32_t M, uint32_t* __restrict__ parent, uint32_t* __restrict__ cuts,
                                                 uint32_t cut_cap)
{
    const uint32_t i = blockIdx.x * BT + threadIdx.x;
    mesh_of_one_root(desc, M, parent, cuts, cut_cap, i);
    if (i >= n)
        return;
    const int* scene_bits = bits_all + 8 * mesh_of_prim(pmesh, i);
    const float lox = ord2f(scene_bits[0]), loy = ord2f(scene_bits[1]);
    const float hix = ord2f(scene_bits[3]), hiy = ord2f(scene_bits[4]);
    const float4 r0 = leafrec[3 * i], r1 = leafrec[3 * i + 1]; // lo.xy . hi.x | hi.y
    float px = (r0.w + r0.x) * 0.5f, py = (r1.x + r0.y) * 0.5f;
    px -= lox;
    py -= loy;
    px /= (hix - lox);
    py /= (hiy - loy);
    codes[i] = morton2(px, py);
    idx[i] = i;
}

// bvh.cuh:292-302 and morton_code.cuh:61-70
__global__ __launch_bounds__(BT) void k_morton(const float4* __restrict__ leafrec, const int* __restrict__ bits_all,
                                               uint32_t n, uint32_t* __restrict__ codes, uint32_t* __restrict__ idx,
                                               const uint32_t* __restrict__ pmesh, const MeshDesc* __restrict__ desc,
                                               uint32_t M, uint32_t* __restrict__ parent, uint32_t* __restrict__ cuts,
                                               uint32_t cut_cap)
{
    const uint32_t i = blockIdx.x * BT + threadIdx.x;
    mesh_of_one_root(desc, M, parent, cuts, cut_cap, i);
    if (i >= n)
        return;
    const int* scene_bits = bits_all + 8 * mesh_of_prim(pmesh, i);
    const float lox = ord2f(scene_bits[0]), loy = ord2f(scene_bits[1]), loz = ord2f(scene_bits[2]);
    const float hix = ord2f(scene_bits[3]), hiy = ord2f(scene_bits[4]), hiz = ord2f(scene_bits[5]);
    const float4 r0 = leafrec[3 * i], r1 = leafrec[3 * i + 1]; // lo.xyz hi.x | hi.yz
    box3 b;
    b.lo = mk(r0.x, r0.y, r0.z);
    b.hi = mk(r0.w, r1.x, r1.y);
    v3 p = box_centroid(b);
    p.x -= lox;
    p.y -= loy;
    p.z -= loz;
    p.x /= (hix - lox);
    p.y /= (hiy - loy);
    p.z /= (hiz - loz);
    codes[i] = morton3(p);
    idx[i] = i;
}

// The sorted codes around one k_karras block in LDS: positions [lo, lo + KWL)
// (the block's own nodes and KW on either side).  Most searches of a node
// stay within a few hundred positions of it; the rest read memory.
constexpr int KW = 512;
constexpr int KWL = 256 + 2 * KW;
struct CodeWindow {
    const uint32_t* lds;  // sorted codes of positions [lo, lo + KWL)
    const uint32_t* lidx; // their primitive ids (ties between equal codes)
    int lo;
    __device__ __forceinline__ uint32_t at(const uint32_t* __restrict__ codes_all, uint32_t pos) const
    {
        const uint32_t u = (uint32_t)((int)pos - lo);
        return u < (uint32_t)KWL ? lds[u] : codes_all[pos];
    }
    __device__ __forceinline__ uint32_t id(const uint32_t* __restrict__ idx_all, uint32_t pos) const
    {
        const uint32_t u = (uint32_t)((int)pos - lo);
        return u < (uint32_t)KWL ? lidx[u] : idx_all[pos];
    }
};

// delta of the virtual 64-bit key (code << 32 | local index) of the mesh's
// sorted positions a and b; codes_all and idx_all are scene-wide, idx holds
// global primitive ids
__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ codes_all, const uint32_t* __restrict__ idx_all,
                                      const CodeWindow& w, uint32_t off, uint32_t a, uint32_t b)
{
    const uint32_t ca = w.at(codes_all, off + a), cb = w.at(codes_all, off + b);
    if (ca != cb)
        return __clz(ca ^ cb);
    return 32 + __clz((w.id(idx_all, off + a) - off) ^ (w.id(idx_all, off + b) - off));
}

constexpr uint32_t RT = 1024;  // leaves of one refit subtree ("cut")
constexpr uint32_t RB = 512;   // threads of one k_refit workgroup (RT / RB leaves each)
constexpr uint32_t RM = 16384; // leaves of a mid set's subtree (the tops below it: k_refit_mid)
constexpr uint32_t MID_CAP = 1024; // members of a mid set (at most 960, see k_refit_mid)


// a list of at most two node ids in registers (an array indexed by its count
// would live in private memory)
struct List2 {
    uint32_t a = INVALID, b = INVALID, n = 0;
    __device__ __forceinline__ void push(bool yes, uint32_t v)
    {
        b = yes && n == 1u ? v : b;
        a = yes && n == 0u ? v : a;
        n += yes ? 1u : 0u;
    }
    __device__ __forceinline__ uint32_t at(uint32_t j) const { return j == 0u ? a : b; }
};

// bvh.cuh:109-229 (determine_range, find_split, construct_internal_nodes)
// for one internal node: children, leaf range and parent links; the cuts it
// roots (maximal subtrees of at most RT leaves below a node of more), the mid
// sets it roots (maximal subtrees of at most RM leaves below a node of more,
// when they hold more than RT), and its class: an upper (more than RM
// leaves: `top`) or a mid-set member (more than RT: `midtop`)
__device__ __forceinline__ void karras_node(const uint32_t* __restrict__ codes_all,
                                            const uint32_t* __restrict__ idx_all, const CodeWindow& w, uint32_t t,
                                            const MeshDesc* __restrict__ desc, uint32_t M,
                                            uint32_t* __restrict__ parent, uint2* __restrict__ child,
                                            uint2* __restrict__ range, List2& cut, uint32_t& top, List2& mid,
                                            uint32_t& midtop)
{
    const uint32_t m = mesh_search<false>(desc, M, t);
    const MeshDesc D = desc[m];
    const uint32_t i = t - (D.off - m), N = D.n, off = D.off, nb = D.nb;
    if (N < 2 || i >= N - 1)
        return; // not reached for consistent offsets
    uint32_t first, last;
    if (i == 0) {
        first = 0;
        last = N - 1;
        parent[nb] = INVALID;
    } else {
        const int L = kdelta(codes_all, idx_all, w, off, i, i - 1);
        const int R = kdelta(codes_all, idx_all, w, off, i, i + 1);
        const int d = (R > L) ? 1 : -1;
        const int dmin = L < R ? L : R;
        int lmax = 2;
        int itmp = (int)i + d * lmax;
        int delta = (0 <= itmp && itmp < (int)N) ? kdelta(codes_all, idx_all, w, off, i, (uint32_t)itmp) : -1;
        while (delta > dmin) {
            lmax <<= 1;
            itmp = (int)i + d * lmax;
            delta = (0 <= itmp && itmp < (int)N) ? kdelta(codes_all, idx_all, w, off, i, (uint32_t)itmp) : -1;
        }
        int l = 0;
        for (int s = lmax >> 1; s > 0; s >>= 1) {
            itmp = (int)i + (l + s) * d;
            delta = (0 <= itmp && itmp < (int)N) ? kdelta(codes_all, idx_all, w, off, i, (uint32_t)itmp) : -1;
            if (delta > dmin)
                l += s;
        }
        const uint32_t j = (uint32_t)((int)i + l * d);
        first = d < 0 ? j : i;
        last = d < 0 ? i : j;
    }
    // find_split; first != last always holds for internal nodes and the
    // virtual 64-bit keys are unique, so the equal-key branch of the
    // reference (bvh.cuh:175-178) cannot trigger
    const int dnode = kdelta(codes_all, idx_all, w, off, first, last);
    int split = (int)first;
    int stride = (int)(last - first);
    do {
        stride = (stride + 1) >> 1;
        const int mid = split + stride;
        if (mid < (int)last) {
            if (kdelta(codes_all, idx_all, w, off, first, (uint32_t)mid) > dnode)
                split = mid;
        }
    } while (stride > 1);
    const uint32_t g = (uint32_t)split;
    const uint32_t left = nb + (first == g ? N - 1 + g : g);
    const uint32_t right = nb + (last == g + 1 ? N - 1 + g + 1 : g + 1);
    const uint32_t self = nb + i;
    child[self] = make_uint2(left, right);
    range[self] = make_uint2(off + first, off + last);
    parent[left] = self;
    parent[right] = self;
    const uint32_t size = last - first + 1, lsize = g - first + 1, rsize = last - g;
    // cuts: left, right (children of at most RT leaves below a node of more),
    // self (a mesh root of at most RT); mid sets: self (a mesh root of at most
    // RM), left, right (children of at most RM below a node of more)
    const bool big = size > RM, medium = size > RT && !big;
    if (big)
        top = self;
    if (medium)
        midtop = self;
    cut.push(size > RT && lsize <= RT, left);
    cut.push(size > RT && rsize <= RT, right);
    cut.push(size <= RT && i == 0, self);
    mid.push(medium && i == 0, self);
    mid.push(big && lsize > RT && lsize <= RM, left);
    mid.push(big && rsize > RT && rsize <= RM, right);
}

// One thread per internal node of every mesh (internal-record slot t);
// nodes are written with global (mesh-major) ids.  Lists the refit subtrees
// ("cuts"), the uppers ("tops", flags[top] = its list slot), the mid-set
// roots ("mids", flags[root] = its slot, its member count cleared) and the
// mid-set members ("midtops").  mids[mid_cap] counts the roots and
// mids[mid_cap + 1 + c] the members of set c.
__global__ __launch_bounds__(BT) void k_karras(const uint32_t* __restrict__ codes_all,
                                               const uint32_t* __restrict__ idx_all, uint32_t n_internal,
                                               uint32_t n_all, const MeshDesc* __restrict__ desc, uint32_t M,
                                               uint32_t* __restrict__ parent, uint2* __restrict__ child,
                                               uint2* __restrict__ range, uint32_t* __restrict__ flags,
                                               uint32_t* __restrict__ cuts, uint32_t cut_cap,
                                               uint32_t* __restrict__ tops, uint32_t top_cap,
                                               uint32_t* __restrict__ mids, uint32_t mid_cap,
                                               uint32_t* __restrict__ midtops)
{
    static_assert(BT == 256, "the code window spans BT nodes");
    // cuts, tops, mids and midtops of this block: counts, then list offsets
    __shared__ uint32_t s_n[4], s_base[4];
    __shared__ uint32_t win[KWL], widx[KWL];
    if (threadIdx.x < 4)
        s_n[threadIdx.x] = 0u;
    // internal-record slot t of mesh m is the node at sorted position t + m
    const uint32_t t0 = blockIdx.x * BT;
    const int lo = (int)(t0 + mesh_search<false>(desc, M, t0)) - KW;
    // the codes and primitive ids around the block's nodes: sorted Morton
    // codes repeat often (54 % of neighbouring pairs at 8.4M triangles), and
    // every tie reads the ids
    for (int u = threadIdx.x; u < KWL; u += BT) {
        const int pos = lo + u;
        const bool in = pos >= 0 && pos < (int)n_all; // positions outside [0, n_all) are never read
        win[u] = in ? codes_all[pos] : 0u;
        widx[u] = in ? idx_all[pos] : 0u;
    }
    __syncthreads();
    const CodeWindow w{win, widx, lo};
    const uint32_t t = t0 + threadIdx.x;
    List2 cut, mid;
    uint32_t self = INVALID, midtop = INVALID; // self: set for a top
    if (t < n_internal)
        karras_node(codes_all, idx_all, w, t, desc, M, parent, child, range, cut, self, mid, midtop);
    const uint32_t n_cut = cut.n, n_mid = mid.n;
    // one list append per block and list (the counters are shared by all blocks)
    const uint32_t cut_k = n_cut ? atomicAdd(&s_n[0], n_cut) : 0u;
    const uint32_t top_k = self != INVALID ? atomicAdd(&s_n[1], 1u) : 0u;
    const uint32_t mid_k = n_mid ? atomicAdd(&s_n[2], n_mid) : 0u;
    const uint32_t mt_k = midtop != INVALID ? atomicAdd(&s_n[3], 1u) : 0u;
    __syncthreads();
    if (threadIdx.x < 4) {
        const uint32_t n = s_n[threadIdx.x];
        uint32_t* ctr = threadIdx.x == 0 ? cuts + cut_cap
                        : threadIdx.x == 1 ? tops + top_cap
                        : threadIdx.x == 2 ? mids + mid_cap
                                           : midtops + top_cap;
        s_base[threadIdx.x] = n ? atomicAdd(ctr, n) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < 2u; ++j) {
        const uint32_t k = s_base[2] + mid_k + j;
        if (j < n_mid && k < mid_cap) {
            mids[k] = mid.at(j);
            mids[mid_cap + 1 + k] = 0u;
            flags[mid.at(j)] = k;
        }
    }
    if (midtop != INVALID) {
        const uint32_t k = s_base[3] + mt_k;
        if (k < top_cap)
            midtops[k] = midtop;
    }
#pragma unroll
    for (uint32_t j = 0; j < 2u; ++j) {
        const uint32_t k = s_base[0] + cut_k + j;
        if (j < n_cut && k < cut_cap)
            cuts[k] = cut.at(j);
    }
    if (self != INVALID) {
        const uint32_t k = s_base[1] + top_k;
        if (k < top_cap)
            tops[k] = self;
        flags[self] = k;
    }
}




__device__ __forceinline__ box3 ldbox(const float4* nb, uint32_t n)
{
    const float4 a = nb[2 * n], c = nb[2 * n + 1];
    box3 b;
    b.lo = mk(a.x, a.y, a.z);
    b.hi = mk(c.x, c.y, c.z);
    return b;
}
__device__ __forceinline__ cone3 ldcone(const float4* nc, uint32_t n)
{
    const float4 a = nc[2 * n], c = nc[2 * n + 1];
    cone3 r;
    r.axis = mk(a.x, a.y, a.z);
    r.half_angle = a.w;
    r.radius = c.x;
    return r;
}

// Bottom-up refit in two launches, with no device-wide fence.
//
// k_karras lists the maximal subtrees of at most RT leaves ("cuts"); they are
// disjoint and cover every leaf of every mesh.  k_refit's workgroups take the
// cuts in turn (an atomic counter); a workgroup keeps the topology of its cut
// and the box and cone of its live nodes in LDS.  It merges level by level:
// nodes with two leaf children are ready at once, one with one leaf child
// starts with one arrival (an LDS arrival bit per internal node), the second
// arrival queues the parent, and each round merges the queued nodes with
// consecutive threads, so the merge math runs on full wavefronts; a barrier
// separates the rounds.  k_refit_top does the same for the nodes above the cuts ("tops",
// about 2N/RT of them, listed by k_karras) in one workgroup, with the boxes
// in LDS when they fit and in memory otherwise (and, for very large scenes,
// a climb from the cut roots with arrival counters in memory).  All of it
// is touched by one workgroup, so workgroup scope orders it, and the launch
// boundary publishes the first launch's results.

// ids of one mesh: global node id <-> local id / sorted position
struct MeshIds {
    uint32_t off, n, nb;
    __device__ __forceinline__ bool internal(uint32_t node) const { return node - nb < n - 1; }
    // sorted position of a leaf node, or Karras position of an internal node
    __device__ __forceinline__ uint32_t pos(uint32_t node) const
    {
        return internal(node) ? off + (node - nb) : off + (node - nb - (n - 1));
    }
    __device__ __forceinline__ uint32_t ref(uint32_t node) const
    {
        return internal(node) ? node - nb : LEAF_BIT | (node - nb - (n - 1));
    }
};

// leaf run of an internal child covering at most leaf_size leaves
__device__ __forceinline__ uint32_t run_of(const MeshIds& id, uint32_t node, const uint2* __restrict__ range,
                                           uint32_t leaf_size)
{
    if (leaf_size <= 1 || !id.internal(node))
        return INVALID;
    const uint2 rg = range[node];
    const uint32_t cnt = rg.y - rg.x + 1;
    return cnt <= leaf_size ? ((cnt - 1) << RUN_SHIFT) | (rg.x - id.off) : INVALID;
}

// the parent box and cone of two children (bvh.cuh:520-604), and the records
// of internal node p
template <int DIM>
__device__ __forceinline__ void refit_write(const MeshIds& id, uint32_t p, uint2 ch, const box3& lb, const box3& rb,
                                            const cone3& lc, const cone3& rc, int conservative,
                                            const uint2* __restrict__ range, uint32_t leaf_size, box3& pb,
                                            cone3& pc, float4* nbox, float4* ncone, float4* __restrict__ tnode,
                                            float4* __restrict__ tcone, bool own = true)
{
    pb = box_merge(lb, rb);
    pc = DIM == 3 ? cone_merge(lc, rc, box_centroid(lb), box_centroid(rb), box_centroid(pb), conservative != 0)
                  : cone_merge2d(lc, rc, box_centroid(lb), box_centroid(rb), box_centroid(pb), conservative != 0);
    if (own) { // the node's own box and cone in the reference layout (else they live in its parent's record)
        nbox[2 * p] = make_float4(pb.lo.x, pb.lo.y, pb.lo.z, 0.0f);
        nbox[2 * p + 1] = make_float4(pb.hi.x, pb.hi.y, pb.hi.z, 0.0f);
        ncone[2 * p] = make_float4(pc.axis.x, pc.axis.y, pc.axis.z, pc.half_angle);
        ncone[2 * p + 1] = make_float4(pc.radius, 0.0f, 0.0f, 0.0f);
    }
    const uint32_t r = p - id.off; // internal record: (off - m) + (p - nb)
    tnode[4 * r + 0] = make_float4(lb.lo.x, lb.lo.y, lb.lo.z, lb.hi.x);
    tnode[4 * r + 1] = make_float4(lb.hi.y, lb.hi.z, rb.lo.x, rb.lo.y);
    tnode[4 * r + 2] = make_float4(rb.lo.z, rb.hi.x, rb.hi.y, rb.hi.z);
    tnode[4 * r + 3] = make_float4(__uint_as_float(id.ref(ch.x)), __uint_as_float(id.ref(ch.y)),
                                   __uint_as_float(run_of(id, ch.x, range, leaf_size)),
                                   __uint_as_float(run_of(id, ch.y, range, leaf_size)));
    tcone[4 * r + 0] = make_float4(lc.axis.x, lc.axis.y, lc.axis.z, lc.half_angle);
    tcone[4 * r + 1] = make_float4(rc.axis.x, rc.axis.y, rc.axis.z, rc.half_angle);
    tcone[4 * r + 2] = make_float4(lc.radius, rc.radius, 0.0f, 0.0f);
    // tcone[4 r + 3] (sin, cos of the half angles) is filled by tcone_trig
    // once the merges are done: off the chain of dependent merges
}

// sin and cos of both children's half angles of internal record r, read
// back from the record's first two entries
__device__ __forceinline__ void tcone_trig(float4* __restrict__ tcone, uint32_t r)
{
    const float hl = tcone[4 * r + 0].w, hr = tcone[4 * r + 1].w;
    float ls, lcs, rs, rcs;
    dm::sincosf_(hl, &ls, &lcs);
    dm::sincosf_(hr, &rs, &rcs);
    tcone[4 * r + 3] = make_float4(ls, lcs, rs, rcs);
}

// box + cone of live nodes in LDS, structure of arrays over RT slots.  A node
// lives in the slot of its first leaf (range.x - first): a parent takes over
// its left child's slot once both children are merged, so live nodes (those
// whose parent is not merged yet) always hold distinct slots.
struct NodeLds {
    float v[11][RT];
    __device__ __forceinline__ void put(uint32_t s, const box3& b, const cone3& c)
    {
        v[0][s] = b.lo.x; v[1][s] = b.lo.y; v[2][s] = b.lo.z;
        v[3][s] = b.hi.x; v[4][s] = b.hi.y; v[5][s] = b.hi.z;
        v[6][s] = c.axis.x; v[7][s] = c.axis.y; v[8][s] = c.axis.z;
        v[9][s] = c.half_angle; v[10][s] = c.radius;
    }
    __device__ __forceinline__ void get(uint32_t s, box3& b, cone3& c) const
    {
        b.lo = mk(v[0][s], v[1][s], v[2][s]);
        b.hi = mk(v[3][s], v[4][s], v[5][s]);
        c.axis = mk(v[6][s], v[7][s], v[8][s]);
        c.half_angle = v[9][s];
        c.radius = v[10][s];
    }
};

// Work queue of the nodes to merge in one round.  Each node is queued once,
// so the queue is one array; round r merges queue[qs, qe) and appends the
// parents whose second child it merged at qe + n[r % 3].  Three counters
// rotate: the one a round appends to is cleared two rounds earlier, after
// every thread has read it (the barrier between rounds orders both).
struct RoundQueue {
    uint32_t n[3];
    __device__ __forceinline__ void push(uint16_t* q, uint32_t qe, uint32_t round, uint32_t item)
    {
        q[qe + atomicAdd(&n[round % 3], 1u)] = (uint16_t)item;
    }
};

constexpr uint16_t NO_PARENT = 0xFFFFu;

// parent p of a node of the cut [first, last], as a Karras position - first
// (NO_PARENT unless it is an internal node of the cut's range: never the case
// below the cut root of a Karras tree)
__device__ __forceinline__ uint16_t local_parent(const MeshIds& id, uint32_t p, uint32_t first, uint32_t last)
{
    return id.internal(p) && id.pos(p) - first <= last - first ? (uint16_t)(id.pos(p) - first) : NO_PARENT;
}

// A cut's internal node at local position k (Karras position first + k), in
// one LDS word: left and right child as local positions (bit 10: a leaf) and
// the node's slot (its first leaf - first).
__device__ __forceinline__ uint32_t cut_child_local(const MeshIds& id, uint32_t c, uint32_t first)
{
    return id.internal(c) ? id.pos(c) - first : (1u << 10) | (id.pos(c) - first);
}
__device__ __forceinline__ uint32_t cut_child_global(const MeshIds& id, uint32_t lc, uint32_t first)
{
    const uint32_t j = first + (lc & 0x3FFu) - id.off; // mesh-local position
    return lc & (1u << 10) ? id.nb + (id.n - 1) + j : id.nb + j;
}

// Members of the mid sets: each top of at most RM leaves climbs to the root
// of its set (the last ancestor of at most RM leaves; flags[root] = the set's
// slot, written by k_karras) and appends itself to the set's member list.
// Run by k_refit's threads (index k0, stride) before they take cuts; the
// lists are read by k_refit_mid, the next launch.
__device__ __forceinline__ void collect_mid_members(uint32_t k0, uint32_t stride,
                                                    const uint32_t* __restrict__ midtops, uint32_t midtop_cap,
                                                    const uint32_t* __restrict__ parent,
                                                    const uint2* __restrict__ range,
                                                    const uint32_t* __restrict__ flags, uint32_t* __restrict__ mids,
                                                    uint32_t mid_cap, uint32_t* __restrict__ members)
{
    const uint32_t n = min(midtops[midtop_cap], midtop_cap);
    if (k0 >= n)
        return;
    const uint32_t n_mid = min(mids[mid_cap], mid_cap);
    uint32_t* mcount = mids + mid_cap + 1;
    for (uint32_t k = k0; k < n; k += stride) {
        const uint32_t t = midtops[k];
        uint32_t node = t;
        for (uint32_t guard = 0; guard < 64; ++guard) {
            const uint32_t p = parent[node];
            if (p == INVALID || range[p].y - range[p].x + 1 > RM)
                break;
            node = p;
        }
        const uint32_t c = flags[node];
        if (c < n_mid) {
            const uint32_t j = atomicAdd(mcount + c, 1u);
            if (j < MID_CAP)
                members[c * MID_CAP + j] = t;
        }
    }
}

template <int DIM>
__global__ __launch_bounds__(RB) void k_refit(const float4* __restrict__ V, const int32_t* __restrict__ F,
                                              const uint32_t* __restrict__ estart,
                                              const uint32_t* __restrict__ sidx, const float4* __restrict__ leafrec,
                                              const uint32_t* __restrict__ parent,
                                              const uint2* __restrict__ child, const uint2* __restrict__ range,
                                              const MeshDesc* __restrict__ desc, uint32_t M, uint32_t N_all,
                                              int conservative, uint32_t leaf_size, float4* nbox, float4* ncone,
                                              float4* __restrict__ tnode, float4* __restrict__ tcone,
                                              float4* __restrict__ tleaf, uint32_t* __restrict__ cuts,
                                              uint32_t cut_cap, const uint32_t* __restrict__ midtops,
                                              uint32_t midtop_cap, uint32_t* __restrict__ mids, uint32_t mid_cap,
                                              uint32_t* __restrict__ members, const uint32_t* __restrict__ flags)
{
    static_assert(RT == 1024, "10-bit local positions");
    collect_mid_members(blockIdx.x * RB + threadIdx.x, gridDim.x * RB, midtops, midtop_cap, parent, range, flags,
                        mids, mid_cap, members);
    constexpr uint32_t LPT = RT / RB; // leaves per thread
    __shared__ NodeLds nd;
    __shared__ uint32_t lnode[RT];    // children and slot of the internal node at local position k
    __shared__ uint16_t lpar[RT];     // its parent (local position), or NO_PARENT for the cut root
    __shared__ uint32_t larr[RT / 32]; // first arrival at internal node k (bit k)
    __shared__ uint16_t queue[RT];
    __shared__ RoundQueue rq;
    __shared__ uint32_t s_cut;
    const uint32_t t = threadIdx.x;
    const uint32_t count = min(cuts[cut_cap], cut_cap);
    if (t < 3)
        rq.n[t] = 0u;
    if (t < RT / 32)
        larr[t] = 0u;
    for (;;) {
        // workgroups take cuts in turn (cut sizes vary)
        if (t == 0)
            s_cut = atomicAdd(cuts + cut_cap + 1, 1u);
        __syncthreads(); // also orders the cleared counters and arrival bits
        const uint32_t c = s_cut;
        if (c >= count)
            break;
        const uint32_t root = cuts[c];
        const uint32_t m = mesh_search<true>(desc, M, root);
        const MeshDesc D = desc[m];
        const MeshIds id{D.off, D.n, D.nb};
        uint32_t first, last;
        if (id.internal(root)) {
            const uint2 rg = range[root];
            first = rg.x;
            last = rg.y;
        } else {
            first = last = id.pos(root);
        }
        const bool ok = last - first < RT && last < N_all; // holds for every cut of a Karras tree
        // the cut's topology, one coalesced load per array: the internal
        // nodes of the cut are the Karras positions in [first, last] whose
        // range lies inside it.  A node with two leaf children is ready at
        // once; one with one leaf child has had its first arrival.
#pragma unroll
        for (uint32_t h = 0; h < LPT; ++h) {
            const uint32_t u = t + h * RB, j = first + u;
            if (ok && j <= last && j - D.off < D.n - 1) {
                const uint32_t ni = D.nb + (j - D.off);
                const uint2 rg = range[ni];
                if (rg.x >= first && rg.y <= last) {
                    const uint2 ch = child[ni];
                    const uint32_t lc = cut_child_local(id, ch.x, first), rc = cut_child_local(id, ch.y, first);
                    lnode[u] = lc | (rc << 11) | ((rg.x - first) << 22);
                    lpar[u] = ni == root ? NO_PARENT : local_parent(id, parent[ni], first, last);
                    const uint32_t leaves = (lc >> 10) + (rc >> 10);
                    if (leaves == 2u)
                        rq.push(queue, 0u, 0u, u);
                    else if (leaves == 1u)
                        atomicOr(&larr[u >> 5], 1u << (u & 31u));
                }
            }
        }
        bool bad = !ok;
#pragma unroll
        for (uint32_t h = 0; h < LPT; ++h) {
            const uint32_t u = t + h * RB, j = first + u;
            if (!(ok && j <= last))
                continue;
            const uint32_t orig = sidx[j];
            if (orig - D.off >= D.n) {
                bad = true; // not a permutation of the mesh's primitives: the cut stays unwritten
                continue;
            }
            const uint32_t node = D.nb + (D.n - 1) + (j - D.off);
            const float4 r0 = leafrec[3 * orig], r1 = leafrec[3 * orig + 1], r2 = leafrec[3 * orig + 2];
            const float4 b0 = make_float4(r0.x, r0.y, r0.z, 0.0f), b1 = make_float4(r0.w, r1.x, r1.y, 0.0f);
            const float4 c0 = make_float4(r1.z, r1.w, r2.x, r2.y), c1 = make_float4(r2.z, 0.0f, 0.0f, 0.0f);
            if (node == root) { // a cut of one leaf; other leaves live in their parents' records
                nbox[2 * node] = b0;
                nbox[2 * node + 1] = b1;
                ncone[2 * node] = c0;
                ncone[2 * node + 1] = c1;
            }
            {
                box3 b;
                b.lo = mk(b0.x, b0.y, b0.z);
                b.hi = mk(b1.x, b1.y, b1.z);
                cone3 cc;
                cc.axis = mk(c0.x, c0.y, c0.z);
                cc.half_angle = c0.w;
                cc.radius = c1.x;
                nd.put(u, b, cc);
            }
            const uint32_t es = estart[orig], ec = estart[orig + 1] - es;
            const uint32_t lo = orig - D.off, les = es - D.eoff;
            if (DIM == 3) {
                const float4 a = V[F[3 * orig]], b = V[F[3 * orig + 1]], d = V[F[3 * orig + 2]];
                tleaf[3 * j + 0] = make_float4(a.x, a.y, a.z, __uint_as_float(lo));
                tleaf[3 * j + 1] = make_float4(b.x, b.y, b.z, __uint_as_float(les));
                tleaf[3 * j + 2] = make_float4(d.x, d.y, d.z, __uint_as_float(ec));
            } else {
                const float4 a = V[F[2 * orig]], b = V[F[2 * orig + 1]];
                tleaf[3 * j + 0] = make_float4(a.x, a.y, 0.0f, __uint_as_float(lo));
                tleaf[3 * j + 1] = make_float4(b.x, b.y, 0.0f, __uint_as_float(les));
                tleaf[3 * j + 2] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(ec));
            }
        }
        // topology, leaf slots and the ready nodes
        uint32_t qs = 0, qe = __syncthreads_or(bad ? 1 : 0) ? 0u : rq.n[0];
        for (uint32_t round = 1; qs < qe && round < RT; ++round) {
            if (t == 0)
                rq.n[(round + 1) % 3] = 0u;
            for (uint32_t i = qs + t; i < qe; i += RB) {
                const uint32_t k = queue[i];
                const uint32_t wd = lnode[k];
                const uint32_t lc = wd & 0x7FFu, rc = (wd >> 11) & 0x7FFu;
                const uint32_t ls = wd >> 22, rs = rc & 0x3FFu; // the right child's first leaf is its position
                const uint2 ch = make_uint2(cut_child_global(id, lc, first), cut_child_global(id, rc, first));
                box3 lb, rb, pb;
                cone3 lc3, rc3, pc;
                nd.get(ls, lb, lc3);
                nd.get(rs, rb, rc3);
                const uint32_t p = D.nb + (first + k - D.off);
                refit_write<DIM>(id, p, ch, lb, rb, lc3, rc3, conservative, range, leaf_size, pb, pc, nbox, ncone,
                                 tnode, tcone, p == root);
                nd.put(ls, pb, pc);
                const uint32_t kp = lpar[k];
                if (kp != NO_PARENT && (atomicOr(&larr[kp >> 5], 1u << (kp & 31u)) >> (kp & 31u)) & 1u)
                    rq.push(queue, qe, round, kp);
            }
            __syncthreads();
            const uint32_t added = rq.n[round % 3];
            qs = qe;
            qe += added;
        }
        // every merged node is in queue[0, qe); the barrier above orders its
        // record's writes before these reads
        for (uint32_t i = t; i < qe; i += RB)
            tcone_trig(tcone, D.nb + (first + queue[i] - D.off) - id.off);
        // clear for the next cut (after the barrier that ends this one's rounds)
        if (t < 3)
            rq.n[t] = 0u;
        if (t < RT / 32)
            larr[t] = 0u;
        __syncthreads(); // the next cut reuses the LDS
    }
}

// Above the cuts, the tops are refit as "sets": a set is a group of tops
// merged by one workgroup, with the topology of its members in LDS.  A
// member's children are members (merged in an earlier round of the same
// workgroup) or nodes whose box and cone are already in memory (nbox/ncone:
// cut roots written by k_refit, mid-set roots written by k_refit_mid).
//
//   mid sets  the tops of one subtree of at most RM leaves whose parent has
//             more (or a mesh root of at most RM leaves): k_refit_mid, one
//             workgroup per set, many sets at once (k_refit lists the
//             members first);
//   uppers    the tops of more than RM leaves: k_refit_top, one workgroup.
//
// Size of a mid set: the common prefix of a node's 64-bit keys grows by at
// least one bit from a node to its children, so a root-to-leaf path holds at
// most 64 internal nodes; the members without member children own disjoint
// subtrees of more than RT leaves, at most RM / (RT + 1) = 15 of them, and
// every member lies on a path from the set root to one of them: at most
// 15 * 64 = 960 members, so MID_CAP members always fit.
constexpr uint32_t MB = 256;              // threads of one k_refit_mid workgroup
constexpr uint32_t NO_MEMBER = 0xFFFFu;
constexpr uint32_t SET_ROUNDS = 128;      // a set is at most 64 levels deep
static_assert(RM / (RT + 1) * 64 <= MID_CAP, "mid-set bound");

// LDS arrays of a set, `cap` entries each (box: 11 rows, boxes in LDS only)
struct SetLds {
    uint32_t* node;  // member's node id
    uint2* ch;       // its children's node ids
    uint32_t* cidx;  // children's member indices, left | right << 16 (NO_MEMBER: in memory)
    uint32_t* word;  // parent's member index << 16 | member children not merged yet
    uint32_t* mesh;  // its mesh
    uint16_t* queue;
    float* box;
    uint32_t cap;
    __device__ __forceinline__ void put(uint32_t k, const box3& b, const cone3& c) const
    {
        float* v = box + k;
        v[0 * cap] = b.lo.x; v[1 * cap] = b.lo.y; v[2 * cap] = b.lo.z;
        v[3 * cap] = b.hi.x; v[4 * cap] = b.hi.y; v[5 * cap] = b.hi.z;
        v[6 * cap] = c.axis.x; v[7 * cap] = c.axis.y; v[8 * cap] = c.axis.z;
        v[9 * cap] = c.half_angle; v[10 * cap] = c.radius;
    }
    __device__ __forceinline__ void get(uint32_t k, box3& b, cone3& c) const
    {
        const float* v = box + k;
        b.lo = mk(v[0 * cap], v[1 * cap], v[2 * cap]);
        b.hi = mk(v[3 * cap], v[4 * cap], v[5 * cap]);
        c.axis = mk(v[6 * cap], v[7 * cap], v[8 * cap]);
        c.half_angle = v[9 * cap];
        c.radius = v[10 * cap];
    }
};

__device__ __forceinline__ uint32_t range_size(const uint2* __restrict__ range, uint32_t node)
{
    const uint2 rg = range[node];
    return rg.y - rg.x + 1;
}

// Refit the n members of a set, listed in s.node with flags[node] = member
// index (written before the call and ordered by a barrier).  A child is a
// member when it is internal with more than lo_thr leaves; a parent when it
// is internal with at most hi_thr leaves.  NT threads; rq cleared.
template <int DIM, bool BOX_LDS, uint32_t NT>
__device__ void refit_set(const SetLds& s, uint32_t n, uint32_t lo_thr, uint32_t hi_thr, RoundQueue& rq,
                          const uint32_t* __restrict__ parent, const uint2* __restrict__ child,
                          const uint2* __restrict__ range, const uint32_t* flags, const MeshDesc* __restrict__ desc,
                          uint32_t M, int conservative, uint32_t leaf_size, float4* nbox, float4* ncone,
                          float4* __restrict__ tnode, float4* __restrict__ tcone)
{
    const uint32_t t = threadIdx.x;
    for (uint32_t k = t; k < n; k += NT) {
        const uint32_t p = s.node[k];
        const uint32_t m = mesh_search<true>(desc, M, p);
        const MeshDesc D = desc[m];
        const MeshIds id{D.off, D.n, D.nb};
        const uint2 ch = child[p];
        const bool lm = id.internal(ch.x) && range_size(range, ch.x) > lo_thr;
        const bool rm = id.internal(ch.y) && range_size(range, ch.y) > lo_thr;
        const uint32_t li = lm ? flags[ch.x] & 0xFFFFu : NO_MEMBER;
        const uint32_t ri = rm ? flags[ch.y] & 0xFFFFu : NO_MEMBER;
        const uint32_t par = parent[p];
        const bool pm = par != INVALID && id.internal(par) && range_size(range, par) <= hi_thr;
        s.ch[k] = ch;
        s.cidx[k] = li | (ri << 16);
        s.word[k] = ((pm ? flags[par] & 0xFFFFu : NO_MEMBER) << 16) | ((lm ? 1u : 0u) + (rm ? 1u : 0u));
        s.mesh[k] = m;
    }
    __syncthreads();
    // each member is queued once (when its last member child is merged), so
    // the queue holds at most n entries; the bounds below only guard LDS
    // against an inconsistent tree
    auto push = [&](uint32_t qe, uint32_t round, uint32_t k) {
        const uint32_t at = qe + atomicAdd(&rq.n[round % 3], 1u);
        if (at < n)
            s.queue[at] = (uint16_t)k;
    };
    for (uint32_t k = t; k < n; k += NT)
        if ((s.word[k] & 0xFFFFu) == 0u)
            push(0u, 0u, k);
    __syncthreads();
    uint32_t qs = 0, qe = min(rq.n[0], n);
    for (uint32_t round = 1; qs < qe && round < SET_ROUNDS; ++round) {
        if (t == 0)
            rq.n[(round + 1) % 3] = 0u;
        for (uint32_t i = qs + t; i < qe; i += NT) {
            const uint32_t k = s.queue[i];
            const uint32_t p = s.node[k];
            const uint2 ch = s.ch[k];
            const uint32_t ci = s.cidx[k];
            const MeshDesc D = desc[min(s.mesh[k], M - 1)];
            const MeshIds id{D.off, D.n, D.nb};
            box3 lb, rb, pb;
            cone3 lc, rc, pc;
            if (BOX_LDS && (ci & 0xFFFFu) != NO_MEMBER) {
                s.get(min(ci & 0xFFFFu, s.cap - 1), lb, lc);
            } else {
                lb = ldbox(nbox, ch.x);
                lc = ldcone(ncone, ch.x);
            }
            if (BOX_LDS && (ci >> 16) != NO_MEMBER) {
                s.get(min(ci >> 16, s.cap - 1), rb, rc);
            } else {
                rb = ldbox(nbox, ch.y);
                rc = ldcone(ncone, ch.y);
            }
            refit_write<DIM>(id, p, ch, lb, rb, lc, rc, conservative, range, leaf_size, pb, pc, nbox, ncone, tnode,
                             tcone);
            if (BOX_LDS)
                s.put(k, pb, pc);
            const uint32_t pk = s.word[k] >> 16;
            if (pk < n && (atomicSub(&s.word[pk], 1u) & 0xFFFFu) == 1u)
                push(qe, round, pk);
        }
        __syncthreads(); // this round's boxes (LDS or memory) precede the next round's reads
        const uint32_t added = rq.n[round % 3];
        qs = qe;
        qe = min(qe + added, n);
    }
    // the sin and cos of the children's half angles, once every record is written
    for (uint32_t k = t; k < n; k += NT)
        tcone_trig(tcone, s.node[k] - desc[min(s.mesh[k], M - 1)].off);
}

// One workgroup per mid set (a grid-stride loop over the sets), boxes in LDS.
template <int DIM>
__global__ __launch_bounds__(MB) void k_refit_mid(const uint32_t* __restrict__ parent,
                                                  const uint2* __restrict__ child, const uint2* __restrict__ range,
                                                  uint32_t* flags, const uint32_t* __restrict__ mids,
                                                  uint32_t mid_cap, const uint32_t* __restrict__ members,
                                                  const MeshDesc* __restrict__ desc, uint32_t M, int conservative,
                                                  uint32_t leaf_size, float4* nbox, float4* ncone,
                                                  float4* __restrict__ tnode, float4* __restrict__ tcone)
{
    __shared__ uint32_t node[MID_CAP], cidx[MID_CAP], word[MID_CAP], mesh[MID_CAP];
    __shared__ uint2 ch[MID_CAP];
    __shared__ uint16_t queue[MID_CAP];
    __shared__ float box[11 * MID_CAP];
    __shared__ RoundQueue rq;
    const SetLds s{node, ch, cidx, word, mesh, queue, box, MID_CAP};
    const uint32_t t = threadIdx.x;
    const uint32_t n_mid = min(mids[mid_cap], mid_cap);
    const uint32_t* mcount = mids + mid_cap + 1;
    for (uint32_t c = blockIdx.x; c < n_mid; c += gridDim.x) {
        const uint32_t n = min(mcount[c], MID_CAP);
        if (t < 3)
            rq.n[t] = 0u;
        for (uint32_t k = t; k < n; k += MB) {
            const uint32_t p = members[c * MID_CAP + k];
            node[k] = p;
            flags[p] = k; // this workgroup's members only: read back after the barrier
        }
        __syncthreads();
        refit_set<DIM, true, MB>(s, n, RT, RM, rq, parent, child, range, flags, desc, M, conservative, leaf_size,
                                 nbox, ncone, tnode, tcone);
        __syncthreads(); // the next set reuses the LDS
    }
}

// The uppers (tops of more than RM leaves, flags[top] = its list slot from
// k_karras) in one workgroup: boxes in LDS up to TOP_BOX members, topology
// alone in LDS up to TOP_TOPO (boxes read back from memory), beyond that a
// climb with arrival counters in memory.  force: 1 or 2 picks the second or
// third layout whatever the count (they give the same records).
constexpr uint32_t TOP_BOX = 2048;
constexpr uint32_t TOP_TOPO = 6016;
constexpr uint32_t TOP_SMEM = TOP_TOPO * 26;
static_assert(TOP_BOX * (26 + 44) <= TOP_SMEM, "LDS layouts of k_refit_top");

template <int DIM>
__global__ __launch_bounds__(1024) void k_refit_top(const uint32_t* __restrict__ parent,
                                                    const uint2* __restrict__ child, const uint2* __restrict__ range,
                                                    uint32_t* flags, const uint32_t* __restrict__ tops,
                                                    uint32_t top_cap, const MeshDesc* __restrict__ desc, uint32_t M,
                                                    int conservative, uint32_t leaf_size, float4* nbox, float4* ncone,
                                                    float4* __restrict__ tnode, float4* __restrict__ tcone,
                                                    uint32_t force)
{
    __shared__ __align__(16) unsigned char smem[TOP_SMEM];
    __shared__ RoundQueue rq;
    const uint32_t t = threadIdx.x;
    const uint32_t n = min(tops[top_cap], top_cap);
    if (t < 3)
        rq.n[t] = 0u;
    auto layout = [&](uint32_t cap, bool boxes) {
        // 8-byte children first, then 4-byte arrays, the boxes, the 2-byte queue
        SetLds s;
        s.cap = cap;
        s.ch = reinterpret_cast<uint2*>(smem);
        s.node = reinterpret_cast<uint32_t*>(smem + 8 * cap);
        s.cidx = s.node + cap;
        s.word = s.cidx + cap;
        s.mesh = s.word + cap;
        s.box = reinterpret_cast<float*>(s.mesh + cap);
        s.queue = reinterpret_cast<uint16_t*>(s.box + (boxes ? 11 * cap : 0));
        return s;
    };
    if ((n <= TOP_BOX && force == 0) || (n <= TOP_TOPO && force <= 1)) {
        const bool boxes = n <= TOP_BOX && force == 0;
        const SetLds s = layout(boxes ? TOP_BOX : TOP_TOPO, boxes);
        for (uint32_t k = t; k < n; k += 1024)
            s.node[k] = tops[k];
        __syncthreads();
        if (boxes)
            refit_set<DIM, true, 1024>(s, n, RM, 0xFFFFFFFFu, rq, parent, child, range, flags, desc, M,
                                       conservative, leaf_size, nbox, ncone, tnode, tcone);
        else
            refit_set<DIM, false, 1024>(s, n, RM, 0xFFFFFFFFu, rq, parent, child, range, flags, desc, M,
                                        conservative, leaf_size, nbox, ncone, tnode, tcone);
        return;
    }
    // climb: flags[top] = its upper children not merged yet; a top with none
    // starts a climb, and the thread that merges a top's last child goes on
    // to merge the top
    auto upper_children = [&](uint32_t p, const MeshIds& id) {
        const uint2 ch = child[p];
        return (id.internal(ch.x) && range_size(range, ch.x) > RM ? 1u : 0u) +
               (id.internal(ch.y) && range_size(range, ch.y) > RM ? 1u : 0u);
    };
    for (uint32_t k = t; k < n; k += 1024) {
        const uint32_t p = tops[k];
        const MeshDesc D = desc[mesh_search<true>(desc, M, p)];
        flags[p] = upper_children(p, MeshIds{D.off, D.n, D.nb});
    }
    __syncthreads();
    for (uint32_t k = t; k < n; k += 1024) {
        uint32_t p = tops[k];
        const MeshDesc D = desc[mesh_search<true>(desc, M, p)];
        const MeshIds id{D.off, D.n, D.nb};
        if (upper_children(p, id) != 0u)
            continue;
        for (uint32_t guard = 0; guard < SET_ROUNDS; ++guard) {
            const uint2 ch = child[p];
            box3 pb;
            cone3 pc;
            refit_write<DIM>(id, p, ch, ldbox(nbox, ch.x), ldbox(nbox, ch.y), ldcone(ncone, ch.x),
                             ldcone(ncone, ch.y), conservative, range, leaf_size, pb, pc, nbox, ncone, tnode, tcone);
            const uint32_t par = parent[p];
            if (par == INVALID || !id.internal(par))
                break;
            // release this top's records, acquire the sibling's
            if (__hip_atomic_fetch_add(flags + par, 0xFFFFFFFFu, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) != 1u)
                break;
            p = par;
        }
    }
    __syncthreads(); // the records' writes precede their reads
    for (uint32_t k = t; k < n; k += 1024) {
        const uint32_t p = tops[k];
        tcone_trig(tcone, p - desc[mesh_search<true>(desc, M, p)].off);
    }
}


inline uint32_t blocks(uint32_t n) { return (n + BT - 1) / BT; }

#define SNCH_DISPATCH_REFIT(dim, ...) \
    do {                              \
        if ((dim) == 3) {             \
            constexpr int DIM = 3;    \
            __VA_ARGS__;              \
        } else {                      \
            constexpr int DIM = 2;    \
            __VA_ARGS__;              \
        }                             \
    } while (0)

} // namespace

hipError_t build_tree(snch_scene* s, uint32_t flags, uint32_t leaf_size, hipStream_t st, hipEvent_t* phase)
{
    auto mark = [&](int k) {
        if (phase)
            (void)hipEventRecord(phase[k], st);
    };
    const uint32_t N = s->nF, M = s->n_meshes, NN = 2 * N - M, NI = N - M;
    const uint32_t E = s->dim == 3 ? s->nE : s->n_internal_sil;
    const int conservative = (flags & SNCH_BUILD_REFERENCE_RADIUS) ? 0 : 1;
    const MeshDesc* desc = s->dMesh.as<MeshDesc>();
    const uint32_t* pmesh = M > 1 ? s->dPmesh.as<uint32_t>() : nullptr;
    BuildScratch& w = s->scratch;
    hipError_t e;
#define SNCH_TRY(x)              \
    do {                         \
        if ((e = (x)) != hipSuccess) \
            return e;            \
    } while (0)
    SNCH_TRY(w.leafbox.ensure(sizeof(float4) * 3 * N)); // the leaf records (box + cone)
    SNCH_TRY(w.scene_bits.ensure(sizeof(int) * 8 * M));
    SNCH_TRY(w.boxpart.ensure(sizeof(float) * 8 * blocks(N)));
    SNCH_TRY(w.codes.ensure(sizeof(uint32_t) * N));
    SNCH_TRY(w.idx.ensure(sizeof(uint32_t) * N));
    SNCH_TRY(w.codes_alt.ensure(sizeof(uint32_t) * N));
    SNCH_TRY(w.idx_alt.ensure(sizeof(uint32_t) * N));
    SNCH_TRY(w.parent.ensure(sizeof(uint32_t) * NN));
    SNCH_TRY(w.child.ensure(sizeof(uint2) * NN));
    SNCH_TRY(w.range.ensure(sizeof(uint2) * NN));
    SNCH_TRY(w.flags.ensure(sizeof(uint32_t) * NN));
    const uint32_t cut_cap = N; // cuts are disjoint subtrees of at least one leaf
    SNCH_TRY(w.top.ensure(sizeof(uint32_t) * (cut_cap + 2))); // cuts, their count, k_refit's next cut
    uint32_t* cuts = w.top.as<uint32_t>();
    const uint32_t top_cap = NI > 0 ? NI : 1u; // tops are internal nodes
    SNCH_TRY(w.tops.ensure(sizeof(uint32_t) * (top_cap + 1)));
    uint32_t* tops = w.tops.as<uint32_t>();
    SNCH_TRY(w.midtops.ensure(sizeof(uint32_t) * (top_cap + 1)));
    uint32_t* midtops = w.midtops.as<uint32_t>();
    // mid sets are disjoint subtrees of more than RT leaves
    const uint32_t mid_cap = N / (RT + 1) + 1;
    SNCH_TRY(w.mids.ensure(sizeof(uint32_t) * (2 * mid_cap + 1)));
    uint32_t* mids = w.mids.as<uint32_t>();
    SNCH_TRY(w.members.ensure(sizeof(uint32_t) * mid_cap * MID_CAP));
    SNCH_TRY(s->tnode.ensure(sizeof(float4) * 4 * (NI > 0 ? NI : 1)));
    SNCH_TRY(s->tcone.ensure(sizeof(float4) * 4 * (NI > 0 ? NI : 1)));
    SNCH_TRY(s->tleaf.ensure(sizeof(float4) * 3 * N));
    SNCH_TRY(s->tedge.ensure(sizeof(float4) * 4 * (E > 0 ? E : 1)));
    SNCH_TRY(s->nbox.ensure(sizeof(float4) * 2 * NN));
    SNCH_TRY(s->ncone.ensure(sizeof(float4) * 2 * NN));

    // stable sort by (mesh, code, index): code passes, then mesh-id passes;
    // k_init_scene_box clears the sort's counters
    int mesh_bits = 0;
    while (mesh_bits < 32 && ((M - 1) >> mesh_bits) != 0)
        ++mesh_bits;
    SNCH_TRY(w.hist.ensure(radix_sort_scratch_bytes(N, 4 + (pmesh ? (mesh_bits + 7) / 8 : 0))));
    mark(0);
    hipLaunchKernelGGL(k_init_scene_box, dim3((8 * M + 255) / 256), dim3(256), 0, st, w.scene_bits.as<int>(), M,
                       cuts + cut_cap, tops + top_cap, mids + mid_cap, midtops + top_cap, w.hist.as<uint32_t>(),
                       (uint32_t)radix_sort_clear_words());
    if (s->dim == 3) {
        hipLaunchKernelGGL(k_leafgen, dim3(blocks(N)), dim3(BT), 0, st, s->dV.as<float4>(), s->dF.as<int32_t>(),
                           s->dSlots.as<int4>(), s->dEstart.as<uint32_t>(), N, conservative, w.leafbox.as<float4>(),
                           s->tedge.as<float4>(), w.scene_bits.as<int>(), pmesh, desc,
                           w.boxpart.as<float>());
        hipLaunchKernelGGL(k_box_final, dim3((blocks(N) + BF - 1) / BF), dim3(BF), 0, st, w.boxpart.as<float>(),
                           blocks(N), w.scene_bits.as<int>());
        hipLaunchKernelGGL(k_morton, dim3(blocks(N)), dim3(BT), 0, st, w.leafbox.as<float4>(),
                           w.scene_bits.as<int>(), N, w.codes.as<uint32_t>(), w.idx.as<uint32_t>(), pmesh, desc, M,
                           w.parent.as<uint32_t>(), cuts, cut_cap);
    } else {
        hipLaunchKernelGGL(k_leafgen2d, dim3(blocks(N)), dim3(BT), 0, st, s->dV.as<float4>(), s->dF.as<int32_t>(),
                           s->dSlots.as<int4>(), s->dEstart.as<uint32_t>(), N, w.leafbox.as<float4>(),
                           s->tedge.as<float4>(), w.scene_bits.as<int>(), pmesh,
                           w.boxpart.as<float>());
        hipLaunchKernelGGL(k_box_final, dim3((blocks(N) + BF - 1) / BF), dim3(BF), 0, st, w.boxpart.as<float>(),
                           blocks(N), w.scene_bits.as<int>());
        hipLaunchKernelGGL(k_morton2d, dim3(blocks(N)), dim3(BT), 0, st, w.leafbox.as<float4>(),
                           w.scene_bits.as<int>(), N, w.codes.as<uint32_t>(), w.idx.as<uint32_t>(), pmesh, desc, M,
                           w.parent.as<uint32_t>(), cuts, cut_cap);
    }
    mark(1);
    bool in_alt = false;
    SNCH_TRY(radix_sort_pairs(w.codes.as<uint32_t>(), w.idx.as<uint32_t>(), w.codes_alt.as<uint32_t>(),
                              w.idx_alt.as<uint32_t>(), N, 30, w.hist, st, &in_alt, pmesh, pmesh ? mesh_bits : 0,
                              true));
    const uint32_t* codes = in_alt ? w.codes_alt.as<uint32_t>() : w.codes.as<uint32_t>();
    const uint32_t* sidx = in_alt ? w.idx_alt.as<uint32_t>() : w.idx.as<uint32_t>();
    mark(2);
    if (NI > 0)
        hipLaunchKernelGGL(k_karras, dim3(blocks(NI)), dim3(BT), 0, st, codes, sidx, NI, N, desc, M,
                           w.parent.as<uint32_t>(), w.child.as<uint2>(), w.range.as<uint2>(), w.flags.as<uint32_t>(),
                           cuts, cut_cap, tops, top_cap, mids, mid_cap, midtops);
    // about 2N / RT cuts (each holds at least RT / 2 leaves but for one per
    // side of a large node, plus one per small mesh); workgroups loop over them
    const uint32_t n_cu = s->ctx && s->ctx->n_cu > 0 ? (uint32_t)s->ctx->n_cu : 256u;
    bool any_big = false, any_huge = false; // a mesh of more than RT (RM) leaves
    for (uint32_t m = 0; m < M; ++m) {
        any_big = any_big || s->mesh_off[m + 1] - s->mesh_off[m] > RT;
        any_huge = any_huge || s->mesh_off[m + 1] - s->mesh_off[m] > RM;
    }
    const uint32_t top_layout = (flags & SNCH_BUILD_TOPS_CLIMB) ? 2u : (flags & SNCH_BUILD_TOPS_IN_MEMORY) ? 1u : 0u;
    mark(3);
    SNCH_DISPATCH_REFIT(s->dim, {
        // as many workgroups as are resident at once (LDS-bound), or one per cut
        static int per_cu = 0;
        if (per_cu == 0 &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_refit<DIM>, (int)RB, 0) != hipSuccess)
            per_cu = 1;
        const uint32_t refit_blocks =
            std::max(1u, std::min(2u * ((N + RT - 1) / RT) + M, (uint32_t)std::max(per_cu, 1) * n_cu));
        hipLaunchKernelGGL(k_refit<DIM>, dim3(refit_blocks), dim3(RB), 0, st, s->dV.as<float4>(),
                           s->dF.as<int32_t>(), s->dEstart.as<uint32_t>(), sidx, w.leafbox.as<float4>(),
                           w.parent.as<uint32_t>(), w.child.as<uint2>(), w.range.as<uint2>(),
                           desc, M, N, conservative, leaf_size, s->nbox.as<float4>(), s->ncone.as<float4>(),
                           s->tnode.as<float4>(), s->tcone.as<float4>(), s->tleaf.as<float4>(), cuts, cut_cap,
                           midtops, top_cap, mids, mid_cap, w.members.as<uint32_t>(), w.flags.as<uint32_t>());
        if (any_big)
            hipLaunchKernelGGL(k_refit_mid<DIM>, dim3(std::min(mid_cap, 2u * n_cu)), dim3(MB), 0, st,
                               w.parent.as<uint32_t>(), w.child.as<uint2>(), w.range.as<uint2>(),
                               w.flags.as<uint32_t>(), mids, mid_cap, w.members.as<uint32_t>(), desc, M,
                               conservative, leaf_size, s->nbox.as<float4>(), s->ncone.as<float4>(),
                               s->tnode.as<float4>(), s->tcone.as<float4>());
        if (any_huge)
            hipLaunchKernelGGL(k_refit_top<DIM>, dim3(1), dim3(1024), 0, st, w.parent.as<uint32_t>(),
                               w.child.as<uint2>(), w.range.as<uint2>(), w.flags.as<uint32_t>(), tops, top_cap,
                               desc, M, conservative, leaf_size, s->nbox.as<float4>(), s->ncone.as<float4>(),
                               s->tnode.as<float4>(), s->tcone.as<float4>(), top_layout);
    });
    mark(4);
    s->sorted_in_alt = in_alt;
#undef SNCH_TRY
    return hipGetLastError();
}

} // namespace snch
