// Per-lane building blocks of the traversal queries (node records, leaf
// tests, the sampler's descent) and plain per-thread traversals for users'
// own HIP kernels (include/snch_lbvh/device.hpp).  Header-only, device code.
#pragma once
#include <hip/hip_runtime.h>

#include "geom.hpp"
#include "tree_view.hpp"

namespace snch {

// per-query visit counters (summed per wave by the library's kernels)
struct Counts {
    uint32_t box = 0, leaf = 0, edge = 0, node = 0, cone = 0;
};

struct NodeRec {
    box3 L, R;
    uint32_t lref, rref, lrun, rrun;
};

__device__ __forceinline__ NodeRec load_node(const float4* __restrict__ tnode, uint32_t i)
{
    const float4 a = tnode[4 * i], b = tnode[4 * i + 1], c = tnode[4 * i + 2], r = tnode[4 * i + 3];
    NodeRec n;
    n.L.lo = mk(a.x, a.y, a.z);
    n.L.hi = mk(a.w, b.x, b.y);
    n.R.lo = mk(b.z, b.w, c.x);
    n.R.hi = mk(c.y, c.z, c.w);
    n.lref = __float_as_uint(r.x);
    n.rref = __float_as_uint(r.y);
    n.lrun = __float_as_uint(r.z);
    n.rrun = __float_as_uint(r.w);
    return n;
}

// a child the traversal treats as a leaf: a true leaf, or a run of sorted
// leaves [first, first + count) under an internal node (leaf size > 1)
__device__ __forceinline__ bool as_leaf(uint32_t ref, uint32_t run, uint32_t& first, uint32_t& count)
{
    if (ref & LEAF_BIT) {
        first = ref & ~LEAF_BIT;
        count = 1;
        return true;
    }
    if (run != INVALID) {
        first = run & RUN_FIRST_MASK;
        count = (run >> RUN_SHIFT) + 1;
        return true;
    }
    return false;
}

__device__ __forceinline__ box3 ldbox2(const float4* nb, uint32_t n)
{
    const float4 a = nb[2 * n], c = nb[2 * n + 1];
    box3 b;
    b.lo = mk(a.x, a.y, a.z);
    b.hi = mk(c.x, c.y, c.z);
    return b;
}

template <int DIM>
__device__ __forceinline__ v3 load_point(const float* __restrict__ q, uint32_t i)
{
    return DIM == 3 ? mk(q[3 * i], q[3 * i + 1], q[3 * i + 2]) : mk(q[2 * i], q[2 * i + 1], 0.0f);
}

template <int DIM>
__device__ __forceinline__ void cp_leaf(const TreeView& t, uint32_t j, v3 p, float& best, uint32_t& best_idx)
{
    if (j >= t.n_leaves)
        return;
    const float4 a = t.tleaf[3 * j], b = t.tleaf[3 * j + 1];
    float d;
    if (DIM == 3) {
        const float4 c = t.tleaf[3 * j + 2];
        d = closest_point_triangle(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(c.x, c.y, c.z), p);
    } else {
        d = closest_point_segment2d(mk(a.x, a.y, 0.0f), mk(b.x, b.y, 0.0f), p); // scene.cuh:504-515
    }
    d *= d; // query.cuh:284-285
    const uint32_t o = __float_as_uint(a.w);
    if (d < best || (d == best && o < best_idx)) {
        best = d;
        best_idx = o;
    }
}

// one leaf: its owned silhouettes (edges in 3D, scene.cuh:978-1003 and
// :788-824; vertices in 2D, scene.cuh:517-541 and :354-386).  The record's
// w holds the silhouette id reported to the caller.
template <int DIM>
__device__ __forceinline__ void sil_leaf(const TreeView& t, uint32_t j, v3 p, bool flip, float& best,
                                         uint32_t& best_e, Counts& cnt)
{
    if (j >= t.n_leaves)
        return;
    const uint32_t es = __float_as_uint(t.tleaf[3 * j + 1].w);
    const uint32_t ec = __float_as_uint(t.tleaf[3 * j + 2].w);
    cnt.leaf += 1;
    cnt.edge += ec;
    for (uint32_t k = 0; k < ec; ++k) {
        const uint32_t e = es + k;
        const float mr2 = best * best;
        if (!(0.0f < mr2))
            break; // min_radius_squared (0) >= max_radius_squared
        const float4 A = t.tedge[4 * e], B = t.tedge[4 * e + 1];
        const uint32_t id = __float_as_uint(A.w);
        const uint32_t fl = __float_as_uint(B.w);
        float d;
        bool sil = (fl & 3u) != 3u; // a boundary edge / end vertex is always a silhouette
        if (DIM == 3) {
            const v3 pa = mk(A.x, A.y, A.z), pb = mk(B.x, B.y, B.z);
            v3 cp;
            d = closest_point_segment(pa, pb, p, cp);
            if (d * d > mr2)
                continue;
            if (!sil) {
                const float4 N0 = t.tedge[4 * e + 2], N1 = t.tedge[4 * e + 3];
                sil = is_silhouette_edge(pa, pb, mk(N0.x, N0.y, N0.z), mk(N1.x, N1.y, N1.z),
                                         mk(p.x - cp.x, p.y - cp.y, p.z - cp.z), d, flip);
            }
        } else {
            const float vx = p.x - A.x, vy = p.y - A.y;
            d = len2(vx, vy);
            if (d * d > mr2)
                continue;
            if (!sil) {
                const float4 N0 = t.tedge[4 * e + 2], N1 = t.tedge[4 * e + 3];
                sil = is_silhouette_vertex(mk(N0.x, N0.y, 0.0f), mk(N1.x, N1.y, 0.0f), vx, vy, d, flip);
            }
        }
        if (sil && (d < best || (d == best && id < best_e))) {
            best = d;
            best_e = id;
        }
    }
}

template <int DIM>
__device__ __forceinline__ bool overlap_dim(const float4& C, float radius, v3 p, const box3& b, float md)
{
    return DIM == 3 ? cone_overlap(mk(C.x, C.y, C.z), C.w, radius, p, b, md)
                    : cone_overlap2d(mk(C.x, C.y, 0.0f), C.w, radius, p, b, md);
}

// SampleTriangleInSphere for one sphere (sample.cuh:23-92): the reference's
// single descent with its FP32 arithmetic (u0 rescaled at each level), then
// the sampled point from u1, u2.  obj = -1 and pdf = 0 when nothing is hit.
template <int DIM>
__device__ __forceinline__ void sample_one(const TreeView& t, const float4* __restrict__ V,
                                           const int32_t* __restrict__ F, v3 c, float r, float u, float u1,
                                           float u2, int32_t& obj, float& pdf, v3& pt)
{
    float npdf = 1.0f;
    obj = -1;
    pdf = 0.0f;
    uint32_t ref = t.n_leaves == 1 ? LEAF_BIT : 0u;
    for (int depth = 0; depth < 96; ++depth) { // a single descent: at most the tree depth
        if (ref & LEAF_BIT) {
            const uint32_t j = ref & ~LEAF_BIT;
            if (j >= t.n_leaves)
                break;
            const float4 a = t.tleaf[3 * j], b = t.tleaf[3 * j + 1], d = t.tleaf[3 * j + 2];
            const v3 pa = mk(a.x, a.y, a.z), pb = mk(b.x, b.y, b.z), pc = mk(d.x, d.y, d.z);
            if (DIM == 3 ? tri_sphere(pa, pb, pc, c, r) : sphere_segment(pa, pb, c, r)) {
                obj = (int32_t)__float_as_uint(a.w);
                pdf = npdf;
                pdf /= DIM == 3 ? tri_area(pa, pb, pc) : len2(pa.x - pb.x, pa.y - pb.y); // scene.cuh:409-418
            }
            break;
        }
        if (ref >= t.n_leaves - 1)
            break;
        const NodeRec nr = load_node(t.tnode, ref);
        const float Lw = box_sphere(nr.L, c, r)
                             ? (DIM == 3 ? green_weight(c, box_centroid(nr.L)) : green_weight2d(c, box_centroid(nr.L)))
                             : 0.0f;
        const float Rw = box_sphere(nr.R, c, r)
                             ? (DIM == 3 ? green_weight(c, box_centroid(nr.R)) : green_weight2d(c, box_centroid(nr.R)))
                             : 0.0f;
        const float tw = Lw + Rw;
        if (!(tw > 0.0f))
            break;
        const float Lp = Lw / tw;
        if (u < Lp) {
            u /= Lp;
            npdf = Lp * npdf;
            ref = nr.lref;
        } else {
            const float Rp = 1.0f - Lp;
            u = (u - Lp) / Rp;
            npdf = Rp * npdf;
            ref = nr.rref;
        }
    }
    if (obj < 0) {
        pdf = 0.0f;
        pt = mk(0.0f, 0.0f, 0.0f);
    } else if (DIM == 3) {
        const float4 a = V[F[3 * obj]], b = V[F[3 * obj + 1]], d = V[F[3 * obj + 2]];
        pt = sample_triangle(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(d.x, d.y, d.z), u1, u2);
    } else { // sample_line (scene.cuh:29-32)
        const float4 a = V[F[2 * obj]], b = V[F[2 * obj + 1]];
        pt = mk(a.x + u1 * (b.x - a.x), a.y + u1 * (b.y - a.y), 0.0f);
    }
}

// ---------------------------------------------------------------------------
// Plain per-thread traversals (query.cuh:79-423 in a user's kernel): near
// child first, far child on a private stack of 64 entries (a root-to-leaf
// path of a Karras tree holds at most 64 internal nodes, so it never
// overflows), the library's pruning tests, ties to the smallest index.  The
// answers equal the library's batched queries.
constexpr uint32_t DEV_STACK = 64;

// closest primitive (query.cuh:238-318): squared distance, original index
template <int DIM>
__device__ inline void dev_nearest(const TreeView& t, v3 p, float& best, uint32_t& best_idx)
{
    best = F_INF;
    best_idx = INVALID;
    if (t.n_leaves == 0)
        return;
    if (t.n_leaves == 1) {
        cp_leaf<DIM>(t, 0, p, best, best_idx);
        return;
    }
    uint32_t sn[DEV_STACK];
    float sk[DEV_STACK];
    uint32_t sp = 0, node = 0;
    for (uint32_t it = 0; it < 2 * t.n_leaves + 64 && node < t.n_leaves - 1; ++it) {
        const NodeRec nr = load_node(t.tnode, node);
        const float Lmd = box_mindist(nr.L, p), Rmd = box_mindist(nr.R, p);
        const float Lmm = box_minmaxdist(nr.L, p), Rmm = box_minmaxdist(nr.R, p);
        const bool cL = Lmd <= Rmm, cR = Rmd <= Lmm; // query.cuh:279, :297
        uint32_t lf, lc, rf, rc;
        const bool lLeaf = as_leaf(nr.lref, nr.lrun, lf, lc), rLeaf = as_leaf(nr.rref, nr.rrun, rf, rc);
        if (cL && lLeaf && Lmd <= best)
            for (uint32_t k = 0; k < lc; ++k)
                cp_leaf<DIM>(t, lf + k, p, best, best_idx);
        if (cR && rLeaf && Rmd <= best)
            for (uint32_t k = 0; k < rc; ++k)
                cp_leaf<DIM>(t, rf + k, p, best, best_idx);
        const bool gL = cL && !lLeaf && Lmd <= best, gR = cR && !rLeaf && Rmd <= best;
        if (gL && gR && sp < DEV_STACK) {
            const bool lfirst = Lmd <= Rmd;
            sn[sp] = lfirst ? nr.rref : nr.lref;
            sk[sp++] = lfirst ? Rmd : Lmd;
            node = lfirst ? nr.lref : nr.rref;
            continue;
        }
        if (gL || gR) {
            node = gL ? nr.lref : nr.rref;
            continue;
        }
        bool found = false;
        while (sp > 0) {
            --sp;
            if (sk[sp] <= best) { // query.cuh:259
                node = sn[sp];
                found = true;
                break;
            }
        }
        if (!found)
            break;
    }
}

// closest silhouette (query.cuh:325-423): distance and silhouette id
template <int DIM>
__device__ inline void dev_nearest_silhouette(const TreeView& t, v3 p, bool flip, float& best, uint32_t& best_e)
{
    best = F_INF;
    best_e = INVALID;
    Counts cnt;
    if (t.n_leaves == 0)
        return;
    if (t.n_leaves == 1) {
        const box3 b = ldbox2(t.nbox, 0);
        const float4 c0 = t.ncone[0], c1 = t.ncone[1];
        if (cone_valid(c0.w) && overlap_dim<DIM>(c0, c1.x, p, b, box_mindist(b, p)))
            sil_leaf<DIM>(t, 0, p, flip, best, best_e, cnt);
        return;
    }
    uint32_t sn[DEV_STACK];
    float sk[DEV_STACK];
    uint32_t sp = 0, node = 0;
    for (uint32_t it = 0; it < 2 * t.n_leaves + 64 && node < t.n_leaves - 1; ++it) {
        const NodeRec nr = load_node(t.tnode, node);
        const float Lmd = box_mindist(nr.L, p), Rmd = box_mindist(nr.R, p);
        const float b2 = best * best;
        bool hL = Lmd <= b2, hR = Rmd <= b2;
        if (hL || hR) { // query.cuh:366-367
            const float4 CL = t.tcone[4 * node], CR = t.tcone[4 * node + 1], Cr = t.tcone[4 * node + 2];
            if (DIM == 3 && (t.flags & TREE_FAST_CONE)) {
                const float4 SC = t.tcone[4 * node + 3];
                hL = hL && cone_valid(CL.w) &&
                     cone_overlap_fast(mk(CL.x, CL.y, CL.z), CL.w, SC.x, SC.y, Cr.x, p, nr.L, Lmd);
                hR = hR && cone_valid(CR.w) &&
                     cone_overlap_fast(mk(CR.x, CR.y, CR.z), CR.w, SC.z, SC.w, Cr.y, p, nr.R, Rmd);
            } else {
                hL = hL && cone_valid(CL.w) && overlap_dim<DIM>(CL, Cr.x, p, nr.L, Lmd);
                hR = hR && cone_valid(CR.w) && overlap_dim<DIM>(CR, Cr.y, p, nr.R, Rmd);
            }
        }
        uint32_t lf, lc, rf, rc;
        const bool lLeaf = as_leaf(nr.lref, nr.lrun, lf, lc), rLeaf = as_leaf(nr.rref, nr.rrun, rf, rc);
        if (hL && lLeaf)
            for (uint32_t k = 0; k < lc; ++k)
                sil_leaf<DIM>(t, lf + k, p, flip, best, best_e, cnt);
        if (hR && rLeaf)
            for (uint32_t k = 0; k < rc; ++k)
                sil_leaf<DIM>(t, rf + k, p, flip, best, best_e, cnt);
        const float nb2 = best * best;
        const bool gL = hL && !lLeaf && Lmd <= nb2, gR = hR && !rLeaf && Rmd <= nb2;
        if (gL && gR && sp < DEV_STACK) {
            const bool lfirst = Lmd <= Rmd;
            sn[sp] = lfirst ? nr.rref : nr.lref;
            sk[sp++] = lfirst ? Rmd : Lmd;
            node = lfirst ? nr.lref : nr.rref;
            continue;
        }
        if (gL || gR) {
            node = gL ? nr.lref : nr.rref;
            continue;
        }
        bool found = false;
        while (sp > 0) {
            --sp;
            if (sk[sp] <= best * best) { // query.cuh:347
                node = sn[sp];
                found = true;
                break;
            }
        }
        if (!found)
            break;
    }
}

// ray (query.cuh:79-169): the closest hit before max_dist (or any hit), with
// its parameter t, barycentric (u, v) (2D: the segment parameter in u) and
// original index
template <int DIM>
__device__ inline bool dev_ray(const TreeView& t, v3 ro, v3 rd, float max_dist, bool any_hit, float& tbest,
                               float& bu, float& bv, uint32_t& bidx)
{
    const v3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z); // aabb.cuh:308-311 (2D: z slab inert)
    tbest = F_INF;
    bu = bv = 0.0f;
    bidx = INVALID;
    bool found = false;
    auto leaf = [&](uint32_t j) {
        if (j >= t.n_leaves)
            return;
        const float4 a = t.tleaf[3 * j], b = t.tleaf[3 * j + 1];
        float th, uh, vh = 0.0f;
        const uint32_t o = __float_as_uint(a.w);
        bool h;
        if (DIM == 3) {
            const float4 c = t.tleaf[3 * j + 2];
            h = ray_triangle(ro, rd, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(c.x, c.y, c.z), th, uh, vh);
        } else {
            h = ray_segment(ro, rd, mk(a.x, a.y, 0.0f), mk(b.x, b.y, 0.0f), th, uh);
        }
        if (h && th < max_dist && (th < tbest || (th == tbest && o < bidx))) { // query.cuh:115
            tbest = th;
            bu = uh;
            bv = vh;
            bidx = o;
            found = true;
        }
    };
    if (t.n_leaves == 0)
        return false;
    if (t.n_leaves == 1) {
        leaf(0);
        return found;
    }
    uint32_t sn[DEV_STACK];
    float sk[DEV_STACK];
    uint32_t sp = 0, node = 0;
    for (uint32_t it = 0; it < 2 * t.n_leaves + 64 && node < t.n_leaves - 1; ++it) {
        const NodeRec nr = load_node(t.tnode, node);
        float Ld = F_INF, Rd = F_INF;
        const bool Lh = box_ray(nr.L, ro, inv, max_dist, Ld) && Ld <= tbest;
        const bool Rh = box_ray(nr.R, ro, inv, max_dist, Rd) && Rd <= tbest;
        uint32_t lf, lc, rf, rc;
        const bool lLeaf = as_leaf(nr.lref, nr.lrun, lf, lc), rLeaf = as_leaf(nr.rref, nr.rrun, rf, rc);
        const bool rfirst = Lh && Rh && Rd < Ld; // nearer child first, as query.cuh:136-149
        if (rfirst) {
            if (Rh && rLeaf)
                for (uint32_t k = 0; k < rc; ++k)
                    leaf(rf + k);
            if (Lh && lLeaf && Ld <= tbest)
                for (uint32_t k = 0; k < lc; ++k)
                    leaf(lf + k);
        } else {
            if (Lh && lLeaf)
                for (uint32_t k = 0; k < lc; ++k)
                    leaf(lf + k);
            if (Rh && rLeaf && Rd <= tbest)
                for (uint32_t k = 0; k < rc; ++k)
                    leaf(rf + k);
        }
        if (any_hit && found)
            break;
        const bool gL = Lh && !lLeaf && Ld <= tbest, gR = Rh && !rLeaf && Rd <= tbest;
        if (gL && gR && sp < DEV_STACK) {
            const bool lnear = !(Rd < Ld);
            sn[sp] = lnear ? nr.rref : nr.lref;
            sk[sp++] = lnear ? Rd : Ld;
            node = lnear ? nr.lref : nr.rref;
            continue;
        }
        if (gL || gR) {
            node = gL ? nr.lref : nr.rref;
            continue;
        }
        bool more = false;
        while (sp > 0) {
            --sp;
            if (sk[sp] <= tbest) { // query.cuh:106
                node = sn[sp];
                more = true;
                break;
            }
        }
        if (!more)
            break;
    }
    return found;
}

} // namespace snch
