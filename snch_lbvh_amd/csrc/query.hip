// Query kernels for triangle scenes on gfx950: closest primitive, closest
// silhouette, ray (closest hit / any hit) and SampleTriangleInSphere.
//
// One lane = one query (the reference: one thread per query through
// query_device, query.cuh:79-423, and sample_object_in_sphere,
// sample.cuh:23-92).  Differences in HOW, not WHAT:
//  * Node layout: one 64-B record per internal node holds both children's
//    boxes and references; the silhouette query reads a 48-B cone record from
//    a separate stream, so the closest-point and ray queries never touch it.
//  * Stack: closest point, closest silhouette and rays keep a short per-lane
//    stack in LDS (16 packed entries, a column per lane so a wave's 64 lanes
//    hit distinct banks) with a restart trail (TrailStack): nothing in
//    private memory; a far child pushed out of the full stack is found again
//    from the trail.  The generic overlap and line queries, which must emit
//    hits in the reference's visit order, keep 16 LDS entries over a private
//    overflow array (the tree depth is at most 64, so it never loses one).
//  * Order: the nearer child is descended first and the farther is pushed,
//    and a child is skipped when its lower bound exceeds the best answer so
//    far.  The reference's own pruning predicates (minmaxdist for closest
//    point, cone overlap for silhouettes, slab tests for rays) are evaluated
//    unchanged; the added bound checks are conservative.
//  * Ties: among candidates at exactly the same distance the smallest index
//    wins (the reference keeps the last visited one for closest point and
//    silhouette, query.cuh:286/:383, and the first for rays, :115).  The
//    answer is then a function of the candidate set alone, independent of
//    visit order, block size or a seeded search radius.
//  * Blocks are remapped so that each of the 8 XCDs processes one contiguous
//    range of queries (blocks are dispatched round-robin over XCDs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.hpp"
#include "traverse.hpp"

namespace snch {

namespace {

constexpr int SS = 16;  // default LDS stack entries per lane
constexpr int OVF = 64; // private overflow entries per lane (generic queries)

__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
    const uint32_t X = 8, q = nb / X, r = nb % X, x = b % X, k = b / X;
    return x * q + (x < r ? x : r) + k;
}

// A short stack with a restart trail (Laine 2010, "Restart trail for
// stackless BVH traversal"), for searches whose pruning bound only tightens:
// SSZ 4-B entries in a circular LDS column, and nothing in private memory.
// An entry is node << 8 | q for trees of at most 2^24 leaves (internal node
// ids < 2^24); q is an 8-bit lower bound of the key (a non-negative squared
// distance): the key's float bits >> 20 (exponent and 3 mantissa bits, the
// key truncated to 8 steps per octave) minus a base chosen per tree so that
// the 255 levels span 31 octaves below the squared scene diagonal; keys
// below that range store 0 (decoded as 0), keys above saturate at 255, so
// the decoded value never exceeds the key and a popped entry is never
// skipped wrongly.  A push onto a full stack overwrites the oldest
// entry (the shallowest level); the lost far children are found again
// through the trail and revisited from the root.  Per lane, one bit per tree
// level (root 0; a path holds at most 64 internal nodes):
//   T  at this level the near child is finished and the far child taken;
//   S  at this level only one child was descended (nothing left behind);
//   M  the levels whose far child sits in the stack (levels grow from the
//      stack's bottom to its top, so the top entry's level is M's highest).
// A level of the current path with neither T nor S set, and not in M, left
// its far child behind; when the stack runs empty the deepest such level j
// gets T and the search restarts from the root: levels above j follow the
// trail (T: far child; else near, re-pushing a surviving far child), and
// level j takes its far child.  Children survive only if they pass the
// pruning tests, which can only fail more often as the bound tightens, so a
// subtree left by the trail holds nothing new.
template <int BS, int SSZ>
struct TrailStack {
    uint32_t* lds;
    uint32_t top; // position of the next push (mod SSZ)
    uint32_t qbase;
    uint64_t T, S, M;
    uint32_t lv; // level of the current node

    __device__ __forceinline__ void init(uint32_t* base, uint32_t qb)
    {
        lds = base;
        qbase = qb;
        reset();
    }
    __device__ __forceinline__ void reset()
    {
        top = 0;
        T = S = M = 0ull;
        lv = 0;
    }
    __device__ __forceinline__ uint32_t enc(uint32_t n, float d) const
    {
        const uint32_t v = __float_as_uint(d) >> 20;
        const uint32_t q = v > qbase ? min(v - qbase, 255u) : 0u;
        return (n << 8) | q;
    }
    __device__ __forceinline__ void push(uint32_t n, float d, uint32_t k)
    {
        lds[(top % SSZ) * BS] = enc(n, d);
        ++top;
        if (__popcll(M) >= SSZ)
            M &= M - 1ull; // the oldest entry was just overwritten
        M |= 1ull << k;
    }
    // pop the top entry (M != 0): its node and decoded key; the current
    // path ends at its level, whose near child is finished
    __device__ __forceinline__ void pop(uint32_t& n, float& d)
    {
        --top;
        const uint32_t e = lds[(top % SSZ) * BS];
        n = e >> 8;
        const uint32_t q = e & 255u;
        d = q ? __uint_as_float((q + qbase) << 20) : 0.0f;
        const uint32_t k = 63u - (uint32_t)__clzll(M);
        const uint64_t kb = 1ull << k;
        M &= ~kb;
        T = (T & (kb - 1ull)) | kb;
        S &= kb - 1ull;
        lv = k;
    }
    // the stack is empty: restart from the root for the deepest level of the
    // current path that left its far child behind; false when there is none
    __device__ __forceinline__ bool restart()
    {
        const uint64_t path = lv >= 63u ? ~1ull : ((2ull << lv) - 1ull) & ~1ull; // levels 1..lv
        const uint64_t Z = ~(T | S) & path;
        if (!Z)
            return false;
        const uint32_t j = 63u - (uint32_t)__clzll(Z);
        const uint64_t jb = 1ull << j;
        T = (T & (jb - 1ull)) | jb;
        S &= jb - 1ull;
        top = 0;
        lv = 0;
        return true;
    }
    // choose the child to descend from the current node (children at level
    // lv + 1), given which of the near (N) and far (F) children survive;
    // false: nothing to descend
    __device__ __forceinline__ bool descend(bool gN, bool gF, uint32_t nN, uint32_t nF, float dF, uint32_t& node)
    {
        if (!gN && !gF)
            return false;
        const uint32_t k = lv + 1u; // at most 63: the caller stops deeper paths
        const uint64_t kb = 1ull << k;
        if (T & kb) { // near finished before: the far child, if it survives
            if (!gF)
                return false;
            node = nF;
        } else if (gN && gF) {
            push(nF, dF, k);
            node = nN;
        } else if (gN) {
            S |= kb;
            node = nN;
        } else if (gF) { // off the trail: the near child no longer survives
            T = (T & (kb - 1ull)) | kb;
            S &= kb - 1ull;
            node = nF;
        } else {
            return false;
        }
        lv = k;
        return true;
    }
};

// per-query visit counters (Counts), summed over the wave before one atomic per wave
__device__ __forceinline__ void flush_counts(const Counts& c, unsigned long long* stats)
{
    unsigned long long v[5] = {c.box, c.leaf, c.edge, c.node, c.cone};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            v[k] += __shfl_xor(v[k], o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k)
            atomicAdd(stats + k, v[k]);
    }
}

// ------------------------------------------------------------ closest point

// query_device(nearest) (query.cuh:238-318) for one query point; best is the
// squared distance
template <int BS, int SSZ, bool STATS, int DIM>
__device__ __forceinline__ void cp_query(const TreeView& t, v3 p, uint32_t* stk_col, float& best,
                                         uint32_t& best_idx, Counts& cnt, bool& overflow)
{
    best = F_INF;
    best_idx = INVALID;
    if (t.n_leaves == 1) {
        cp_leaf<DIM>(t, 0, p, best, best_idx);
        return;
    }
    TrailStack<BS, SSZ> st;
    st.init(stk_col, t.qbase);
    uint32_t node = 0;
    // a pass visits each internal node at most once, restarts add the paths
    // above lost far children; the cap and the range checks only guard
    // against a malformed tree (hitting the cap is reported as an overflow)
    const uint32_t cap_iters = 8 * t.n_leaves + 8192;
    uint32_t it = 0;
    for (; it < cap_iters && node < t.n_leaves - 1; ++it) {
        const NodeRec nr = load_node(t.tnode, node);
        const float Lmd = box_mindist(nr.L, p), Rmd = box_mindist(nr.R, p);
        const float Lmm = box_minmaxdist(nr.L, p), Rmm = box_minmaxdist(nr.R, p);
        const bool cL = Lmd <= Rmm, cR = Rmd <= Lmm; // query.cuh:279, :297
        uint32_t lf, lc, rf, rc;
        const bool lLeaf = as_leaf(nr.lref, nr.lrun, lf, lc), rLeaf = as_leaf(nr.rref, nr.rrun, rf, rc);
        if (STATS) {
            cnt.node += 1;
            cnt.box += 2;
        }
        if (cL && lLeaf && Lmd <= best) {
            if (STATS)
                cnt.leaf += lc;
            for (uint32_t k = 0; k < lc; ++k)
                cp_leaf<DIM>(t, lf + k, p, best, best_idx);
        }
        if (cR && rLeaf && Rmd <= best) {
            if (STATS)
                cnt.leaf += rc;
            for (uint32_t k = 0; k < rc; ++k)
                cp_leaf<DIM>(t, rf + k, p, best, best_idx);
        }
        const bool gL = cL && !lLeaf && Lmd <= best;
        const bool gR = cR && !rLeaf && Rmd <= best;
        if ((gL || gR) && st.lv >= 63u) {
            overflow = true; // deeper than a Karras tree can be
            break;
        }
        const bool lfirst = Lmd <= Rmd;
        if (st.descend(lfirst ? gL : gR, lfirst ? gR : gL, lfirst ? nr.lref : nr.rref, lfirst ? nr.rref : nr.lref,
                       lfirst ? Rmd : Lmd, node))
            continue;
        bool found = false;
        while (st.M) {
            uint32_t nn;
            float dd;
            st.pop(nn, dd);
            if (dd <= best) { // query.cuh:259
                node = nn;
                found = true;
                break;
            }
        }
        if (found)
            continue;
        if (!st.restart())
            break;
        node = 0;
    }
    overflow = overflow || it >= cap_iters;
}

template <int BS, bool STATS, int DIM, int SSZ>
__global__ __launch_bounds__(BS) void k_closest_point(TreeView t, const float* __restrict__ q, uint32_t n,
                                                      uint32_t* __restrict__ out_idx, float* __restrict__ out_dist,
                                                      const uint32_t* __restrict__ perm, uint32_t* __restrict__ err,
                                                      unsigned long long* stats)
{
    __shared__ uint32_t stk[SSZ * BS];
    const uint32_t slot = xcd_remap(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    Counts cnt;
    if (t.n_dev)
        n = min(n, *t.n_dev);
    if (slot >= n) {
        if (STATS)
            flush_counts(cnt, stats);
        return;
    }
    const uint32_t qi = perm ? min(perm[slot], n - 1) : slot;
    float best;
    uint32_t best_idx;
    bool ovf = false;
    cp_query<BS, SSZ, STATS, DIM>(t, load_point<DIM>(q, qi), stk + threadIdx.x, best, best_idx, cnt, ovf);
    if (ovf && err)
        atomicOr(err, 1u);
    out_idx[qi] = best_idx;
    out_dist[qi] = dm::sqrt_(best);
    if (STATS)
        flush_counts(cnt, stats);
}

// ------------------------------------------------------------ silhouette

// Per-wave LDS of the flattened leaf tests: one round of 64 edge slots (the
// edge, then its distance with the sign bit set when it is not a silhouette),
// the owner lane of each slot, and each owner's query point and bound.
struct EdgeLds {
    uint32_t edge[64];
    float d[64];
    uint8_t owner[64];
    float px[64], py[64], pz[64], bound[64];
};

// The leaf tests of the lanes with `mine` set, spread over every lane of the
// wave that runs this (all looping lanes, wave-uniformly): each owner lists
// the edges of its (at most two) leaves as a run of slots, and in rounds of
// 64 slots every lane computes the distance and the silhouette test of one,
// then each owner applies its results in its original edge order, so ties and
// the reference's early-out at distance 0 (scene.cuh:791) come out as in
// sil_leaf.  3D, leaves of one primitive each.
__device__ __forceinline__ void sil_leaves_wave(const TreeView& t, EdgeLds& L, bool mine, uint32_t la, bool ha,
                                                uint32_t lb, bool hb, v3 p, bool flip, float& best,
                                                uint32_t& best_e, Counts& cnt)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t ea = 0, eb = 0, sa = 0, sb = 0;
    if (mine && ha && la < t.n_leaves) {
        sa = __float_as_uint(t.tleaf[3 * la + 1].w);
        ea = __float_as_uint(t.tleaf[3 * la + 2].w);
    }
    if (mine && hb && lb < t.n_leaves) {
        sb = __float_as_uint(t.tleaf[3 * lb + 1].w);
        eb = __float_as_uint(t.tleaf[3 * lb + 2].w);
    }
    ea = min(ea, 3u);
    eb = min(eb, 3u);
    const uint32_t my = ea + eb; // 0..6
    const uint64_t b0 = __ballot(my & 1u), b1 = __ballot(my & 2u), b2 = __ballot(my & 4u);
    const uint32_t off = (uint32_t)__popcll(b0 & lt) + 2u * (uint32_t)__popcll(b1 & lt) + 4u * (uint32_t)__popcll(b2 & lt);
    const uint32_t total = (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
    if (mine) {
        L.px[lane] = p.x;
        L.py[lane] = p.y;
        L.pz[lane] = p.z;
        L.bound[lane] = best;
        cnt.leaf += (ha ? 1u : 0u) + (hb ? 1u : 0u);
        cnt.edge += my;
    }
    const uint64_t helpers = __ballot(true);
    const uint32_t hr = (uint32_t)__popcll(helpers & lt), nh = (uint32_t)__popcll(helpers);
    bool stop = false; // best reached 0: the owner's remaining edges are skipped
    for (uint32_t base = 0; base < total; base += 64) {
        // this round's slots [base, base + 64) of every owner
        for (uint32_t k = 0; mine && k < my; ++k) {
            const uint32_t sl = off + k;
            if (sl >= base && sl < base + 64) {
                L.edge[sl - base] = k < ea ? sa + k : sb + (k - ea);
                L.owner[sl - base] = (uint8_t)lane;
            }
        }
        // every lane of the wave (LDS operations of one wave complete in order)
        const uint32_t n_round = min(64u, total - base);
        for (uint32_t r = hr; r < n_round; r += nh) {
            const uint32_t e = L.edge[r], o = L.owner[r];
            const v3 q = mk(L.px[o], L.py[o], L.pz[o]);
            const float b = L.bound[o], mr2 = b * b;
            const float4 A = t.tedge[4 * e], B = t.tedge[4 * e + 1];
            const uint32_t fl = __float_as_uint(B.w);
            const v3 pa = mk(A.x, A.y, A.z), pb = mk(B.x, B.y, B.z);
            v3 cp;
            const float d = closest_point_segment(pa, pb, q, cp);
            bool sil = (fl & 3u) != 3u; // a boundary edge is always a silhouette
            if (!sil && d * d <= mr2) {
                const float4 N0 = t.tedge[4 * e + 2], N1 = t.tedge[4 * e + 3];
                sil = is_silhouette_edge(pa, pb, mk(N0.x, N0.y, N0.z), mk(N1.x, N1.y, N1.z),
                                         mk(q.x - cp.x, q.y - cp.y, q.z - cp.z), d, flip);
            }
            L.d[r] = sil ? d : -d; // d >= 0: the sign bit carries "not a silhouette"
        }
        for (uint32_t k = 0; mine && !stop && k < my; ++k) {
            const uint32_t sl = off + k;
            if (sl < base || sl >= base + 64)
                continue;
            const float mr2 = best * best;
            if (!(0.0f < mr2)) {
                stop = true; // min_radius_squared (0) >= max_radius_squared
                break;
            }
            const float sd = L.d[sl - base];
            const float d = dm::fabs_(sd);
            if (d * d > mr2 || dm::signbit_(sd))
                continue;
            const uint32_t e = k < ea ? sa + k : sb + (k - ea);
            const uint32_t id = __float_as_uint(t.tedge[4 * e].w);
            if (d < best || (d == best && id < best_e)) {
                best = d;
                best_e = id;
            }
        }
    }
}


// query_device(nearest_silhouette) (query.cuh:325-423) for one query point.
// best / best_e receive the answer.  The stack is a TrailStack: SSZ entries
// in LDS and no private memory.  A finite radius `cap` prunes internal nodes
// farther than cap as well.  When the search within cap ends with
// best <= cap, every node that could hold a silhouette at least as close has
// been visited; otherwise (HINT) it searches again from the root with no cap
// and the bound found so far, or (hard cap, !HINT) it stops.  With index
// tie-breaking the answer equals the uncapped search.
template <int BS, int SSZ, bool STATS, int DIM, bool HINT, bool PARK>
__device__ __forceinline__ void sil_query(const TreeView& t, v3 p, bool flip, float cap, uint32_t* stk_col,
                                          EdgeLds* elds, float& best, uint32_t& best_e, Counts& cnt,
                                          bool& overflow)
{
    best = F_INF;
    best_e = INVALID;
    if (t.n_leaves == 1) {
        const box3 b = ldbox2(t.nbox, 0);
        const float4 c0 = t.ncone[0], c1 = t.ncone[1];
        if (cone_valid(c0.w) && overlap_dim<DIM>(c0, c1.x, p, b, box_mindist(b, p)))
            sil_leaf<DIM>(t, 0, p, flip, best, best_e, cnt);
        return;
    }
    if (!(cap >= 0.0f))
        cap = F_INF;
    float cap2 = cap * cap;
    TrailStack<BS, SSZ> st;
    st.init(stk_col, t.qbase);
    uint32_t node = 0;
    // a pass visits each internal node at most once; restarts add the paths
    // above lost far children (a few per query); the bound only guards
    // against an inconsistent tree, and hitting it is reported as an overflow
    const uint32_t cap_iters = 8 * t.n_leaves + 8192;
    uint32_t it = 0;
    // A lane whose node has leaf children to evaluate parks them (phase
    // `leaves`) until a quarter of the wave's looping lanes, or all of them,
    // are parked; the parked lanes then evaluate their leaves together and
    // decide the node's internal children with the bound found, exactly as
    // if the leaves had been evaluated at once.  The leaf code then runs on
    // many lanes at a time instead of on the few that reached a leaf (about
    // 5 of 64 on the 1M-triangle torus: 99.6 -> 89 ms).  Hard-capped searches
    // (!HINT: short, as in walk on stars) evaluate leaves at once.
    bool leaves = false;
    float Lmd = 0.0f, Rmd = 0.0f;
    bool hL = false, hR = false, lLeaf = false, rLeaf = false;
    uint32_t lref = 0, rref = 0, lf = 0, lc = 0, rf = 0, rc = 0;
    for (; it < cap_iters && node < t.n_leaves - 1;) {
        bool decide = false;
        if (!leaves) {
            ++it;
            const NodeRec nr = load_node(t.tnode, node);
            Lmd = box_mindist(nr.L, p);
            Rmd = box_mindist(nr.R, p);
            const float b2 = best * best;
            hL = Lmd <= b2;
            hR = Rmd <= b2;
            if (STATS) {
                cnt.node += 1;
                cnt.box += 2;
            }
            if (hL || hR) {
                if (STATS)
                    cnt.cone += 1;
                const float4 CL = t.tcone[4 * node], CR = t.tcone[4 * node + 1], Cr = t.tcone[4 * node + 2];
                // query.cuh:366-367
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
            lref = nr.lref;
            rref = nr.rref;
            lLeaf = as_leaf(nr.lref, nr.lrun, lf, lc);
            rLeaf = as_leaf(nr.rref, nr.rrun, rf, rc);
            leaves = (hL && lLeaf) || (hR && rLeaf);
            decide = !leaves;
        }
        const uint64_t parked = __ballot(leaves), looping = __ballot(true);
        if (parked &&
            (!PARK || 4u * (uint32_t)__popcll(parked) >= (uint32_t)__popcll(looping) || parked == looping)) {
            // wave-uniform: the parked lanes' leaves are tested now
            const bool la = hL && lLeaf, ra = hR && rLeaf;
            const bool runs = __ballot(leaves && ((la && lc > 1) || (ra && rc > 1))) != 0;
            if (DIM == 3 && elds && !runs) {
                sil_leaves_wave(t, *elds, leaves, lf, la, rf, ra, p, flip, best, best_e, cnt);
            } else if (leaves) {
                if (la)
                    for (uint32_t k = 0; k < lc; ++k)
                        sil_leaf<DIM>(t, lf + k, p, flip, best, best_e, cnt);
                if (ra)
                    for (uint32_t k = 0; k < rc; ++k)
                        sil_leaf<DIM>(t, rf + k, p, flip, best, best_e, cnt);
            }
            decide = decide || leaves;
            leaves = false;
        }
        if (!decide)
            continue;
        // internal children to descend: within the bound and the cap
        const float nb2 = best * best;
        const bool gL = hL && !lLeaf && Lmd <= nb2 && Lmd <= cap2;
        const bool gR = hR && !rLeaf && Rmd <= nb2 && Rmd <= cap2;
        if ((gL || gR) && st.lv >= 63u) {
            overflow = true; // deeper than a Karras tree can be
            break;
        }
        const bool lfirst = Lmd <= Rmd;
        if (st.descend(lfirst ? gL : gR, lfirst ? gR : gL, lfirst ? lref : rref, lfirst ? rref : lref,
                       lfirst ? Rmd : Lmd, node))
            continue;
        bool found = false;
        while (st.M) {
            uint32_t nn;
            float dd;
            st.pop(nn, dd);
            if (dd <= best * best) { // query.cuh:347
                node = nn;
                found = true;
                break;
            }
        }
        if (found)
            continue;
        if (st.restart()) {
            node = 0;
            continue;
        }
        // the search within the cap is over: settled, or (HINT) search again
        // from the root with no cap and the bound found so far
        if (cap2 == F_INF || (best_e != INVALID && best <= cap) || !HINT)
            break;
        cap2 = F_INF;
        st.reset();
        node = 0;
    }
    overflow = overflow || it >= cap_iters;
}

// MODE 0: plain search.  MODE 1: search within the radius seed[qi] first; a
// query not settled there searches again from the root with no cap and the
// bound found so far (bit-identical to MODE 0).  MODE 2: search within the
// radius only (a hard cap).  Keeping the second search in the same launch
// lets the rare unsettled queries (the costliest ones) overlap with the rest
// of the grid; a separate second launch over them measured ~30 ms slower on
// the 1M-triangle torus, all of it the tail of those few queries.  The plain
// kernel is held to 80 VGPRs (6 waves per SIMD), the others to 96.
template <int BS, bool STATS, int DIM, int MODE, int SSZ>
__global__ __launch_bounds__(BS, MODE == 0 && !STATS ? 6 : 5) void k_closest_silhouette(TreeView t, const float* __restrict__ q, uint32_t n,
                                                           uint32_t nq, int flip, const float* __restrict__ seed,
                                                           uint32_t* __restrict__ out_e, float* __restrict__ out_d,
                                                           const uint32_t* __restrict__ perm,
                                                           uint32_t* __restrict__ err, unsigned long long* stats)
{
    __shared__ uint32_t stk[SSZ * BS];
    __shared__ EdgeLds elds[DIM == 3 ? BS / 64 : 1];
    const uint32_t slot = xcd_remap(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    Counts cnt;
    if (t.n_dev)
        n = min(n, *t.n_dev);
    const uint32_t qi = slot < n ? (perm ? perm[slot] : slot) : nq;
    if (qi >= nq) {
        if (STATS)
            flush_counts(cnt, stats);
        return;
    }
    float cap = MODE != 0 ? seed[qi] : F_INF;
    if (!(cap >= 0.0f))
        cap = F_INF;
    float best;
    uint32_t best_e;
    bool ovf = false;
    const v3 p = load_point<DIM>(q, qi);
    sil_query<BS, SSZ, STATS, DIM, MODE != 2, MODE != 2>(t, p, flip != 0, cap, stk + threadIdx.x,
                                                         DIM == 3 ? &elds[threadIdx.x / 64] : nullptr, best,
                                                         best_e, cnt, ovf);
    if (ovf && err)
        atomicOr(err, 1u);
    if (out_e)
        out_e[qi] = best_e;
    out_d[qi] = best;
    if (STATS)
        flush_counts(cnt, stats);
}

// Seeded search with re-queued unsettled queries.  A wave's time is that of
// its slowest lane, so a query not settled within its seed radius, searched
// again in place, holds 63 settled lanes idle.  Here the launch is
// persistent: each wave takes 64-query chunks of the visiting order (each
// XCD a contiguous range of chunks first, then the other XCDs' ranges) and
// searches them within their seeds (a hard cap); a query not settled there is
// appended to a queue.  A wave that finds no chunk left takes up to 64 queued
// queries at a time and searches them with no cap, until it finds the queue
// empty.  Every appended entry is taken: its appender itself goes on to drain
// the queue after its last append.  A taken slot may still be being written
// by its appender, which has reserved it and is running, so the wait for it
// is short (and bounded).  Each XCD appends to its own queue and drains it
// first: its waves work through a contiguous stretch of the visiting order,
// so a wave of its queued queries stays in one region of the scene, whose
// nodes its L2 holds (one queue for all mixed the regions of all eight).
// ctr: [0, 8) next chunk of each XCD's range, [8 + 2x] and [9 + 2x] head and
// tail of XCD x's queue (RQ_CTR_WORDS words in all); queue: 8 queues of n
// entries, RQ_EMPTY at launch.  Results equal the unseeded search's.
constexpr uint32_t RQ_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t RQ_SPIN = 1u << 20;
constexpr uint32_t RQ_CTR_WORDS = 24;

__global__ __launch_bounds__(256) void k_requeue_init(uint32_t* __restrict__ ctr, uint32_t* __restrict__ queue,
                                                      uint32_t n)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < RQ_CTR_WORDS)
        ctr[i] = 0u;
    for (uint32_t k = i; k < n; k += gridDim.x * 256)
        queue[k] = RQ_EMPTY;
}

// wave-uniform copy of lane 0's value (a scalar register)
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// the next 64-query chunk of nchunks for a one-wave block of a persistent
// launch: its own XCD's contiguous range first (x0: the block's XCD), then
// the other XCDs' ranges; xi counts the ranges exhausted (wave-uniform).
// ctr: [0, 8) the ranges' counters.  INVALID when none is left.
__device__ __forceinline__ uint32_t take_chunk(uint32_t* ctr, uint32_t nchunks, uint32_t x0, uint32_t& xi)
{
    while (xi < 8u) {
        const uint32_t xx = (x0 + xi) & 7u;
        const uint32_t lo = (uint32_t)((uint64_t)nchunks * xx / 8u);
        const uint32_t hi = (uint32_t)((uint64_t)nchunks * (xx + 1u) / 8u);
        uint32_t k = 0;
        if (threadIdx.x == 0)
            k = atomicAdd(ctr + xx, 1u);
        k = uniform(k);
        if (k < hi - lo)
            return lo + k;
        ++xi;
    }
    return INVALID;
}

template <int SSZ, bool STATS, int DIM>
__global__ __launch_bounds__(64, STATS ? 5 : 6) void k_sil_requeue(TreeView t, const float* __restrict__ q,
                                                                 uint32_t n, int flip, const float* __restrict__ seed,
                                                                 uint32_t* __restrict__ out_e,
                                                                 float* __restrict__ out_d,
                                                                 const uint32_t* __restrict__ perm, uint32_t* ctr,
                                                                 uint32_t* queue, uint32_t* __restrict__ err,
                                                                 unsigned long long* stats)
{
    __shared__ uint32_t stk[SSZ * 64];
    __shared__ EdgeLds elds[1];
    const uint32_t lane = threadIdx.x;
    const uint32_t nchunks = (n + 63) / 64;
    const uint32_t x0 = blockIdx.x % 8u; // blocks are dispatched round-robin over the 8 XCDs
    uint32_t xi = 0;                     // ranges exhausted so far, own range first (wave-uniform)
    uint32_t di = 0;                     // queues found empty while draining, own XCD's first
    bool draining = false, ovf = false;
    Counts cnt;
    for (;;) {
        uint32_t qi = INVALID;
        float cap = F_INF;
        if (!draining) {
            uint32_t c = INVALID;
            while (xi < 8u) {
                const uint32_t xx = (x0 + xi) & 7u;
                const uint32_t lo = (uint32_t)((uint64_t)nchunks * xx / 8u);
                const uint32_t hi = (uint32_t)((uint64_t)nchunks * (xx + 1u) / 8u);
                uint32_t k = 0;
                if (lane == 0)
                    k = atomicAdd(ctr + xx, 1u);
                k = uniform(k);
                if (k < hi - lo) {
                    c = lo + k;
                    break;
                }
                ++xi;
            }
            if (c == INVALID) {
                draining = true;
                continue;
            }
            const uint32_t slot = c * 64u + lane;
            if (slot < n) {
                qi = perm ? perm[slot] : slot;
                cap = seed[qi];
                if (!(cap >= 0.0f))
                    cap = F_INF;
            }
        } else {
            // take up to 64 queued queries of queue dq (this XCD's first):
            // [base, base + k) below its tail
            const uint32_t dq = (x0 + di) & 7u;
            uint32_t* head = ctr + 8 + 2 * dq;
            uint32_t base = 0, k = 0, empty = 0;
            if (lane == 0) {
                uint32_t h = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                for (uint32_t tries = 0; tries < 64u; ++tries) {
                    const uint32_t tl = __hip_atomic_load(head + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (h >= tl) {
                        empty = 1u;
                        break;
                    }
                    const uint32_t want = min(64u, tl - h);
                    const uint32_t old = atomicCAS(head, h, h + want); // a failure means another wave took some
                    if (old == h) {
                        base = h;
                        k = want;
                        break;
                    }
                    h = old;
                }
            }
            if (uniform(empty)) {
                if (++di >= 8u)
                    break;
                continue;
            }
            base = uniform(base);
            k = uniform(k);
            if (k == 0u)
                continue; // contended: try again
            if (lane < k) {
                uint32_t v = RQ_EMPTY;
                for (uint32_t s = 0; s < RQ_SPIN && v == RQ_EMPTY; ++s) {
                    v = __hip_atomic_load(queue + (size_t)dq * n + base + lane, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                    if (v == RQ_EMPTY)
                        __builtin_amdgcn_s_sleep(2);
                }
                if (v < n)
                    qi = v;
                else
                    ovf = true; // not expected: the slot's writer had reserved it
            }
        }
        float best = F_INF;
        uint32_t best_e = INVALID;
        if (qi < n)
            sil_query<64, SSZ, STATS, DIM, false, true>(t, load_point<DIM>(q, qi), flip != 0, cap, stk + lane,
                                                        DIM == 3 ? elds : nullptr, best, best_e, cnt, ovf);
        const bool unsettled = qi < n && cap < F_INF && !(best_e != INVALID && best <= cap);
        if (qi < n && !unsettled) {
            if (out_e)
                out_e[qi] = best_e;
            out_d[qi] = best;
        }
        const uint64_t um = __ballot(unsettled);
        if (um) {
            uint32_t b = 0;
            if (lane == 0)
                b = atomicAdd(ctr + 9 + 2 * x0, (uint32_t)__popcll(um)); // this XCD's queue
            b = uniform(b);
            // rank among the unsettled lanes below this one
            const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(um >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)um, 0u));
            if (unsettled)
                __hip_atomic_store(queue + (size_t)x0 * n + b + r, qi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (ovf && err)
        atomicOr(err, 1u);
    if (STATS)
        flush_counts(cnt, stats);
}

// Plain search on a persistent launch: each wave takes 64-query chunks of the
// visiting order, its own XCD's contiguous range of chunks first, then the
// other XCDs' ranges, until none is left.  A wave's time is that of its
// slowest lane, and waves of one launch differ widely; taking chunks as waves
// free up balances that over the whole launch, and the per-XCD ranges keep
// the queries an XCD searches close together in space (its L2 holds their
// nodes).  On the 1M-triangle torus, 4M queries: 75.3 -> 70.7 ms against one
// 64-query block per chunk.  Handing the few lanes left in a wave to fresh
// waves (suspend, resume from the trail) was measured slower: 81-90 ms with
// the suspended queries kept in visiting order, ~140 ms in suspension order.
// CAPPED: the search within the radius seed[qi] only (as k_closest_silhouette's
// MODE 2; walk on stars), over the first min(n, *t.n_dev) queries when the
// count is on the device.  ctr: [0, 8) the chunk counters of the XCD ranges
// (zero at launch).
template <int SSZ, bool STATS, int DIM, bool CAPPED>
__global__ __launch_bounds__(64, STATS || CAPPED ? 5 : 6) void k_sil_persist(TreeView t, const float* __restrict__ q,
                                                                 uint32_t n, int flip, const float* __restrict__ seed,
                                                                 uint32_t* __restrict__ out_e,
                                                                 float* __restrict__ out_d,
                                                                 const uint32_t* __restrict__ perm, uint32_t* ctr,
                                                                 uint32_t* __restrict__ err,
                                                                 unsigned long long* stats)
{
    __shared__ uint32_t stk[SSZ * 64];
    __shared__ EdgeLds elds[1];
    const uint32_t lane = threadIdx.x;
    if (CAPPED && t.n_dev)
        n = min(n, *t.n_dev);
    const uint32_t nchunks = (n + 63) / 64;
    const uint32_t x0 = blockIdx.x % 8u; // blocks are dispatched round-robin over the 8 XCDs
    uint32_t xi = 0;                     // ranges exhausted so far, own range first (wave-uniform)
    bool ovf = false;
    Counts cnt;
    for (;;) {
        const uint32_t c = take_chunk(ctr, nchunks, x0, xi);
        if (c == INVALID)
            break;
        const uint32_t slot = c * 64u + lane;
        const uint32_t qi = slot < n ? (perm ? perm[slot] : slot) : INVALID;
        if (qi < n) {
            float best;
            uint32_t best_e;
            float cap = F_INF;
            if (CAPPED) {
                cap = seed[qi];
                if (!(cap >= 0.0f))
                    cap = F_INF;
            }
            // hard-capped searches are short: leaves are tested at once (no parking)
            sil_query<64, SSZ, STATS, DIM, false, !CAPPED>(t, load_point<DIM>(q, qi), flip != 0, cap, stk + lane,
                                                           DIM == 3 ? elds : nullptr, best, best_e, cnt, ovf);
            if (out_e)
                out_e[qi] = best_e;
            out_d[qi] = best;
        }
    }
    if (ovf && err)
        atomicOr(err, 1u);
    if (STATS)
        flush_counts(cnt, stats);
}

// ------------------------------------------------------------ ray
template <int BS, bool STATS, int DIM>
__global__ __launch_bounds__(BS) void k_ray(TreeView t, const float* __restrict__ o, const float* __restrict__ d,
                                            const float* __restrict__ tmax, uint32_t n, int any_hit,
                                            uint8_t* __restrict__ out_hit, float* __restrict__ out_t,
                                            float* __restrict__ out_uv, uint32_t* __restrict__ out_idx,
                                            uint32_t* __restrict__ err, unsigned long long* stats)
{
    __shared__ uint32_t stk[SS * BS];
    const uint32_t qi = xcd_remap(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    Counts cnt;
    if (t.n_dev)
        n = min(n, *t.n_dev);
    if (qi >= n) {
        if (STATS)
            flush_counts(cnt, stats);
        return;
    }
    const v3 ro = load_point<DIM>(o, qi);
    const v3 rd = load_point<DIM>(d, qi);
    const v3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z); // aabb.cuh:308-311 (2D: z slab inert)
    const float max_dist = tmax ? tmax[qi] : F_INF;
    float tbest = F_INF, bu = 0.0f, bv = 0.0f;
    uint32_t bidx = INVALID;
    bool found = false;
    auto leaf = [&](uint32_t j) {
        if (j >= t.n_leaves)
            return;
        if (STATS)
            cnt.leaf += 1;
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
    if (t.n_leaves == 1) {
        leaf(0);
    } else {
        // the closest hit only shrinks, so the restart-trail stack applies
        TrailStack<BS, SS> st;
        st.init(stk + threadIdx.x, t.qbase);
        uint32_t node = 0;
        const uint32_t cap_iters = 8 * t.n_leaves + 8192;
        uint32_t it = 0;
        bool deep = false;
        for (; it < cap_iters && node < t.n_leaves - 1; ++it) {
            const NodeRec nr = load_node(t.tnode, node);
            float Ld = F_INF, Rd = F_INF;
            if (STATS) {
                cnt.node += 1;
                cnt.box += 2;
            }
            const bool bL = box_ray(nr.L, ro, inv, max_dist, Ld), bR = box_ray(nr.R, ro, inv, max_dist, Rd);
            const bool Lh = bL && Ld <= tbest;
            const bool Rh = bR && Rd <= tbest;
            uint32_t lf, lc, rf, rc;
            const bool lLeaf = as_leaf(nr.lref, nr.lrun, lf, lc), rLeaf = as_leaf(nr.rref, nr.rrun, rf, rc);
            // nearer child first, as query.cuh:136-149
            const bool rfirst = Lh && Rh && Rd < Ld;
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
            const bool gL = Lh && !lLeaf && Ld <= tbest;
            const bool gR = Rh && !rLeaf && Rd <= tbest;
            if ((gL || gR) && st.lv >= 63u) {
                deep = true; // deeper than a Karras tree can be
                break;
            }
            // near and far by geometry alone (a box the ray misses is far), so
            // every pass orders a node's children alike; with both boxes hit
            // this is query.cuh:136-149's order
            const bool lnear = bL && (!bR || !(Rd < Ld));
            if (st.descend(lnear ? gL : gR, lnear ? gR : gL, lnear ? nr.lref : nr.rref, lnear ? nr.rref : nr.lref,
                           lnear ? Rd : Ld, node))
                continue;
            bool more = false;
            while (st.M) {
                uint32_t nn;
                float dd;
                st.pop(nn, dd);
                if (dd <= tbest) { // query.cuh:106
                    node = nn;
                    more = true;
                    break;
                }
            }
            if (more)
                continue;
            if (!st.restart())
                break;
            node = 0;
        }
        if ((deep || it >= cap_iters) && err)
            atomicOr(err, 1u);
    }
    if (out_hit)
        out_hit[qi] = found ? 1 : 0;
    if (STATS)
        flush_counts(cnt, stats);
    if (any_hit)
        return;
    if (out_t)
        out_t[qi] = tbest;
    if (out_uv) {
        if (DIM == 3) {
            out_uv[2 * qi] = bu;
            out_uv[2 * qi + 1] = bv;
        } else {
            out_uv[qi] = bu; // segment parameter s
        }
    }
    if (out_idx)
        out_idx[qi] = bidx;
}

// ------------------------------------------------------------ batches
// One persistent launch over the queries of every mesh of a batch, given as
// one concatenated query array; the queries of each mesh are cut into chunks of `qpw` (a multiple of 64), and chunks are handed out in
// schedule order (meshes with deeper trees first) through one atomic counter.
// Queries of mesh m are [qoff[m], qoff[m + 1]).
// Each wave takes a chunk, answers it 64 queries at a time against that
// mesh's tree, and takes the next, until the counter passes the last chunk:
// every wave reaches that exit.
struct BatchTable {
    const TreeView* views;  // per mesh
    const uint32_t* qoff;   // per mesh, n_meshes + 1 entries
    const uint32_t* sched;  // schedule entry -> mesh
    const uint32_t* cstart; // schedule entry -> first chunk (n_sched + 1 entries)
    const uint32_t* perm;   // optional visiting order of all queries (null: input order)
    uint32_t* counter;
    uint32_t n_sched, total_chunks, qpw;
};

// SEEDED (closest silhouette): seed[qi] is the query's pruning radius (a
// learned bound); the search within it is exact, a query not settled there
// searches again from the root, so results equal the unseeded ones
template <int BS, int SSZ, int KIND, bool SEEDED>
__global__ __launch_bounds__(BS, 5) void k_batch(BatchTable bt, const float* __restrict__ q, int flip,
                                              const float* __restrict__ seed, uint32_t* __restrict__ out_idx,
                                              float* __restrict__ out_dist, uint32_t* __restrict__ err)
{
    __shared__ uint32_t stk[SSZ * BS];
    __shared__ EdgeLds elds[BS / 64];
    const uint32_t lane = threadIdx.x & 63;
    bool ovf = false;
    Counts cnt;
    for (uint32_t guard = 0; guard < bt.total_chunks + 1; ++guard) {
        uint32_t c = 0;
        if (lane == 0)
            c = atomicAdd(bt.counter, 1u);
        c = __shfl(c, 0, 64);
        if (c >= bt.total_chunks)
            break;
        uint32_t lo = 0, hi = bt.n_sched; // last entry with cstart <= c
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (bt.cstart[mid] <= c)
                lo = mid;
            else
                hi = mid;
        }
        const uint32_t m = bt.sched[lo];
        const TreeView tv = bt.views[m];
        const uint32_t q0 = bt.qoff[m] + (c - bt.cstart[lo]) * bt.qpw;
        const uint32_t q1 = min(q0 + bt.qpw, bt.qoff[m + 1]);
        for (uint32_t base = q0; base < q1; base += 64) {
            const uint32_t slot = base + lane;
            if (slot >= q1)
                continue;
            const uint32_t qi = bt.perm ? bt.perm[slot] : slot;
            const v3 p = load_point<3>(q, qi);
            float best = F_INF;
            uint32_t bi = INVALID;
            if (KIND == 0) {
                cp_query<BS, SSZ, false, 3>(tv, p, stk + threadIdx.x, best, bi, cnt, ovf);
                best = dm::sqrt_(best);
            } else {
                float cap = SEEDED ? seed[qi] : F_INF;
                if (!(cap >= 0.0f))
                    cap = F_INF;
                sil_query<BS, SSZ, false, 3, SEEDED, true>(tv, p, flip != 0, cap, stk + threadIdx.x,
                                                           &elds[threadIdx.x / 64], best, bi, cnt, ovf);
            }
            if (out_idx)
                out_idx[qi] = bi;
            out_dist[qi] = best;
        }
    }
    if (ovf && err)
        atomicOr(err, 1u);
}

// Morton code of each query in its own mesh's root box (batches)
__global__ __launch_bounds__(256) void k_batch_codes(const TreeView* __restrict__ views,
                                                     const uint32_t* __restrict__ qmesh, const float* __restrict__ q,
                                                     uint32_t n, uint32_t* __restrict__ codes,
                                                     uint32_t* __restrict__ idx)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const TreeView& t = views[qmesh[i]];
    const float4 lo = t.nbox[0], hi = t.nbox[1];
    const v3 p = load_point<3>(q, i);
    codes[i] = morton3(mk((p.x - lo.x) / (hi.x - lo.x), (p.y - lo.y) / (hi.y - lo.y), (p.z - lo.z) / (hi.z - lo.z)));
    idx[i] = i;
}

__global__ void k_zero_counter(uint32_t* c) { *c = 0u; }

// mesh of each query of a batch (last m with qoff[m] <= i)
__global__ __launch_bounds__(256) void k_query_mesh(const uint32_t* __restrict__ qoff, uint32_t M, uint32_t n,
                                                    uint32_t* __restrict__ qmesh)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    uint32_t lo = 0, hi = M;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (qoff[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    qmesh[i] = lo;
}

// ------------------------------------------------------------ sampling
// sample.cuh:23-92: a single root-to-leaf descent, bit-identical to the
// reference's FP32 arithmetic (u rescaled at each level).
template <int DIM>
__global__ __launch_bounds__(256) void k_sample(TreeView t, const float4* __restrict__ V,
                                                const int32_t* __restrict__ F, const float* __restrict__ cc,
                                                const float* __restrict__ rr, const float* __restrict__ uu,
                                                uint32_t n, int32_t* __restrict__ out_idx,
                                                float* __restrict__ out_pdf, float* __restrict__ out_pt)
{
    const uint32_t qi = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (t.n_dev)
        n = min(n, *t.n_dev);
    if (qi >= n)
        return;
    const v3 c = load_point<DIM>(cc, qi);
    int32_t obj;
    float pdf;
    v3 pt;
    sample_one<DIM>(t, V, F, c, rr[qi], uu[3 * qi], uu[3 * qi + 1], uu[3 * qi + 2], obj, pdf, pt);
    out_idx[qi] = obj;
    out_pdf[qi] = pdf;
    if (out_pt) {
        if (DIM == 3) {
            out_pt[3 * qi] = pt.x;
            out_pt[3 * qi + 1] = pt.y;
            out_pt[3 * qi + 2] = pt.z;
        } else {
            out_pt[2 * qi] = pt.x;
            out_pt[2 * qi + 1] = pt.y;
        }
    }
}

template <int DIM>
__global__ __launch_bounds__(256) void k_sample_on_object(const float4* __restrict__ V, const int32_t* __restrict__ F,
                                                          uint32_t nF, const int32_t* __restrict__ idx,
                                                          const float* __restrict__ uv, uint32_t n,
                                                          float* __restrict__ out_pt)
{
    const uint32_t qi = blockIdx.x * 256 + threadIdx.x;
    if (qi >= n)
        return;
    const int32_t o = idx[qi];
    v3 pt = mk(0.0f, 0.0f, 0.0f);
    if (o >= 0 && (uint32_t)o < nF) {
        if (DIM == 3) {
            const float4 a = V[F[3 * o]], b = V[F[3 * o + 1]], d = V[F[3 * o + 2]];
            pt = sample_triangle(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(d.x, d.y, d.z), uv[2 * qi],
                                 uv[2 * qi + 1]);
        } else {
            const float4 a = V[F[2 * o]], b = V[F[2 * o + 1]];
            const float u1 = uv[2 * qi];
            pt = mk(a.x + u1 * (b.x - a.x), a.y + u1 * (b.y - a.y), 0.0f);
        }
    }
    if (DIM == 3) {
        out_pt[3 * qi] = pt.x;
        out_pt[3 * qi + 1] = pt.y;
        out_pt[3 * qi + 2] = pt.z;
    } else {
        out_pt[2 * qi] = pt.x;
        out_pt[2 * qi + 1] = pt.y;
    }
}

// ------------------------------------------------------------ overlap / line
// query_device(bvh, overlaps(box), out, max) (query.cuh:176-231) and
// query_device(bvh, line_intersect(line), element_test, out, max)
// (query.cuh:12-77): every primitive whose leaf box meets the query box
// (aabb.cuh:54-82) or the infinite line (slab test, aabb.cuh:330-364; the
// element test is the scene's ray test without t >= 0), emitted in the
// reference's order: pop a node, test its left then its right child, emit a
// leaf at once, push an internal child.  The first max_out ids (and line
// parameters) land in the query's row; count = all found.  The node ids are
// the reference's, so the order is the reference's bit for bit.
template <int BS, int SSZ>
struct IdStack {
    uint32_t* lds; // this lane's column: entry s at lds[s * BS]
    int sp;
    uint32_t ov[OVF];
    bool overflow;
    __device__ __forceinline__ void init(uint32_t* base)
    {
        lds = base;
        sp = 0;
        overflow = false;
    }
    __device__ __forceinline__ void push(uint32_t n)
    {
        const int s = sp % SSZ;
        if (sp >= SSZ) {
            if (sp - SSZ >= OVF) {
                overflow = true;
                return;
            }
            ov[sp - SSZ] = lds[s * BS];
        }
        lds[s * BS] = n;
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop()
    {
        --sp;
        const int s = sp % SSZ;
        const uint32_t n = lds[s * BS];
        if (sp >= SSZ)
            lds[s * BS] = ov[sp - SSZ];
        return n;
    }
};

template <int DIM>
__device__ __forceinline__ bool overlap_box(const box3& b, const box3& q)
{
    if (b.hi.x < q.lo.x || q.hi.x < b.lo.x || b.hi.y < q.lo.y || q.hi.y < b.lo.y)
        return false;
    return DIM == 2 || !(b.hi.z < q.lo.z || q.hi.z < b.lo.z);
}

template <int DIM>
__device__ __forceinline__ bool line_box(const box3& b, v3 o, v3 inv)
{
    float t1 = (b.lo.x - o.x) * inv.x, t2 = (b.hi.x - o.x) * inv.x;
    float tmin = fmn(t1, t2), tmax = fmx(t1, t2);
    t1 = (b.lo.y - o.y) * inv.y;
    t2 = (b.hi.y - o.y) * inv.y;
    tmin = fmx(tmin, fmn(t1, t2));
    tmax = fmn(tmax, fmx(t1, t2));
    if (DIM == 3) {
        t1 = (b.lo.z - o.z) * inv.z;
        t2 = (b.hi.z - o.z) * inv.z;
        tmin = fmx(tmin, fmn(t1, t2));
        tmax = fmn(tmax, fmx(t1, t2));
    }
    return tmax >= tmin;
}

template <int BS, int DIM, int KIND>
__global__ __launch_bounds__(BS) void k_generic(TreeView t, const float* __restrict__ qa, const float* __restrict__ qb,
                                                uint32_t n, uint32_t max_out, uint32_t* __restrict__ out_count,
                                                uint32_t* __restrict__ out_idx, float* __restrict__ out_t,
                                                uint32_t* __restrict__ err)
{
    __shared__ uint32_t stk[SS * BS];
    const uint32_t qi = xcd_remap(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (qi >= n)
        return;
    // KIND 0: the box [qa, qb]; KIND 1: the line through qa along qb
    const v3 a = load_point<DIM>(qa, qi), b = load_point<DIM>(qb, qi);
    box3 qbox;
    qbox.lo = a;
    qbox.hi = b;
    const v3 inv = mk(1.0f / b.x, 1.0f / b.y, DIM == 3 ? 1.0f / b.z : 0.0f); // line(o, dir) (aabb.cuh:288-294)
    uint32_t found = 0;
    const size_t row = (size_t)qi * max_out;
    auto hit_box = [&](const box3& bx) { return KIND == 0 ? overlap_box<DIM>(bx, qbox) : line_box<DIM>(bx, a, inv); };
    auto emit = [&](uint32_t j) {
        if (j >= t.n_leaves)
            return;
        const float4 la = t.tleaf[3 * j], lb = t.tleaf[3 * j + 1];
        float tt = 0.0f;
        bool ok = true;
        if (KIND == 1) {
            float uh, vh;
            if (DIM == 3) {
                const float4 lc = t.tleaf[3 * j + 2];
                ok = line_triangle(a, b, mk(la.x, la.y, la.z), mk(lb.x, lb.y, lb.z), mk(lc.x, lc.y, lc.z), tt, uh, vh);
            } else {
                ok = line_segment(a, b, mk(la.x, la.y, 0.0f), mk(lb.x, lb.y, 0.0f), tt, uh);
            }
        }
        if (!ok)
            return;
        if (found < max_out) {
            out_idx[row + found] = __float_as_uint(la.w);
            if (KIND == 1 && out_t)
                out_t[row + found] = tt;
        }
        ++found;
    };
    if (t.n_leaves == 1) {
        if (hit_box(ldbox2(t.nbox, 0)))
            emit(0);
    } else {
        IdStack<BS, SS> st;
        st.init(stk + threadIdx.x);
        st.push(0);
        // each internal node is pushed at most once
        for (uint32_t it = 0; it < t.n_leaves && st.sp > 0; ++it) {
            const NodeRec nr = load_node(t.tnode, st.pop());
            if (hit_box(nr.L)) {
                if (nr.lref & LEAF_BIT)
                    emit(nr.lref & ~LEAF_BIT);
                else
                    st.push(nr.lref);
            }
            if (hit_box(nr.R)) {
                if (nr.rref & LEAF_BIT)
                    emit(nr.rref & ~LEAF_BIT);
                else
                    st.push(nr.rref);
            }
        }
        if (st.overflow && err)
            atomicOr(err, 1u);
    }
    out_count[qi] = found;
}

// Morton code of each query point in the root box (clamped to it), used to
// visit queries in spatial order: neighbouring lanes then walk mostly the
// same nodes, which keeps a wave's memory accesses and branches together
template <int DIM>
__global__ __launch_bounds__(256) void k_query_codes(const float4* __restrict__ nbox, const float* __restrict__ q,
                                                     uint32_t n, uint32_t* __restrict__ codes,
                                                     uint32_t* __restrict__ idx)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const float4 lo = nbox[0], hi = nbox[1];
    const v3 p = load_point<DIM>(q, i);
    const float x = (p.x - lo.x) / (hi.x - lo.x), y = (p.y - lo.y) / (hi.y - lo.y);
    codes[i] = DIM == 3 ? morton3(mk(x, y, (p.z - lo.z) / (hi.z - lo.z))) : morton2(x, y);
    idx[i] = i;
}

inline uint32_t nblocks(uint32_t n, uint32_t bs) { return (n + bs - 1) / bs; }

} // namespace

#define SNCH_DISPATCH_DIM(dim, ...) \
    do {                            \
        if ((dim) == 3) {           \
            constexpr int DIM = 3;  \
            __VA_ARGS__;            \
        } else {                    \
            constexpr int DIM = 2;  \
            __VA_ARGS__;            \
        }                           \
    } while (0)

hipError_t query_order(const TreeView& t, int dim, const float* q, uint32_t n, QueryScratch& w, hipStream_t st,
                       const uint32_t** perm)
{
    *perm = nullptr;
    if (n < 2)
        return hipSuccess;
    hipError_t e;
    if ((e = w.codes.ensure(sizeof(uint32_t) * n)) != hipSuccess || (e = w.idx.ensure(sizeof(uint32_t) * n)) ||
        (e = w.codes_alt.ensure(sizeof(uint32_t) * n)) || (e = w.idx_alt.ensure(sizeof(uint32_t) * n)))
        return e;
    SNCH_DISPATCH_DIM(dim, hipLaunchKernelGGL(k_query_codes<DIM>, dim3(nblocks(n, 256)), dim3(256), 0, st, t.nbox,
                                              q, n, w.codes.as<uint32_t>(), w.idx.as<uint32_t>()));
    bool in_alt = false;
    if ((e = radix_sort_pairs(w.codes.as<uint32_t>(), w.idx.as<uint32_t>(), w.codes_alt.as<uint32_t>(),
                              w.idx_alt.as<uint32_t>(), n, 30, w.hist, st, &in_alt)) != hipSuccess)
        return e;
    *perm = in_alt ? w.idx_alt.as<uint32_t>() : w.idx.as<uint32_t>();
    return hipGetLastError();
}

size_t query_order_scratch_bytes(uint32_t n)
{
    return sizeof(uint32_t) * 4 * (size_t)n + radix_sort_scratch_bytes(n, 4);
}

hipError_t query_order_in(const TreeView& t, int dim, const float* q, uint32_t n, uint32_t* words, hipStream_t st,
                          const uint32_t** perm, uint32_t* err)
{
    *perm = nullptr;
    if (n < 2)
        return hipSuccess;
    uint32_t *codes = words, *idx = words + n, *codes_alt = words + 2 * (size_t)n, *idx_alt = words + 3 * (size_t)n;
    SNCH_DISPATCH_DIM(dim, hipLaunchKernelGGL(k_query_codes<DIM>, dim3(nblocks(n, 256)), dim3(256), 0, st, t.nbox,
                                              q, n, codes, idx));
    bool in_alt = false;
    const hipError_t e = radix_sort_pairs_in(codes, idx, codes_alt, idx_alt, n, 30, words + 4 * (size_t)n, st,
                                             &in_alt, err);
    if (e != hipSuccess)
        return e;
    *perm = in_alt ? idx_alt : idx;
    return hipGetLastError();
}

template <int BS, int DIM, int SSZ>
static void launch_cp_cfg(const TreeView& t, const float* q, uint32_t n, uint32_t* idx, float* dist,
                          const uint32_t* perm, unsigned long long* stats, hipStream_t st)
{
    if (stats)
        hipLaunchKernelGGL((k_closest_point<BS, true, DIM, SSZ>), dim3(nblocks(n, BS)), dim3(BS), 0, st, t, q, n, idx,
                           dist, perm, traversal_error_word(), stats);
    else
        hipLaunchKernelGGL((k_closest_point<BS, false, DIM, SSZ>), dim3(nblocks(n, BS)), dim3(BS), 0, st, t, q, n,
                           idx, dist, perm, traversal_error_word(), stats);
}

template <int BS, int DIM>
static void launch_cp_bs(const TreeView& t, const float* q, uint32_t n, uint32_t* idx, float* dist, uint32_t depth,
                         const uint32_t* perm, unsigned long long* stats, hipStream_t st)
{
    if (DIM == 3 && depth == 8)
        launch_cp_cfg<BS, DIM, 8>(t, q, n, idx, dist, perm, stats, st);
    else
        launch_cp_cfg<BS, DIM, 16>(t, q, n, idx, dist, perm, stats, st);
}

hipError_t launch_closest_point(const TreeView& t, int dim, const float* q, uint32_t n, uint32_t* idx, float* dist,
                                uint32_t block, uint32_t depth, const uint32_t* perm, uint64_t* stats, hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    auto* sp = reinterpret_cast<unsigned long long*>(stats);
    SNCH_DISPATCH_DIM(dim, {
        switch (block) {
        case 128: launch_cp_bs<128, DIM>(t, q, n, idx, dist, depth, perm, sp, st); break;
        case 256: launch_cp_bs<256, DIM>(t, q, n, idx, dist, depth, perm, sp, st); break;
        case 512: launch_cp_bs<512, DIM>(t, q, n, idx, dist, depth, perm, sp, st); break;
        default: launch_cp_bs<64, DIM>(t, q, n, idx, dist, depth, perm, sp, st); break;
        }
    });
    return hipGetLastError();
}

template <int BS, int DIM, int SSZ>
static void launch_sil_cfg(const TreeView& t, const float* q, uint32_t n, int flip, const float* seed, bool capped,
                           uint32_t* sil, float* dist, const uint32_t* perm, unsigned long long* stats,
                           uint32_t* err, hipStream_t st)
{
    const dim3 g(nblocks(n, BS)), b(BS);
#define SNCH_SIL(S_, M_, ...) \
    hipLaunchKernelGGL((k_closest_silhouette<BS, S_, DIM, M_, SSZ>), g, b, 0, st, __VA_ARGS__)
    if (!seed) {
        if (stats)
            SNCH_SIL(true, 0, t, q, n, n, flip, seed, sil, dist, perm, err, stats);
        else
            SNCH_SIL(false, 0, t, q, n, n, flip, seed, sil, dist, perm, err, stats);
    } else if (capped) {
        if (stats)
            SNCH_SIL(true, 2, t, q, n, n, flip, seed, sil, dist, perm, err, stats);
        else
            SNCH_SIL(false, 2, t, q, n, n, flip, seed, sil, dist, perm, err, stats);
    } else {
        if (stats)
            SNCH_SIL(true, 1, t, q, n, n, flip, seed, sil, dist, perm, err, stats);
        else
            SNCH_SIL(false, 1, t, q, n, n, flip, seed, sil, dist, perm, err, stats);
    }
#undef SNCH_SIL
}

template <int BS, int DIM>
static void launch_sil_bs(const TreeView& t, const float* q, uint32_t n, int flip, const float* seed, bool capped,
                          uint32_t* sil, float* dist, uint32_t depth, const uint32_t* perm,
                          unsigned long long* stats, uint32_t* err, hipStream_t st)
{
    if (DIM == 2 || depth == 16) // scene<2> uses the default depth only
        launch_sil_cfg<BS, DIM, 16>(t, q, n, flip, seed, capped, sil, dist, perm, stats, err, st);
    else if (depth == 8)
        launch_sil_cfg<BS, DIM, 8>(t, q, n, flip, seed, capped, sil, dist, perm, stats, err, st);
    else if (depth == 24)
        launch_sil_cfg<BS, DIM, 24>(t, q, n, flip, seed, capped, sil, dist, perm, stats, err, st);
    else
        launch_sil_cfg<BS, DIM, 32>(t, q, n, flip, seed, capped, sil, dist, perm, stats, err, st);
}

hipError_t launch_closest_silhouette(const TreeView& t, int dim, const float* q, uint32_t n, int flip,
                                     const float* seed, bool capped, uint32_t* sil, float* dist, uint32_t block,
                                     uint32_t depth, const uint32_t* perm, uint64_t* stats, hipStream_t st,
                                     uint32_t* err)
{
    if (n == 0)
        return hipSuccess;
    auto* sp = reinterpret_cast<unsigned long long*>(stats);
    if (!err)
        err = traversal_error_word();
    SNCH_DISPATCH_DIM(dim, {
        switch (block) {
        case 128: launch_sil_bs<128, DIM>(t, q, n, flip, seed, capped, sil, dist, depth, perm, sp, err, st); break;
        case 256: launch_sil_bs<256, DIM>(t, q, n, flip, seed, capped, sil, dist, depth, perm, sp, err, st); break;
        case 512: launch_sil_bs<512, DIM>(t, q, n, flip, seed, capped, sil, dist, depth, perm, sp, err, st); break;
        default: launch_sil_bs<64, DIM>(t, q, n, flip, seed, capped, sil, dist, depth, perm, sp, err, st); break;
        }
    });
    return hipGetLastError();
}

template <int DIM, bool STATS>
static void launch_sil_requeue_cfg(const TreeView& t, const float* q, uint32_t n, int flip, const float* seed,
                                   uint32_t* sil, float* dist, const uint32_t* perm, unsigned long long* stats,
                                   uint32_t* ctr, uint32_t* queue, uint32_t n_cu, uint32_t* err, hipStream_t st)
{
    // as many one-wave blocks as are resident at once; more would only wait
    static int per_cu = 0;
    if (per_cu == 0 &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sil_requeue<16, STATS, DIM>, 64, 0) != hipSuccess)
        per_cu = 1;
    const uint32_t grid = std::max(1u, std::min(n_cu * (uint32_t)std::max(per_cu, 1), (n + 63) / 64));
    hipLaunchKernelGGL((k_sil_requeue<16, STATS, DIM>), dim3(grid), dim3(64), 0, st, t, q, n, flip, seed, sil, dist,
                       perm, ctr, queue, err, stats);
}

hipError_t launch_closest_silhouette_requeue(const TreeView& t, int dim, const float* q, uint32_t n, int flip,
                                             const float* seed, uint32_t* sil, float* dist, const uint32_t* perm,
                                             uint64_t* stats, uint32_t* ctr, uint32_t* queue, uint32_t n_cu,
                                             hipStream_t st, uint32_t* err)
{
    if (n == 0)
        return hipSuccess;
    auto* sp = reinterpret_cast<unsigned long long*>(stats);
    if (!err)
        err = traversal_error_word();
    hipLaunchKernelGGL(k_requeue_init, dim3(std::min(nblocks(n, 256), 1024u)), dim3(256), 0, st, ctr, queue, 8 * n);
    SNCH_DISPATCH_DIM(dim, {
        if (sp)
            launch_sil_requeue_cfg<DIM, true>(t, q, n, flip, seed, sil, dist, perm, sp, ctr, queue, n_cu, err, st);
        else
            launch_sil_requeue_cfg<DIM, false>(t, q, n, flip, seed, sil, dist, perm, sp, ctr, queue, n_cu, err, st);
    });
    return hipGetLastError();
}

template <int DIM, bool STATS, bool CAPPED>
static void launch_sil_persist_cfg(const TreeView& t, const float* q, uint32_t n, int flip, const float* seed,
                                   uint32_t* sil, float* dist, const uint32_t* perm, unsigned long long* stats,
                                   uint32_t* ctr, uint32_t n_cu, uint32_t* err, hipStream_t st)
{
    // as many one-wave blocks as are resident at once; more would only wait
    static int per_cu = 0;
    if (per_cu == 0 && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sil_persist<16, STATS, DIM, CAPPED>,
                                                                     64, 0) != hipSuccess)
        per_cu = 1;
    const uint32_t grid = std::max(1u, std::min(n_cu * (uint32_t)std::max(per_cu, 1), (n + 63) / 64));
    hipLaunchKernelGGL((k_sil_persist<16, STATS, DIM, CAPPED>), dim3(grid), dim3(64), 0, st, t, q, n, flip, seed, sil,
                       dist, perm, ctr, err, stats);
}

hipError_t launch_closest_silhouette_persist(const TreeView& t, int dim, const float* q, uint32_t n, int flip,
                                             const float* seed, uint32_t* sil, float* dist, const uint32_t* perm,
                                             uint64_t* stats, uint32_t* ctr, uint32_t n_cu, hipStream_t st,
                                             uint32_t* err)
{
    if (n == 0)
        return hipSuccess;
    auto* sp = reinterpret_cast<unsigned long long*>(stats);
    if (!err)
        err = traversal_error_word();
    hipLaunchKernelGGL(k_requeue_init, dim3(1), dim3(256), 0, st, ctr, nullptr, 0u); // the chunk counters
    SNCH_DISPATCH_DIM(dim, {
        if (sp && seed)
            launch_sil_persist_cfg<DIM, true, true>(t, q, n, flip, seed, sil, dist, perm, sp, ctr, n_cu, err, st);
        else if (sp)
            launch_sil_persist_cfg<DIM, true, false>(t, q, n, flip, seed, sil, dist, perm, sp, ctr, n_cu, err, st);
        else if (seed)
            launch_sil_persist_cfg<DIM, false, true>(t, q, n, flip, seed, sil, dist, perm, sp, ctr, n_cu, err, st);
        else
            launch_sil_persist_cfg<DIM, false, false>(t, q, n, flip, seed, sil, dist, perm, sp, ctr, n_cu, err, st);
    });
    return hipGetLastError();
}

hipError_t launch_ray(const TreeView& t, int dim, const float* o, const float* d, const float* tmax, uint32_t n,
                      int any_hit, uint8_t* hit, float* tout, float* uv, uint32_t* idx, uint64_t* stats,
                      hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    auto* sp = reinterpret_cast<unsigned long long*>(stats);
    SNCH_DISPATCH_DIM(dim, {
        if (stats)
            hipLaunchKernelGGL((k_ray<256, true, DIM>), dim3(nblocks(n, 256)), dim3(256), 0, st, t, o, d, tmax, n,
                               any_hit, hit, tout, uv, idx, traversal_error_word(), sp);
        else
            hipLaunchKernelGGL((k_ray<256, false, DIM>), dim3(nblocks(n, 256)), dim3(256), 0, st, t, o, d, tmax, n,
                               any_hit, hit, tout, uv, idx, traversal_error_word(), sp);
    });
    return hipGetLastError();
}

hipError_t launch_sample(const TreeView& t, int dim, const float4* V, const int32_t* F, const float* c,
                         const float* r, const float* u, uint32_t n, int32_t* idx, float* pdf, float* pt,
                         hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    SNCH_DISPATCH_DIM(dim, hipLaunchKernelGGL(k_sample<DIM>, dim3(nblocks(n, 256)), dim3(256), 0, st, t, V, F, c, r,
                                              u, n, idx, pdf, pt));
    return hipGetLastError();
}

hipError_t launch_generic(const TreeView& t, int dim, int kind, const float* qa, const float* qb, uint32_t n,
                          uint32_t max_out, uint32_t* count, uint32_t* idx, float* tout, hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    uint32_t* err = traversal_error_word();
    SNCH_DISPATCH_DIM(dim, {
        if (kind == 0)
            hipLaunchKernelGGL((k_generic<64, DIM, 0>), dim3(nblocks(n, 64)), dim3(64), 0, st, t, qa, qb, n, max_out,
                               count, idx, tout, err);
        else
            hipLaunchKernelGGL((k_generic<64, DIM, 1>), dim3(nblocks(n, 64)), dim3(64), 0, st, t, qa, qb, n, max_out,
                               count, idx, tout, err);
    });
    return hipGetLastError();
}

hipError_t launch_sample_on_object(int dim, const float4* V, const int32_t* F, uint32_t nF, const int32_t* idx,
                                   const float* uv, uint32_t n, float* pt, hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    SNCH_DISPATCH_DIM(dim, hipLaunchKernelGGL(k_sample_on_object<DIM>, dim3(nblocks(n, 256)), dim3(256), 0, st, V,
                                              F, nF, idx, uv, n, pt));
    return hipGetLastError();
}

template <int BS, int SSZ, int KIND, bool SEEDED>
static hipError_t launch_batch_cfg(const BatchTable& bt, const float* q, int flip, const float* seed,
                                   uint32_t* out_idx, float* out_dist, uint32_t n_cu, hipStream_t st)
{
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch<BS, SSZ, KIND, SEEDED>, BS, 0);
    if (e != hipSuccess)
        return e;
    const uint32_t waves = n_cu * (uint32_t)std::max(per_cu, 1) * (BS / 64);
    // no more waves than chunks
    const uint32_t grid = std::max(1u, std::min(n_cu * (uint32_t)std::max(per_cu, 1),
                                                (std::min(waves, bt.total_chunks) + BS / 64 - 1) / (BS / 64)));
    hipLaunchKernelGGL((k_batch<BS, SSZ, KIND, SEEDED>), dim3(grid), dim3(BS), 0, st, bt, q, flip, seed, out_idx,
                       out_dist, traversal_error_word());
    return hipGetLastError();
}

template <int BS, int KIND, bool SEEDED>
static hipError_t launch_batch_bs(const BatchTable& bt, uint32_t ssz, const float* q, int flip, const float* seed,
                                  uint32_t* out_idx, float* out_dist, uint32_t n_cu, hipStream_t st)
{
    switch (ssz) {
    case 8: return launch_batch_cfg<BS, 8, KIND, SEEDED>(bt, q, flip, seed, out_idx, out_dist, n_cu, st);
    case 24: return launch_batch_cfg<BS, 24, KIND, SEEDED>(bt, q, flip, seed, out_idx, out_dist, n_cu, st);
    case 32: return launch_batch_cfg<BS, 32, KIND, SEEDED>(bt, q, flip, seed, out_idx, out_dist, n_cu, st);
    default: return launch_batch_cfg<BS, 16, KIND, SEEDED>(bt, q, flip, seed, out_idx, out_dist, n_cu, st);
    }
}

template <int KIND, bool SEEDED>
static hipError_t launch_batch_kind(const BatchTable& bt, uint32_t bs, uint32_t ssz, const float* q, int flip,
                                    const float* seed, uint32_t* out_idx, float* out_dist, uint32_t n_cu,
                                    hipStream_t st)
{
    switch (bs) {
    case 64: return launch_batch_bs<64, KIND, SEEDED>(bt, ssz, q, flip, seed, out_idx, out_dist, n_cu, st);
    case 128: return launch_batch_bs<128, KIND, SEEDED>(bt, ssz, q, flip, seed, out_idx, out_dist, n_cu, st);
    case 512: return launch_batch_bs<512, KIND, SEEDED>(bt, ssz, q, flip, seed, out_idx, out_dist, n_cu, st);
    default: return launch_batch_bs<256, KIND, SEEDED>(bt, ssz, q, flip, seed, out_idx, out_dist, n_cu, st);
    }
}

hipError_t launch_batch(const snch_scene* s, int kind, const float* q, const uint32_t* q_off_host, int flip,
                        uint32_t* out_idx, float* out_dist, const BatchConfig& cfg, hipStream_t st,
                        const float* seed)
{
    const uint32_t M = s->n_meshes;
    BatchState& b = s->batch;
    const uint32_t n = q_off_host[M];
    if (n == 0)
        return hipSuccess;
    const uint32_t qpw = cfg.queries_per_wave;
    // schedule: meshes with more leaves (deeper trees, costlier queries) first
    std::vector<uint32_t> order(M);
    for (uint32_t m = 0; m < M; ++m)
        order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
        return s->mesh_off[a + 1] - s->mesh_off[a] > s->mesh_off[c + 1] - s->mesh_off[c];
    });
    // host table: qoff (M + 1) | sched (M) | cstart (M + 1)
    b.host_tab.assign(3 * (size_t)M + 2, 0);
    uint32_t* h_qoff = b.host_tab.data();
    uint32_t* h_sched = h_qoff + M + 1;
    uint32_t* h_cstart = h_sched + M;
    for (uint32_t m = 0; m <= M; ++m)
        h_qoff[m] = q_off_host[m];
    uint32_t ns = 0, total = 0;
    for (uint32_t k = 0; k < M; ++k) {
        const uint32_t m = order[k], cnt = q_off_host[m + 1] - q_off_host[m];
        if (cnt == 0)
            continue;
        h_sched[ns] = m;
        h_cstart[ns] = total;
        total += (cnt + qpw - 1) / qpw;
        ++ns;
    }
    h_cstart[ns] = total;
    hipError_t e;
    if ((e = b.tab.ensure(sizeof(uint32_t) * b.host_tab.size())) != hipSuccess ||
        (e = b.counter.ensure(64)) != hipSuccess)
        return e;
    uint32_t* d_tab = b.tab.as<uint32_t>();
    if ((e = hipMemcpyAsync(d_tab, b.host_tab.data(), sizeof(uint32_t) * b.host_tab.size(), hipMemcpyHostToDevice,
                            st)) != hipSuccess)
        return e;
    BatchTable bt;
    bt.views = b.views.as<TreeView>();
    bt.qoff = d_tab;
    bt.sched = d_tab + M + 1;
    bt.cstart = d_tab + 2 * M + 1;
    bt.perm = nullptr;
    bt.counter = b.counter.as<uint32_t>();
    bt.n_sched = ns;
    bt.total_chunks = total;
    bt.qpw = qpw;
    if (cfg.coherent) {
        QueryScratch& w = b.order;
        if ((e = w.codes.ensure(sizeof(uint32_t) * n)) != hipSuccess ||
            (e = w.idx.ensure(sizeof(uint32_t) * n)) != hipSuccess ||
            (e = w.codes_alt.ensure(sizeof(uint32_t) * n)) != hipSuccess ||
            (e = w.idx_alt.ensure(sizeof(uint32_t) * n)) != hipSuccess ||
            (e = b.qmesh.ensure(sizeof(uint32_t) * n)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_query_mesh, dim3(nblocks(n, 256)), dim3(256), 0, st, d_tab, M, n,
                           b.qmesh.as<uint32_t>());
        hipLaunchKernelGGL(k_batch_codes, dim3(nblocks(n, 256)), dim3(256), 0, st, bt.views, b.qmesh.as<uint32_t>(),
                           q, n, w.codes.as<uint32_t>(), w.idx.as<uint32_t>());
        int mesh_bits = 0;
        while (mesh_bits < 32 && ((M - 1) >> mesh_bits) != 0)
            ++mesh_bits;
        bool in_alt = false;
        if ((e = radix_sort_pairs(w.codes.as<uint32_t>(), w.idx.as<uint32_t>(), w.codes_alt.as<uint32_t>(),
                                  w.idx_alt.as<uint32_t>(), n, 30, w.hist, st, &in_alt,
                                  M > 1 ? b.qmesh.as<uint32_t>() : nullptr, M > 1 ? mesh_bits : 0)) != hipSuccess)
            return e;
        bt.perm = in_alt ? w.idx_alt.as<uint32_t>() : w.idx.as<uint32_t>();
    }
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, bt.counter);
    const uint32_t n_cu = s->ctx && s->ctx->n_cu > 0 ? (uint32_t)s->ctx->n_cu : 256u;
    if (kind == 0)
        return launch_batch_kind<0, false>(bt, cfg.block_size, cfg.stack_depth, q, flip, nullptr, out_idx, out_dist,
                                           n_cu, st);
    if (seed)
        return launch_batch_kind<1, true>(bt, cfg.block_size, cfg.stack_depth, q, flip, seed, out_idx, out_dist, n_cu,
                                          st);
    return launch_batch_kind<1, false>(bt, cfg.block_size, cfg.stack_depth, q, flip, nullptr, out_idx, out_dist, n_cu,
                                       st);
}

} // namespace snch
