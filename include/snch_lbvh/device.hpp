// HIP device API for users' own kernels: the reference's query_device and
// sample_object_in_sphere called one thread per query from device code
// (query.cuh:79-423, sample.cuh:7-92; the pattern of the walk-on-stars
// solvers built on it).  Compile the user's kernel with hipcc for gfx950 and
// -ffp-contract=off (the library's own setting, which the bit-for-bit
// agreement with its batched queries relies on).
//
//   snch_device_view dv;                  // host: after snch_scene_build
//   snch_scene_device_view(scene, 0, &dv);
//   lbvh::device::bvh_view bvh(dv);       // pass by value to the kernel
//   ...
//   __global__ void k(lbvh::device::bvh_view bvh, ...) {
//       auto r = lbvh::device::query_nearest_silhouette(bvh, p, false);
//   }
//
// Each call is a plain per-thread traversal (nearer child first, a private
// stack of 64 entries, which a Karras tree's depth never exceeds) with the
// library's pruning tests; ties go to the smallest index, so the answers
// equal the library's batched queries bit for bit (closest silhouette: up to
// ties at distance exactly 0).  The batched entry points of snch.h are the
// fast path; these exist for code that cannot batch its queries.
#pragma once
#include <hip/hip_runtime.h>

#include "../snch.h"
#include "../../snch_lbvh_amd/csrc/traverse.hpp"

namespace lbvh {
namespace device {

// a mesh's tree as device code reads it (trivially copyable: pass by value)
struct bvh_view {
    snch::TreeView tree;
    const float4* V;
    const int32_t* F;
    uint32_t dim;

    bvh_view() = default;
    explicit bvh_view(const snch_device_view& d)
    {
        tree = snch::TreeView{};
        tree.n_leaves = d.n_leaves;
        tree.flags = d.flags;
        tree.qbase = d.qbase;
        tree.tnode = static_cast<const float4*>(d.tnode);
        tree.tcone = static_cast<const float4*>(d.tcone);
        tree.tleaf = static_cast<const float4*>(d.tleaf);
        tree.tedge = static_cast<const float4*>(d.tedge);
        tree.nbox = static_cast<const float4*>(d.nbox);
        tree.ncone = static_cast<const float4*>(d.ncone);
        V = static_cast<const float4*>(d.V);
        F = d.F;
        dim = d.dim;
    }
};

// query_device(bvh, nearest(p), ...) -> (primitive index, distance)
struct nearest_result {
    uint32_t idx; // 0xFFFFFFFF when the tree is empty (query.cuh:254)
    float dist;
};
// query_device(bvh, nearest_silhouette(p, flip), ...) -> (silhouette id, distance)
struct silhouette_result {
    uint32_t idx; // edge id (3D) or vertex id (2D); 0xFFFFFFFF with dist = +inf when none
    float dist;
};
// query_device(bvh, ray_intersect<false>(ray, max_dist), ...)
struct ray_result {
    bool hit;
    float t;
    float u, v;   // barycentric (3D); the segment parameter in u (2D)
    uint32_t idx; // primitive index
};
// sample_object_in_sphere(bvh, sphere_intersect(s), ...) + the sampled point
struct sample_result {
    int32_t idx; // -1 when nothing was sampled (sample.cuh:45)
    float pdf;   // 0 when nothing was sampled
    float3 point;
};

namespace detail {
__device__ inline snch::v3 v(float3 p) { return snch::mk(p.x, p.y, p.z); }
__device__ inline snch::v3 v(float2 p) { return snch::mk(p.x, p.y, 0.0f); }
template <class P> struct dim_of;
template <> struct dim_of<float3> { static constexpr int value = 3; };
template <> struct dim_of<float2> { static constexpr int value = 2; };
} // namespace detail

// closest primitive (3D: float3 points on a scene<3>, 2D: float2 on a scene<2>)
template <class P>
__device__ inline nearest_result query_nearest(const bvh_view& b, P p)
{
    float best;
    uint32_t idx;
    snch::dev_nearest<detail::dim_of<P>::value>(b.tree, detail::v(p), best, idx);
    return {idx, snch::dm::sqrt_(best)};
}

// closest silhouette; flip inverts the viewing side (nearest_silhouette(p, flip))
template <class P>
__device__ inline silhouette_result query_nearest_silhouette(const bvh_view& b, P p, bool flip)
{
    float best;
    uint32_t idx;
    snch::dev_nearest_silhouette<detail::dim_of<P>::value>(b.tree, detail::v(p), flip, best, idx);
    return {idx, best};
}

// closest hit of the ray o + t d with t < max_dist
template <class P>
__device__ inline ray_result query_ray(const bvh_view& b, P o, P d, float max_dist)
{
    ray_result r;
    r.hit = snch::dev_ray<detail::dim_of<P>::value>(b.tree, detail::v(o), detail::v(d), max_dist, false, r.t, r.u,
                                                     r.v, r.idx);
    return r;
}

// any hit of the ray (ray_intersect<true>): stops at the first one found
template <class P>
__device__ inline bool query_any_hit(const bvh_view& b, P o, P d, float max_dist)
{
    float t, u, v;
    uint32_t idx;
    return snch::dev_ray<detail::dim_of<P>::value>(b.tree, detail::v(o), detail::v(d), max_dist, true, t, u, v, idx);
}

// SampleTriangleInSphere: a primitive inside the sphere (c, r) drawn by the
// reference's Green-weighted descent from u.x, and a point on it from u.y, u.z
template <class P>
__device__ inline sample_result sample_object_in_sphere(const bvh_view& b, P c, float r, float3 u)
{
    sample_result s;
    snch::v3 pt;
    snch::sample_one<detail::dim_of<P>::value>(b.tree, b.V, b.F, detail::v(c), r, u.x, u.y, u.z, s.idx, s.pdf, pt);
    s.point = make_float3(pt.x, pt.y, pt.z);
    return s;
}

} // namespace device
} // namespace lbvh
