// Device layout of a built tree: the records every traversal reads.  Shared
// by the library's kernels and by include/snch_lbvh/device.hpp (users' own
// HIP kernels), so it holds no host-side types.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace snch {

// ---------------------------------------------------------------------------
// HBM layout of a built tree (N leaves, N-1 internal nodes, E silhouette edges)
//
//   tnode[i]  4 x float4 = 64 B per internal node i (64-B aligned): the boxes
//             of BOTH children plus their references, so one visit is one
//             64-B load:
//               [0] L.lo.x L.lo.y L.lo.z L.hi.x
//               [1] L.hi.y L.hi.z R.lo.x R.lo.y
//               [2] R.lo.z R.hi.x R.hi.y R.hi.z
//               [3] ref(L) ref(R) run(L) run(R)
//             ref(c) = c for an internal child, LEAF_BIT | j for the leaf in
//             sorted position j.  run(c) = INVALID, or, when the tree is
//             built with leaf size k > 1 and internal child c covers at most
//             k leaves, (count - 1) << 27 | first: the traversal queries then
//             treat c as one leaf holding sorted leaves [first, first + count).
//   tcone[i]  4 x float4 = 64 B per internal node: the children's normal
//             cones, read only by the silhouette query:
//               [0] L.axis L.half_angle  [1] R.axis R.half_angle
//               [2] L.radius R.radius 0 0
//               [3] sin, cos of L.half_angle, sin, cos of R.half_angle
//   tleaf[j]  3 x float4 = 48 B per leaf, in Morton (sorted) order:
//               [k] vertex k xyz, w = (original triangle, first owned edge,
//               owned edge count) for k = 0, 1, 2
//   tedge[e]  4 x float4 = 64 B per silhouette edge, in edge-id order (the
//             edges owned by one triangle have consecutive ids):
//               [0] pa xyz, edge id  [1] pb xyz, flags (bit0 has face 0,
//               bit1 has face 1)  [2] n0 xyz  [3] n1 xyz (unit face normals)
// The reference layout (node{parent,left,right,object}, aabb, cone per node;
// bvh.cuh:27-33) is kept alongside for export and for the refit's own use.
//
// A scene may hold M meshes (a batch).  Mesh m owns primitives (and sorted
// leaf positions) [off, off + n), silhouette records [eoff, ...), nodes
// [nb, nb + 2n - 1) with nb = 2 off - m (its n - 1 internal nodes first, root
// at nb), and internal records [off - m, off - m + n - 1).  A single mesh is
// the batch of one, with every offset zero.  Records hold mesh-local ids.
// ---------------------------------------------------------------------------
constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t RUN_SHIFT = 27;
constexpr uint32_t RUN_FIRST_MASK = (1u << RUN_SHIFT) - 1u;
constexpr uint32_t MAX_LEAF_SIZE = 8;
constexpr uint32_t MAX_MESH_PRIMS = 1u << 24; // node ids fit the 24-bit field of a packed stack entry

struct MeshDesc {
    uint32_t off, n, eoff, nb;
};

constexpr uint32_t TREE_FAST_CONE = 1u; // TreeView::flags: cone_overlap_fast may replace cone_overlap

struct TreeView {
    uint32_t n_leaves;
    uint32_t flags;
    uint32_t qbase; // traversal-stack key quantisation base (bits >> 20 of D^2, less 247)
    uint32_t pad_;
    const float4* __restrict__ tnode;
    const float4* __restrict__ tcone;
    const float4* __restrict__ tleaf;
    const float4* __restrict__ tedge;
    const float4* __restrict__ nbox;  // own box per node, reference numbering (root = 0)
    const float4* __restrict__ ncone; // own cone per node, reference numbering
    // optional device-side query count (<= the launch's n): lets a captured
    // graph run over a data-dependent number of queries with a fixed grid
    const uint32_t* n_dev = nullptr;
};

} // namespace snch
