// snch_lbvh/scene.hpp -- header-only C++ front end over the C ABI (snch.h),
// mirroring the reference's lbvh::scene<2|3> (scene.cuh:287-703, :705-1268)
// and its query descriptors (predicator.cuh:8-207) for host code.
//
//   lbvh::scene<3> s(vb, ve, ib, ie);   // scene<3>(...) (scene.cuh:1129-1133)
//   s.compute_silhouettes();            // scene.cuh:1167-1204 (done at construction)
//   s.build_bvh();                      // scene.cuh:1205-1229
//   s.query(lbvh::nearest_silhouette(d_points, false), n, d_edge, d_dist, stream);
//
// The reference answers one query per GPU thread from the caller's kernel
// (query_device, query.cuh); here one call answers n queries: a descriptor
// carries DEVICE pointers to the n queries' data (points, rays, spheres,
// boxes, lines), and the outputs are device arrays of n entries.  Launches
// are asynchronous on the given hipStream_t (passed as void*).  Errors throw
// std::runtime_error with the library's message, as the reference throws
// (scene.cuh:1250).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../snch.h"

namespace lbvh {

inline void snch_check(int rc)
{
    if (rc != SNCH_OK)
        throw std::runtime_error(std::string("snch: ") + snch_last_error());
}

// ---- query descriptors (predicator.cuh), each over n queries -------------
// nearest(p) (predicator.cuh:138-175): points, n x dim floats
struct query_nearest {
    const float* points = nullptr;
};
// nearest_silhouette(p, flip) (predicator.cuh:177-207); seed_radius (optional,
// n floats): a per-query pruning radius, e.g. a learned bound; the answers
// equal the unseeded ones
struct query_nearest_silhouette {
    const float* points = nullptr;
    bool flip_normal_orientation = false;
    const float* seed_radius = nullptr;
};
// ray_intersect<TestOnly>(ray(o, d), max_dist) (predicator.cuh:42-79): origins
// and directions n x dim, max_dist n floats (nullptr: unbounded)
template <bool TestOnly = false>
struct query_ray_intersect {
    static constexpr bool test_only = TestOnly;
    const float* origins = nullptr;
    const float* dirs = nullptr;
    const float* max_dist = nullptr;
};
// sphere_intersect(sphere(c, r)) (predicator.cuh:81-113): centres n x dim, radii n
struct query_sphere_intersect {
    const float* centers = nullptr;
    const float* radii = nullptr;
};
// overlaps(aabb) (predicator.cuh:115-136): boxes [lo, hi], n x dim floats each
struct query_overlap {
    const float* lo = nullptr;
    const float* hi = nullptr;
};
// line_intersect(line(o, d)) (predicator.cuh:8-40): origins and directions n x dim
struct query_line_intersect {
    const float* origins = nullptr;
    const float* dirs = nullptr;
};

inline query_nearest nearest(const float* d_points = nullptr) { return query_nearest{d_points}; }
inline query_nearest_silhouette nearest_silhouette(const float* d_points = nullptr, bool flip_normal_orientation = false,
                                                   const float* d_seed_radius = nullptr)
{
    return query_nearest_silhouette{d_points, flip_normal_orientation, d_seed_radius};
}
template <bool TestOnly = false>
inline query_ray_intersect<TestOnly> ray_intersect(const float* d_origins, const float* d_dirs,
                                                   const float* d_max_dist = nullptr)
{
    return query_ray_intersect<TestOnly>{d_origins, d_dirs, d_max_dist};
}
inline query_sphere_intersect sphere_intersect(const float* d_centers, const float* d_radii)
{
    return query_sphere_intersect{d_centers, d_radii};
}
inline query_overlap overlaps(const float* d_lo, const float* d_hi) { return query_overlap{d_lo, d_hi}; }
inline query_line_intersect line_intersect(const float* d_origins, const float* d_dirs)
{
    return query_line_intersect{d_origins, d_dirs};
}

class context {
public:
    explicit context(int device = 0) { snch_check(snch_context_create(device, &ctx_)); }
    ~context() { snch_context_destroy(ctx_); }
    context(const context&) = delete;
    context& operator=(const context&) = delete;
    snch_context* get() const { return ctx_; }

    static context& default_context()
    {
        static context c(0);
        return c;
    }

private:
    snch_context* ctx_ = nullptr;
};

namespace detail {

// what scene<2> and scene<3> share: the handle, the build and the queries
template <unsigned int dim>
class scene_base {
public:
    ~scene_base() { snch_scene_destroy(s_); }
    scene_base(const scene_base&) = delete;
    scene_base& operator=(const scene_base&) = delete;

    void compute_silhouettes() {} // scene.cuh:631-652 / :1167-1204: done at construction

    // build_bvh() (scene.cuh:673-701 / :1205-1229).  reference_radius: the
    // reference's cone radius instead of the enclosing one
    void build_bvh(bool reference_radius = false, void* stream = nullptr, uint32_t leaf_size = 1)
    {
        snch_build_opts o{reference_radius ? SNCH_BUILD_REFERENCE_RADIUS : 0u, leaf_size};
        snch_check(snch_scene_build(s_, &o, stream));
    }

    // query_device(bvh, nearest(p), distance_calculator) (query.cuh:238-318)
    void query(query_nearest q, uint32_t n, uint32_t* d_idx, float* d_dist, void* stream = nullptr) const
    {
        snch_check(snch_closest_point(s_, q.points, n, d_idx, d_dist, stream));
    }
    // query_device(bvh, nearest_silhouette(p, flip), silhouette_distance_calculator)
    // (query.cuh:325-423); d_sil (may be nullptr) also gets the silhouette id
    void query(query_nearest_silhouette q, uint32_t n, uint32_t* d_sil, float* d_dist, void* stream = nullptr) const
    {
        snch_check(snch_closest_silhouette(s_, q.points, n, q.flip_normal_orientation ? 1 : 0, q.seed_radius, d_sil,
                                           d_dist, stream));
    }
    // query_device(bvh, ray_intersect<TestOnly>(ray, max_dist), intersect_test)
    // (query.cuh:79-169); d_uv: 2 floats per ray in 3D (barycentric u, v), 1 in
    // 2D (segment parameter); outputs other than d_hit may be nullptr
    template <bool TestOnly>
    void query(query_ray_intersect<TestOnly> q, uint32_t n, uint8_t* d_hit, float* d_t = nullptr,
               float* d_uv = nullptr, uint32_t* d_idx = nullptr, void* stream = nullptr) const
    {
        snch_check(snch_ray(s_, q.origins, q.dirs, q.max_dist, n, TestOnly ? 1 : 0, d_hit, d_t, d_uv, d_idx, stream));
    }
    // query_device(bvh, overlaps(box), out, max) (query.cuh:176-231): the first
    // max_per_query ids of each query's row of d_idx, all counted in d_count
    void query(query_overlap q, uint32_t n, uint32_t max_per_query, uint32_t* d_count, uint32_t* d_idx,
               void* stream = nullptr) const
    {
        snch_check(snch_overlap(s_, q.lo, q.hi, n, max_per_query, d_count, d_idx, stream));
    }
    // query_device(bvh, line_intersect(line), element_test, out, max) (query.cuh:12-77)
    void query(query_line_intersect q, uint32_t n, uint32_t max_per_query, uint32_t* d_count, uint32_t* d_idx,
               float* d_t = nullptr, void* stream = nullptr) const
    {
        snch_check(snch_line_intersect(s_, q.origins, q.dirs, n, max_per_query, d_count, d_idx, d_t, stream));
    }
    // sample_object_in_sphere(bvh, sphere_intersect(s), ..., u) then
    // sample_on_object(...) (sample.cuh:7-92): d_u3 holds 3 uniforms per query
    // (descent, then the point on the primitive); d_point may be nullptr
    void sample_object_in_sphere(query_sphere_intersect q, const float* d_u3, uint32_t n, int32_t* d_idx,
                                 float* d_pdf, float* d_point = nullptr, void* stream = nullptr) const
    {
        snch_check(snch_sample_in_sphere(s_, q.centers, q.radii, d_u3, n, d_idx, d_pdf, d_point, stream));
    }
    // sample_on_object(bvh, idx, sample_on_object, u, v) (sample.cuh:7-21)
    void sample_on_object(const int32_t* d_idx, const float* d_uv, uint32_t n, float* d_point,
                          void* stream = nullptr) const
    {
        snch_check(snch_sample_on_object(s_, d_idx, d_uv, n, d_point, stream));
    }

    // reference-layout copy of the tree (bvh.cuh:615-622): nodes (parent,
    // left, right, object), boxes (lo xyz, hi xyz), cones (axis xyz, half
    // angle, radius)
    void export_tree(std::vector<uint32_t>& nodes, std::vector<float>& boxes, std::vector<float>& cones) const
    {
        uint32_t d, np, ns;
        snch_check(snch_scene_info(s_, &d, &np, &ns));
        const size_t nn = 2 * (size_t)np - 1;
        nodes.resize(4 * nn);
        boxes.resize(6 * nn);
        cones.resize(5 * nn);
        snch_check(snch_scene_export(s_, nodes.data(), boxes.data(), cones.data(), nullptr, nullptr));
    }
    uint32_t primitives_size() const
    {
        uint32_t d, np, ns;
        snch_check(snch_scene_info(s_, &d, &np, &ns));
        return np;
    }

    snch_scene* handle() const { return s_; }

protected:
    scene_base() = default;
    snch_scene* s_ = nullptr;
};

} // namespace detail

template <unsigned int dim>
class scene;

// scene<3>: triangles.  VIt iterates float3-like vertices (x, y, z members),
// IIt int3-like triangles (x, y, z vertex indices)
template <>
class scene<3> : public detail::scene_base<3> {
public:
    template <typename VIt, typename IIt>
    scene(VIt vb, VIt ve, IIt ib, IIt ie, context& ctx = context::default_context())
    {
        std::vector<float> V;
        std::vector<int32_t> F;
        for (; vb != ve; ++vb) {
            V.push_back(vb->x);
            V.push_back(vb->y);
            V.push_back(vb->z);
        }
        for (; ib != ie; ++ib) {
            F.push_back(ib->x);
            F.push_back(ib->y);
            F.push_back(ib->z);
        }
        snch_check(snch_scene_create_tri(ctx.get(), V.data(), (uint32_t)(V.size() / 3), F.data(),
                                         (uint32_t)(F.size() / 3), &s_));
    }
};

// scene<2>: line segments with silhouette vertices.  VIt iterates float2-like
// vertices (x, y), IIt int2-like segments (x, y vertex indices)
template <>
class scene<2> : public detail::scene_base<2> {
public:
    template <typename VIt, typename IIt>
    scene(VIt vb, VIt ve, IIt ib, IIt ie, context& ctx = context::default_context())
    {
        std::vector<float> V;
        std::vector<int32_t> S;
        for (; vb != ve; ++vb) {
            V.push_back(vb->x);
            V.push_back(vb->y);
        }
        for (; ib != ie; ++ib) {
            S.push_back(ib->x);
            S.push_back(ib->y);
        }
        snch_check(snch_scene_create_seg(ctx.get(), V.data(), (uint32_t)(V.size() / 2), S.data(),
                                         (uint32_t)(S.size() / 2), &s_));
    }
};

// the reference's free-function spelling (sample.cuh:23-32, :7-21)
template <unsigned int dim>
void sample_object_in_sphere(const scene<dim>& s, query_sphere_intersect q, const float* d_u3, uint32_t n,
                             int32_t* d_idx, float* d_pdf, float* d_point = nullptr, void* stream = nullptr)
{
    s.sample_object_in_sphere(q, d_u3, n, d_idx, d_pdf, d_point, stream);
}

} // namespace lbvh
