// snch_lbvh/scene.hpp -- header-only C++ front end over the C ABI (snch.h),
// mirroring the reference's lbvh::scene<3> (scene.cuh:705-1268) and its query
// descriptors (predicator.cuh:138-207) for host code.
//
//   lbvh::scene<3> s(vb, ve, ib, ie);   // scene<3>(...) (scene.cuh:1129-1133)
//   s.compute_silhouettes();            // scene.cuh:1167-1204 (done at construction)
//   s.build_bvh();                      // scene.cuh:1205-1229
//   s.query(lbvh::nearest_silhouette(), d_q, n, d_edge, d_dist, stream);
//
// All query pointers are device pointers; launches are asynchronous on the
// given hipStream_t (passed as void*).  Errors throw std::runtime_error with
// the library's message.
#pragma once

#include <cstdint>
#include <iterator>
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

// query descriptors (predicator.cuh); the batched form carries only flags
struct nearest {};
struct nearest_silhouette {
    bool flip_normal_orientation = false;
};
template <bool TestOnly = false>
struct ray_intersect {
    static constexpr bool test_only = TestOnly;
};

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

template <unsigned int dim>
class scene;

template <>
class scene<3> {
public:
    // VerticesInputIterator over float[3]-like values, IndicesInputIterator
    // over int[3]-like values (x, y, z members, as float3 / int3)
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
    ~scene() { snch_scene_destroy(s_); }
    scene(const scene&) = delete;
    scene& operator=(const scene&) = delete;

    void compute_silhouettes() {}

    void build_bvh(bool reference_radius = false, void* stream = nullptr)
    {
        snch_build_opts o{reference_radius ? SNCH_BUILD_REFERENCE_RADIUS : 0u, 1u};
        snch_check(snch_scene_build(s_, &o, stream));
    }

    // query_device(bvh, nearest(p), distance_calculator) for n points
    void query(nearest, const float* d_points, uint32_t n, uint32_t* d_idx, float* d_dist,
               void* stream = nullptr) const
    {
        snch_check(snch_closest_point(s_, d_points, n, d_idx, d_dist, stream));
    }
    // query_device(bvh, nearest_silhouette(p, flip), silhouette_distance_calculator)
    void query(nearest_silhouette q, const float* d_points, uint32_t n, uint32_t* d_edge, float* d_dist,
               void* stream = nullptr, const float* d_seed_radius = nullptr) const
    {
        snch_check(snch_closest_silhouette(s_, d_points, n, q.flip_normal_orientation ? 1 : 0, d_seed_radius,
                                           d_edge, d_dist, stream));
    }
    // query_device(bvh, ray_intersect<TestOnly>(ray, max_dist), intersect_test)
    template <bool TestOnly>
    void query(ray_intersect<TestOnly>, const float* d_origins, const float* d_dirs, const float* d_tmax, uint32_t n,
               uint8_t* d_hit, float* d_t, float* d_uv, uint32_t* d_idx, void* stream = nullptr) const
    {
        snch_check(snch_ray(s_, d_origins, d_dirs, d_tmax, n, TestOnly ? 1 : 0, d_hit, d_t, d_uv, d_idx, stream));
    }
    // sample_object_in_sphere + sample_on_object (sample.cuh:7-92)
    void sample_in_sphere(const float* d_centers, const float* d_radii, const float* d_u3, uint32_t n, int32_t* d_idx,
                          float* d_pdf, float* d_point, void* stream = nullptr) const
    {
        snch_check(snch_sample_in_sphere(s_, d_centers, d_radii, d_u3, n, d_idx, d_pdf, d_point, stream));
    }

    // reference-layout copy of the tree (bvh.cuh:615-622)
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

    snch_scene* handle() const { return s_; }

private:
    snch_scene* s_ = nullptr;
};

} // namespace lbvh
