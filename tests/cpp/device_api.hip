// A user's own HIP kernel on the device API (include/snch_lbvh/device.hpp):
// one thread per query calls every query, the way the reference's users call
// query_device from their kernels.  Built as libdevice_api.so (linked against
// libsnch.so); tests/test_device_api.py compares its outputs with the
// library's batched queries.
#include <hip/hip_runtime.h>

#include "snch_lbvh/device.hpp"

namespace {

template <class P>
__device__ P load(const float* a, uint32_t i);
template <>
__device__ float3 load<float3>(const float* a, uint32_t i)
{
    return make_float3(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
}
template <>
__device__ float2 load<float2>(const float* a, uint32_t i)
{
    return make_float2(a[2 * i], a[2 * i + 1]);
}

template <class P>
__global__ void k_user(lbvh::device::bvh_view bvh, const float* q, const float* d, const float* tmax,
                       const float* r, const float* u, uint32_t n, int flip, uint32_t* cp_idx, float* cp_dist,
                       uint32_t* sil_idx, float* sil_dist, uint8_t* hit, float* t, float* uv, uint32_t* ray_idx,
                       uint8_t* any_hit, int32_t* smp_idx, float* smp_pdf, float* smp_pt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const P p = load<P>(q, i), dir = load<P>(d, i);
    const auto a = lbvh::device::query_nearest(bvh, p);
    cp_idx[i] = a.idx;
    cp_dist[i] = a.dist;
    const auto s = lbvh::device::query_nearest_silhouette(bvh, p, flip != 0);
    sil_idx[i] = s.idx;
    sil_dist[i] = s.dist;
    const auto h = lbvh::device::query_ray(bvh, p, dir, tmax[i]);
    hit[i] = h.hit ? 1 : 0;
    t[i] = h.t;
    uv[2 * i] = h.u;
    uv[2 * i + 1] = h.v;
    ray_idx[i] = h.idx;
    any_hit[i] = lbvh::device::query_any_hit(bvh, p, dir, tmax[i]) ? 1 : 0;
    const auto m = lbvh::device::sample_object_in_sphere(bvh, p, r[i], make_float3(u[3 * i], u[3 * i + 1], u[3 * i + 2]));
    smp_idx[i] = m.idx;
    smp_pdf[i] = m.pdf;
    smp_pt[3 * i] = m.point.x;
    smp_pt[3 * i + 1] = m.point.y;
    smp_pt[3 * i + 2] = m.point.z;
}

} // namespace

// every pointer but the scene is a device pointer; returns an snch status or
// a HIP error code + 100
extern "C" int run_device_api(const snch_scene* scene, uint32_t mesh, const float* q, const float* d,
                              const float* tmax, const float* r, const float* u, uint32_t n, int flip,
                              uint32_t* cp_idx, float* cp_dist, uint32_t* sil_idx, float* sil_dist, uint8_t* hit,
                              float* t, float* uv, uint32_t* ray_idx, uint8_t* any_hit, int32_t* smp_idx,
                              float* smp_pdf, float* smp_pt, void* stream)
{
    snch_device_view dv;
    const int rc = snch_scene_device_view(scene, mesh, &dv);
    if (rc != SNCH_OK)
        return rc;
    const lbvh::device::bvh_view bvh(dv);
    const dim3 grid((n + 127) / 128), block(128);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (dv.dim == 3)
        hipLaunchKernelGGL(k_user<float3>, grid, block, 0, st, bvh, q, d, tmax, r, u, n, flip, cp_idx, cp_dist,
                           sil_idx, sil_dist, hit, t, uv, ray_idx, any_hit, smp_idx, smp_pdf, smp_pt);
    else
        hipLaunchKernelGGL(k_user<float2>, grid, block, 0, st, bvh, q, d, tmax, r, u, n, flip, cp_idx, cp_dist,
                           sil_idx, sil_dist, hit, t, uv, ray_idx, any_hit, smp_idx, smp_pdf, smp_pt);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : 100 + (int)e;
}
