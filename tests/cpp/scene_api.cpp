// Exercises the header-only C++ front end (include/snch_lbvh/scene.hpp) the
// way a reference user would: construct scene<3>, build, query on device
// buffers.  Prints "ok <checksum>" on success.
#include <cmath>
#include <cstdio>
#include <vector>

#include "snch_lbvh/scene.hpp"

struct f3 { float x, y, z; };
struct i3 { int x, y, z; };

int main()
{
    // unit octahedron
    std::vector<f3> V = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    std::vector<i3> F = {{0, 2, 4}, {2, 1, 4}, {1, 3, 4}, {3, 0, 4}, {2, 0, 5}, {1, 2, 5}, {3, 1, 5}, {0, 3, 5}};
    try {
        lbvh::scene<3> s(V.begin(), V.end(), F.begin(), F.end());
        s.compute_silhouettes();
        bool threw = false;
        try {
            s.query(lbvh::nearest(), nullptr, 1, nullptr, nullptr);
        } catch (const std::runtime_error&) {
            threw = true; // "BVH is not built yet."
        }
        if (!threw)
            return 2;
        s.build_bvh();
        const uint32_t n = 2;
        std::vector<float> q = {2.0f, 0.0f, 0.0f, 0.0f, 0.0f, 3.0f};
        float *dq, *dd;
        uint32_t* di;
        lbvh::snch_check(snch_device_malloc((void**)&dq, sizeof(float) * q.size()));
        lbvh::snch_check(snch_device_malloc((void**)&dd, sizeof(float) * n));
        lbvh::snch_check(snch_device_malloc((void**)&di, sizeof(uint32_t) * n));
        lbvh::snch_check(snch_memcpy_h2d(dq, q.data(), sizeof(float) * q.size(), nullptr));
        s.query(lbvh::nearest(), dq, n, di, dd);
        float d[2];
        lbvh::snch_check(snch_memcpy_d2h(d, dd, sizeof(d), nullptr));
        // distances to the octahedron from (2,0,0) and (0,0,3): vertex distances 1 and 2
        if (std::fabs(d[0] - 1.0f) > 1e-6f || std::fabs(d[1] - 2.0f) > 1e-6f)
            return 3;
        s.query(lbvh::nearest_silhouette{false}, dq, n, di, dd);
        lbvh::snch_check(snch_memcpy_d2h(d, dd, sizeof(d), nullptr));
        // from (2,0,0) the silhouette is the square |y|+|z|=1 at x=0 seen... distance > 1
        if (!(d[0] > 1.0f) || !(d[1] > 2.0f))
            return 4;
        std::vector<uint32_t> nodes;
        std::vector<float> boxes, cones;
        s.export_tree(nodes, boxes, cones);
        if (nodes.size() != 4 * 15 || nodes[3 * 1] != 0xFFFFFFFFu)
            return 5;
        snch_device_free(dq);
        snch_device_free(dd);
        snch_device_free(di);
        std::printf("ok %.6f %.6f\n", d[0], d[1]);
        return 0;
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
}
