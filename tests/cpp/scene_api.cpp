// Drives the header-only C++ front end (include/snch_lbvh/lbvh.hpp) the way a
// reference user would: OBJ files through lbvh::scene_loader<3|2>,
// lbvh::scene<3|2>, build_bvh, then every query descriptor on device buffers.
// Usage: scene_api <mesh3.obj> <mesh2.obj> <dir>.  Reads the query inputs
// <dir>/in3_*.bin, <dir>/in2_*.bin (float32 or int32, raw) and writes every
// output to <dir>/out3_*.bin, <dir>/out2_*.bin; the calling test compares them
// with the oracle.  Prints "ok" on success.
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "snch_lbvh/lbvh.hpp"

namespace {

std::string g_dir;

template <class T>
std::vector<T> read_bin(const std::string& name)
{
    std::ifstream f(g_dir + "/" + name, std::ios::binary | std::ios::ate);
    if (!f)
        throw std::runtime_error("missing input " + name);
    const std::streamsize bytes = f.tellg();
    f.seekg(0);
    std::vector<T> v(bytes / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), bytes);
    return v;
}

template <class T>
void write_bin(const std::string& name, const std::vector<T>& v)
{
    std::ofstream f(g_dir + "/" + name, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), sizeof(T) * v.size());
}

// a device copy of a host vector, freed at scope exit
template <class T>
struct dev {
    T* p = nullptr;
    size_t n = 0;
    explicit dev(size_t count) : n(count) { lbvh::snch_check(snch_device_malloc((void**)&p, sizeof(T) * (n ? n : 1))); }
    explicit dev(const std::vector<T>& h) : dev(h.size())
    {
        lbvh::snch_check(snch_memcpy_h2d(p, h.data(), sizeof(T) * n, nullptr));
    }
    ~dev() { snch_device_free(p); }
    std::vector<T> host() const
    {
        std::vector<T> h(n);
        lbvh::snch_check(snch_memcpy_d2h(h.data(), p, sizeof(T) * n, nullptr));
        return h;
    }
};

template <unsigned int D, class Scene>
void run_all(const Scene& s, const std::string& tag)
{
    const std::vector<float> q = read_bin<float>("in" + tag + "_points.bin");
    const std::vector<float> d = read_bin<float>("in" + tag + "_dirs.bin");
    const std::vector<float> tmax = read_bin<float>("in" + tag + "_tmax.bin");
    const std::vector<float> r = read_bin<float>("in" + tag + "_radii.bin");
    const std::vector<float> u = read_bin<float>("in" + tag + "_u3.bin");
    const std::vector<float> lo = read_bin<float>("in" + tag + "_lo.bin");
    const std::vector<float> hi = read_bin<float>("in" + tag + "_hi.bin");
    const uint32_t n = (uint32_t)(q.size() / D), M = 16;
    dev<float> dq(q), dd(d), dt(tmax), dr(r), du(u), dlo(lo), dhi(hi);
    dev<uint32_t> idx(n), cnt(n), gidx((size_t)n * M);
    dev<float> dist(n), t(n), uv((size_t)n * (D == 3 ? 2 : 1)), pdf(n), pt((size_t)n * D), gt((size_t)n * M);
    dev<uint8_t> hit(n);
    dev<int32_t> sidx(n);

    s.query(lbvh::nearest(dq.p), n, idx.p, dist.p);
    write_bin("out" + tag + "_cp_idx.bin", idx.host());
    write_bin("out" + tag + "_cp_dist.bin", dist.host());

    s.query(lbvh::nearest_silhouette(dq.p, false), n, idx.p, dist.p);
    write_bin("out" + tag + "_sil_idx.bin", idx.host());
    write_bin("out" + tag + "_sil_dist.bin", dist.host());

    s.query(lbvh::ray_intersect<false>(dq.p, dd.p, dt.p), n, hit.p, t.p, uv.p, idx.p);
    write_bin("out" + tag + "_ray_hit.bin", hit.host());
    write_bin("out" + tag + "_ray_t.bin", t.host());
    write_bin("out" + tag + "_ray_idx.bin", idx.host());
    s.query(lbvh::ray_intersect<true>(dq.p, dd.p, dt.p), n, hit.p);
    write_bin("out" + tag + "_anyhit.bin", hit.host());

    lbvh::sample_object_in_sphere(s, lbvh::sphere_intersect(dq.p, dr.p), du.p, n, sidx.p, pdf.p, pt.p);
    write_bin("out" + tag + "_smp_idx.bin", sidx.host());
    write_bin("out" + tag + "_smp_pdf.bin", pdf.host());
    write_bin("out" + tag + "_smp_pt.bin", pt.host());

    s.query(lbvh::overlaps(dlo.p, dhi.p), n, M, cnt.p, gidx.p);
    write_bin("out" + tag + "_ov_cnt.bin", cnt.host());
    write_bin("out" + tag + "_ov_idx.bin", gidx.host());

    s.query(lbvh::line_intersect(dq.p, dd.p), n, M, cnt.p, gidx.p, gt.p);
    write_bin("out" + tag + "_ln_cnt.bin", cnt.host());
    write_bin("out" + tag + "_ln_idx.bin", gidx.host());
    write_bin("out" + tag + "_ln_t.bin", gt.host());
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 4)
        return 2;
    g_dir = argv[3];
    try {
        lbvh::scene_loader<3> l3(argv[1]);
        lbvh::scene<3> s3(l3.get_vertices().begin(), l3.get_vertices().end(), l3.get_indices().begin(),
                          l3.get_indices().end());
        s3.compute_silhouettes();
        bool threw = false;
        try {
            s3.query(lbvh::nearest(nullptr), 1, nullptr, nullptr);
        } catch (const std::runtime_error&) {
            threw = true; // "BVH is not built yet." (scene.cuh:1250)
        }
        if (!threw)
            return 4;
        s3.build_bvh();
        run_all<3>(s3, "3");

        lbvh::scene_loader<2> l2(argv[2]);
        lbvh::scene<2> s2(l2.get_vertices().begin(), l2.get_vertices().end(), l2.get_indices().begin(),
                          l2.get_indices().end());
        s2.build_bvh();
        run_all<2>(s2, "2");
        std::vector<uint32_t> nodes;
        std::vector<float> boxes, cones;
        s2.export_tree(nodes, boxes, cones);
        if (nodes.size() != 4 * (2 * (size_t)s2.primitives_size() - 1))
            return 5;
        std::printf("ok\n");
        return 0;
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
}
