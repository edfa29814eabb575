// Reads an OBJ file with lbvh::scene_loader<dim> (include/snch_lbvh/scene_loader.hpp)
// and prints what it read, one value per line: counts, then the vertex
// coordinates (%.9g), then the indices.  Usage: obj_dump <2|3> <file>.
// Exit status 3 and the message on stdout when the loader throws.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "snch_lbvh/scene_loader.hpp"

int main(int argc, char** argv)
{
    if (argc != 3)
        return 2;
    try {
        if (std::atoi(argv[1]) == 3) {
            lbvh::scene_loader<3> ld(argv[2]);
            std::printf("%zu %zu\n", ld.vertices_size(), ld.primitives_size());
            for (const auto& v : ld.get_vertices())
                std::printf("%.9g\n%.9g\n%.9g\n", v.x, v.y, v.z);
            for (const auto& f : ld.get_indices())
                std::printf("%d\n%d\n%d\n", f.x, f.y, f.z);
        } else {
            lbvh::scene_loader<2> ld(argv[2]);
            std::printf("%zu %zu\n", ld.vertices_size(), ld.primitives_size());
            for (const auto& v : ld.get_vertices())
                std::printf("%.9g\n%.9g\n", v.x, v.y);
            for (const auto& s : ld.get_indices())
                std::printf("%d\n%d\n", s.x, s.y);
        }
    } catch (const std::runtime_error& e) {
        std::printf("%s\n", e.what());
        return 3;
    }
    return 0;
}
