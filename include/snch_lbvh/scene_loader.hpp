// snch_lbvh/scene_loader.hpp -- Wavefront OBJ reader for lbvh::scene<2|3>,
// the host-side counterpart of the reference's scene_loader<2|3>
// (scene_loader.cuh:13-152), header-only, no GPU.
//
//   lbvh::scene_loader<3> ld("mesh.obj");
//   lbvh::scene<3> s(ld.get_vertices().begin(), ld.get_vertices().end(),
//                    ld.get_indices().begin(), ld.get_indices().end());
//
// Format, as the reference reads it: one record per line, the first token the
// record type; "v x y z" a vertex (scene_loader<2> keeps x and y), "f a b c" a
// triangle and "l a b" a segment, with 1-based vertex indices; every other
// record ("vn", "vt", "#", "o", "g", "s", ...) is skipped; a face with more
// than three vertices keeps its first three, as the reference does.  Where
// the reference leaves a value unread (its behaviour is then undefined), this
// reader defines it: "f a/b/c" takes the vertex index before the first '/',
// a negative index counts back from the last vertex read (OBJ relative
// indexing), and a record with too few numbers throws std::runtime_error.  A
// file that cannot be opened throws "Could not open .obj file." like the
// reference.
#pragma once

#include <cstddef>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace lbvh {

struct vec2f {
    float x, y;
};
struct vec3f {
    float x, y, z;
};
struct vec2i {
    int x, y;
};
struct vec3i {
    int x, y, z;
};

namespace detail {

// one vertex reference of an "f" / "l" record: "7", "7/1", "7//3", "7/1/3", "-2"
inline bool obj_index(std::istringstream& ls, std::size_t n_vertices, int& out)
{
    std::string tok;
    if (!(ls >> tok))
        return false;
    const std::size_t slash = tok.find('/');
    const std::string head = tok.substr(0, slash);
    std::size_t used = 0;
    long v = 0;
    try {
        v = std::stol(head, &used);
    } catch (const std::exception&) {
        return false;
    }
    if (used != head.size() || v == 0)
        return false;
    out = v > 0 ? (int)(v - 1) : (int)((long)n_vertices + v);
    return true;
}

inline void obj_fail(const std::string& filename, std::size_t line_no, const char* what)
{
    throw std::runtime_error(filename + ":" + std::to_string(line_no) + ": " + what);
}

} // namespace detail

template <unsigned int dim>
class scene_loader;

template <>
class scene_loader<3> {
public:
    explicit scene_loader(const std::string& filename) { load_from_obj(filename); }
    const std::vector<vec3f>& get_vertices() const { return vertices; }
    const std::vector<vec3i>& get_indices() const { return indices; }
    std::size_t vertices_size() const { return vertices.size(); }
    std::size_t primitives_size() const { return indices.size(); }

private:
    std::vector<vec3f> vertices;
    std::vector<vec3i> indices;

    void load_from_obj(const std::string& filename)
    {
        std::ifstream in(filename);
        if (!in.is_open())
            throw std::runtime_error("Could not open .obj file.");
        std::string line;
        std::size_t line_no = 0;
        while (std::getline(in, line)) {
            ++line_no;
            std::istringstream ls(line);
            std::string prefix;
            ls >> prefix;
            if (prefix == "v") {
                vec3f p;
                if (!(ls >> p.x >> p.y >> p.z))
                    detail::obj_fail(filename, line_no, "a vertex needs three coordinates");
                vertices.push_back(p);
            } else if (prefix == "f") {
                vec3i f;
                if (!detail::obj_index(ls, vertices.size(), f.x) || !detail::obj_index(ls, vertices.size(), f.y) ||
                    !detail::obj_index(ls, vertices.size(), f.z))
                    detail::obj_fail(filename, line_no, "a face needs three vertex indices");
                indices.push_back(f);
            }
        }
    }
};

template <>
class scene_loader<2> {
public:
    explicit scene_loader(const std::string& filename) { load_from_obj(filename); }
    const std::vector<vec2f>& get_vertices() const { return vertices; }
    const std::vector<vec2i>& get_indices() const { return indices; }
    std::size_t vertices_size() const { return vertices.size(); }
    std::size_t primitives_size() const { return indices.size(); }

private:
    std::vector<vec2f> vertices;
    std::vector<vec2i> indices;

    void load_from_obj(const std::string& filename)
    {
        std::ifstream in(filename);
        if (!in.is_open())
            throw std::runtime_error("Could not open .obj file.");
        std::string line;
        std::size_t line_no = 0;
        while (std::getline(in, line)) {
            ++line_no;
            std::istringstream ls(line);
            std::string prefix;
            ls >> prefix;
            if (prefix == "v") {
                vec2f p;
                if (!(ls >> p.x >> p.y))
                    detail::obj_fail(filename, line_no, "a vertex needs two coordinates");
                vertices.push_back(p); // a third coordinate, if present, is ignored
            } else if (prefix == "l") {
                vec2i s;
                if (!detail::obj_index(ls, vertices.size(), s.x) || !detail::obj_index(ls, vertices.size(), s.y))
                    detail::obj_fail(filename, line_no, "a segment needs two vertex indices");
                indices.push_back(s);
            }
        }
    }
};

} // namespace lbvh
