"""Wavefront OBJ reader for Scene / Scene2D: the Python counterpart of the
reference's scene_loader<2|3> (scene_loader.cuh:13-152) and of the C++
header include/snch_lbvh/scene_loader.hpp, with the same rules:

* one record per line, the first token its type; "v x y z" a vertex (2D keeps
  x and y), "f a b c" a triangle, "l a b" a segment, 1-based indices; every
  other record is skipped; a face with more than three vertices keeps its
  first three;
* where the reference leaves a value unread (undefined behaviour there):
  "f a/b/c" takes the index before the first '/', a negative index counts
  back from the last vertex read, and a record with too few numbers raises
  ValueError;
* a file that cannot be opened raises OSError("Could not open .obj file.").

Host-only: no GPU is touched.
"""
import numpy as np


def _index(tok, n_vertices):
    head = tok.split("/", 1)[0]
    v = int(head)
    if v == 0:
        raise ValueError("OBJ indices start at 1")
    return v - 1 if v > 0 else n_vertices + v


def load_obj(path, dim=3):
    """(vertices float32 (nV, dim), primitives int32 (nP, dim)) of an OBJ file:
    triangles for dim 3, line segments for dim 2."""
    try:
        f = open(path, "r")
    except OSError as e:
        raise OSError("Could not open .obj file.") from e
    verts, prims = [], []
    want = "f" if dim == 3 else "l"
    with f:
        for line_no, line in enumerate(f, 1):
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                try:
                    verts.append([float(x) for x in tok[1:1 + dim]])
                except ValueError as e:
                    raise ValueError(f"{path}:{line_no}: a vertex needs {dim} coordinates") from e
                if len(verts[-1]) != dim:
                    raise ValueError(f"{path}:{line_no}: a vertex needs {dim} coordinates")
            elif tok[0] == want:
                try:
                    idx = [_index(t, len(verts)) for t in tok[1:1 + dim]]
                except ValueError as e:
                    raise ValueError(f"{path}:{line_no}: a {'face' if dim == 3 else 'segment'} needs {dim} "
                                     f"vertex indices") from e
                if len(idx) != dim:
                    raise ValueError(f"{path}:{line_no}: a {'face' if dim == 3 else 'segment'} needs {dim} "
                                     f"vertex indices")
                prims.append(idx)
    V = np.asarray(verts, np.float32).reshape(-1, dim)
    P = np.asarray(prims, np.int32).reshape(-1, dim)
    return V, P


class SceneLoader:
    """scene_loader<dim>(filename) (scene_loader.cuh:13-152)."""

    def __init__(self, filename, dim=3):
        self.vertices, self.indices = load_obj(filename, dim)

    def get_vertices(self):
        return self.vertices

    def get_indices(self):
        return self.indices

    def vertices_size(self):
        return len(self.vertices)

    def primitives_size(self):
        return len(self.indices)
