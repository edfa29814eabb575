"""Writes the OBJ fixtures of the loader tests and, from the same arrays (not
by parsing), the vertices and primitives each loader must return.

  python tests/golden/make_obj_fixtures.py

* ico2.obj: a level-2 icosphere (320 triangles), plain "v" / "f a b c";
* features.obj: the same mesh with the records a real exporter writes
  (comments, "o", "s", "vn", "vt", "f a/b/c" and "f a//c" forms, relative
  negative indices, a quad whose first three vertices form the triangle);
* circle_square.obj: a 2D scene, "v x y 0" and "l a b" records.
Coordinates are multiples of 2^-12, exact in float32 and in decimal, so every
parser reads the same bits."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from snch_lbvh_amd import workloads as W  # noqa: E402


def q12(x):
    return np.round(np.asarray(x, np.float64) * 4096.0) / 4096.0


def fmt(x):
    return repr(float(x))


def main():
    out = {}
    V, F = W.icosphere(2, 0.0)
    V = q12(V).astype(np.float32)
    with open(os.path.join(HERE, "ico2.obj"), "w") as f:
        f.write("# level-2 icosphere\n")
        for p in V:
            f.write("v " + " ".join(fmt(c) for c in p) + "\n")
        for t in F:
            f.write("f %d %d %d\n" % tuple(int(i) + 1 for i in t))
    out["ico2.obj"] = {"dim": 3, "vertices": V.tolist(), "indices": F.tolist()}

    with open(os.path.join(HERE, "features.obj"), "w") as f:
        f.write("# exporter-style records\no sphere\ns 1\n\n")
        for k, p in enumerate(V):
            f.write("v " + " ".join(fmt(c) for c in p) + "\n")
            f.write("vn 0 0 1\nvt 0.5 0.5\n")
        nV = len(V)
        exp = []
        for k, t in enumerate(F):
            a, b, c = (int(i) + 1 for i in t)
            if k % 4 == 0:
                f.write("f %d/1/1 %d/1/1 %d/1/1\n" % (a, b, c))
            elif k % 4 == 1:
                f.write("f %d//1 %d//1 %d//1\n" % (a, b, c))
            elif k % 4 == 2:  # relative indices: -1 is the last vertex read
                f.write("f %d %d %d\n" % (a - 1 - nV, b - 1 - nV, c - 1 - nV))
            else:  # a quad: the first three vertices make the triangle
                f.write("f %d %d %d %d\n" % (a, b, c, a))
            exp.append([a - 1, b - 1, c - 1])
        f.write("l 1 2\n")  # a segment record in a 3D file is skipped
    out["features.obj"] = {"dim": 3, "vertices": V.tolist(), "indices": exp}

    V2, S2 = W.circle_in_square(48, 12)
    V2 = q12(V2).astype(np.float32)
    with open(os.path.join(HERE, "circle_square.obj"), "w") as f:
        f.write("# circle in a square, line segments\n")
        for p in V2:
            f.write("v %s %s 0\n" % (fmt(p[0]), fmt(p[1])))
        for s in S2:
            f.write("l %d %d\n" % (int(s[0]) + 1, int(s[1]) + 1))
        f.write("f 1 2 3\n")  # a face record in a 2D file is skipped
    out["circle_square.obj"] = {"dim": 2, "vertices": V2.tolist(), "indices": S2.tolist()}

    with open(os.path.join(HERE, "obj_expected.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
