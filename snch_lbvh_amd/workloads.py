"""Seeded synthetic meshes and query sets.

The reference's benchmark meshes (benchmark.cu:319-324) are not part of its
repository, so every workload here is generated from a committed seed with
numpy's PCG64 generator:

* ``icosphere``  -- icosahedron subdivided ``level`` times, projected to the
  unit sphere, vertices displaced radially by N(0, sigma).
* ``cut_torus``  -- torus (R, r) on an nu x nv grid, cut open along the
  meridian u = 0 (seam vertices duplicated, giving 2*nv boundary edges) and
  displaced along the tube normal by Perlin gradient noise (Perlin 2002,
  "Improving Noise").
* query helpers: uniform points in a box, directions uniform on S^2,
  log-uniform radii.
"""
import numpy as np

__all__ = [
    "icosphere", "cut_torus", "uv_sphere", "padded_box", "uniform_points", "unit_directions",
    "log_uniform", "config_small", "config_large", "circle_in_square", "polyline_soup", "torus_interior_points",
    "split_dirichlet_neumann", "surface_points", "mlp_training_points", "batched_meshes",
]


def icosphere(level=5, sigma=0.02, seed=1):
    t = (1.0 + 5 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    for _ in range(level):
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        e = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.ravel()
        mid = V[uniq[:, 0]] + V[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = V.shape[0]
        V = np.concatenate([V, mid])
        nf = F.shape[0]
        ab, bc, ca = inv[:nf] + base, inv[nf:2 * nf] + base, inv[2 * nf:] + base
        F = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    rng = np.random.default_rng(seed)
    if sigma > 0:
        V = V * (1.0 + rng.normal(0.0, sigma, size=(V.shape[0], 1)))
    return V.astype(np.float32), F.astype(np.int32)


class _Perlin:
    """Improved gradient noise (K. Perlin, SIGGRAPH 2002) with a seeded permutation."""

    _G = np.array([[1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
                   [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], np.float64)

    def __init__(self, seed):
        p = np.random.default_rng(seed).permutation(256)
        self.p = np.concatenate([p, p]).astype(np.int64)

    @staticmethod
    def _fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    def __call__(self, P):
        Pi = np.floor(P).astype(np.int64)
        f = P - Pi
        Pi &= 255
        u, v, w = (self._fade(f[:, k]) for k in range(3))
        p = self.p
        X, Y, Z = Pi[:, 0], Pi[:, 1], Pi[:, 2]

        def grad(h, x, y, z):
            g = self._G[h % 12]
            return g[:, 0] * x + g[:, 1] * y + g[:, 2] * z

        A = p[X] + Y
        AA, AB = p[A] + Z, p[A + 1] + Z
        B = p[X + 1] + Y
        BA, BB = p[B] + Z, p[B + 1] + Z
        x, y, z = f[:, 0], f[:, 1], f[:, 2]
        lerp = lambda t, a, b: a + t * (b - a)  # noqa: E731
        return lerp(w,
                    lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                         lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                    lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                         lerp(u, grad(p[AB + 1], x, y - 1, z - 1), grad(p[BB + 1], x - 1, y - 1, z - 1))))


def cut_torus(nu=1024, nv=512, R=1.0, r=0.3, amp=0.05, freq=4.0, seed=3):
    """Torus with 2*nu*nv triangles, open along u = 0 (2*nv boundary edges)."""
    i = np.arange(nu + 1, dtype=np.float64)
    j = np.arange(nv, dtype=np.float64)
    u = (2 * np.pi * i / nu)[:, None]
    v = (2 * np.pi * j / nv)[None, :]
    cu, su, cv, sv = np.cos(u), np.sin(u), np.cos(v), np.sin(v)
    P = np.stack(np.broadcast_arrays((R + r * cv) * cu, (R + r * cv) * su, r * sv + 0 * u), -1).reshape(-1, 3)
    N = np.stack(np.broadcast_arrays(cv * cu, cv * su, sv + 0 * u), -1).reshape(-1, 3)
    if amp > 0:
        noise = _Perlin(seed)(P * freq + 0.5)
        P = P + (amp * noise)[:, None] * N
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = ii * nv + jj
    b = (ii + 1) * nv + jj
    c = (ii + 1) * nv + (jj + 1) % nv
    d = ii * nv + (jj + 1) % nv
    F = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)], 0)
    # interleave the two triangles of each quad so triangle order follows the grid
    F = F.reshape(2, -1, 3).transpose(1, 0, 2).reshape(-1, 3)
    return P.astype(np.float32), F.astype(np.int32)


def uv_sphere(n_tri, sigma=0.02, seed=1):
    """Closed noised sphere with about n_tri triangles (nu = 2 nv, poles fanned)."""
    nv = max(3, int(round((n_tri / 4.0) ** 0.5)))
    nu = 2 * nv
    theta = np.pi * np.arange(1, nv) / nv
    phi = 2 * np.pi * np.arange(nu) / nu
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    ring = np.stack(np.broadcast_arrays(st * np.cos(phi), st * np.sin(phi), ct + 0 * phi), -1).reshape(-1, 3)
    V = np.concatenate([[[0, 0, 1.0]], ring, [[0, 0, -1.0]]])
    rows = nv - 1
    k = np.arange(nu)
    idx = lambda r_, c_: 1 + r_ * nu + (c_ % nu)  # noqa: E731
    top = np.stack([np.zeros(nu, np.int64), idx(0, k), idx(0, k + 1)], 1)
    r_ = np.arange(rows - 1)[:, None]
    a, b, c, d = idx(r_, k), idx(r_ + 1, k), idx(r_ + 1, k + 1), idx(r_, k + 1)
    mid = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], 2).reshape(-1, 3)
    last = V.shape[0] - 1
    bot = np.stack([np.full(nu, last), idx(rows - 1, k + 1), idx(rows - 1, k)], 1)
    tris = np.concatenate([top, mid, bot])
    rng = np.random.default_rng(seed)
    V = V * (1.0 + rng.normal(0.0, sigma, size=(V.shape[0], 1)))
    return V.astype(np.float32), tris.astype(np.int32)


def padded_box(V, pad=0.1):
    lo, hi = V.min(0).astype(np.float64), V.max(0).astype(np.float64)
    ext = hi - lo
    return lo - pad * ext, hi + pad * ext


def uniform_points(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    return (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)


def unit_directions(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d.astype(np.float32)


def log_uniform(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * rng.random(n)).astype(np.float32)


def config_small(n_query=65536):
    """Plain method, small: level-5 noised icosphere, queries uniform in [-1.5, 1.5]^3."""
    V, F = icosphere(5, 0.02, seed=1)
    q = uniform_points(n_query, [-1.5] * 3, [1.5] * 3, seed=2)
    return V, F, q


def config_large(n_query=1 << 22, nu=1024, nv=512):
    """Single large mesh: 2^20-triangle cut torus, queries uniform in the 10%-padded AABB."""
    V, F = cut_torus(nu, nv, seed=3)
    lo, hi = padded_box(V)
    q = uniform_points(n_query, lo, hi, seed=4)
    return V, F, q


def circle_in_square(n_circle=256, n_side=64, radius=0.5, sigma=0.0, seed=1):
    """2D scene like the reference's circle_in_square case: a square [-1, 1]^2
    boundary (counter-clockwise) around a circle polygon (clockwise), both
    closed loops of line segments."""
    rng = np.random.default_rng(seed)
    t = np.linspace(0, 1, n_side, endpoint=False)
    sq = np.concatenate([np.stack([-1 + 2 * t, -np.ones_like(t)], 1), np.stack([np.ones_like(t), -1 + 2 * t], 1),
                         np.stack([1 - 2 * t, np.ones_like(t)], 1), np.stack([-np.ones_like(t), 1 - 2 * t], 1)])
    a = -2 * np.pi * np.arange(n_circle) / n_circle
    rr = radius * (1 + (rng.normal(0, sigma, n_circle) if sigma > 0 else 0))
    ci = np.stack([rr * np.cos(a), rr * np.sin(a)], 1)
    V = np.concatenate([sq, ci]).astype(np.float32)
    ns = sq.shape[0]
    S = np.concatenate([np.stack([np.arange(ns), (np.arange(ns) + 1) % ns], 1),
                        ns + np.stack([np.arange(n_circle), (np.arange(n_circle) + 1) % n_circle], 1)])
    return V, S.astype(np.int32)


def polyline_soup(n_vert=400, n_seg=600, seed=2):
    rng = np.random.default_rng(seed)
    V = rng.uniform(-1, 1, (n_vert, 2)).astype(np.float32)
    S = rng.integers(0, n_vert, (n_seg, 2)).astype(np.int32)
    S = S[S[:, 0] != S[:, 1]]
    return V, S


def torus_interior_points(n, R=1.0, r_max=0.2, seed=8):
    """Points inside the solid torus (tube radius well inside the displaced surface)."""
    rng = np.random.default_rng(seed)
    u = 2 * np.pi * rng.random(n)
    v = 2 * np.pi * rng.random(n)
    rho = r_max * np.sqrt(rng.random(n))
    return np.stack([(R + rho * np.cos(v)) * np.cos(u), (R + rho * np.cos(v)) * np.sin(u), rho * np.sin(v)],
                    1).astype(np.float32)


def split_dirichlet_neumann(V, F):
    """Dirichlet part: triangles with centroid z >= 0; Neumann: the rest."""
    cz = V[F].mean(1)[:, 2]
    return F[cz >= 0], F[cz < 0]


def surface_points(V, F, n, seed):
    """n points uniform by area on the mesh surface."""
    rng = np.random.default_rng(seed)
    a, b, c = (V[F[:, k]].astype(np.float64) for k in range(3))
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    tri = rng.choice(len(F), size=n, p=area / area.sum())
    u, v = rng.random(n), rng.random(n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    return (a[tri] + u[:, None] * (b[tri] - a[tri]) + v[:, None] * (c[tri] - a[tri])).astype(np.float32)


def mlp_training_points(V, F, n, sigma=0.05, seed=9):
    """Half uniform in the 10%-padded AABB, half near the surface (a surface
    point plus an N(0, sigma) offset per axis)."""
    lo, hi = padded_box(V)
    h = n // 2
    far = uniform_points(n - h, lo, hi, seed)
    near = surface_points(V, F, h, seed + 1)
    near = near + np.random.default_rng(seed + 2).normal(0.0, sigma, near.shape).astype(np.float32)
    return np.concatenate([far, near]).astype(np.float32)


def batched_meshes(n_meshes=64, f_lo=1e3, f_hi=5e5, q_lo=2 ** 12, q_hi=2 ** 20, seed=7):
    """Alternating noised spheres and cut tori with triangle counts
    log-uniform in [f_lo, f_hi] and query counts log-uniform in [q_lo, q_hi];
    queries uniform in each mesh's 10%-padded AABB.  Returns
    (meshes, queries concatenated by mesh, query offsets)."""
    rng = np.random.default_rng(seed)
    f = np.exp(rng.uniform(np.log(f_lo), np.log(f_hi), n_meshes))
    nq = np.exp(rng.uniform(np.log(q_lo), np.log(q_hi), n_meshes)).astype(np.int64)
    meshes, qs = [], []
    for k in range(n_meshes):
        if k % 2 == 0:
            V, F = uv_sphere(int(f[k]), sigma=0.02, seed=100 + k)
        else:
            nu = max(8, int(round(f[k] ** 0.5)))
            V, F = cut_torus(nu, max(4, nu // 2), seed=100 + k)
        V = V + np.float32([3.0 * (k % 8), 3.0 * (k // 8), 0.0])
        meshes.append((V, F))
        lo, hi = padded_box(V)
        qs.append(uniform_points(int(nq[k]), lo, hi, 200 + k))
    q_off = np.zeros(n_meshes + 1, np.uint32)
    q_off[1:] = np.cumsum([len(x) for x in qs])
    return meshes, np.concatenate(qs).astype(np.float32), q_off


def slab_box(n=16, h=0.25):
    """The closed box [-1, 1]^2 x [-h, h] with outward-facing triangles, n x n
    quads per face (two triangles each), split into its four side faces
    (x = +-1, y = +-1) and its two caps (z = +-h).  Returns (V, F_sides,
    F_caps); the two parts share the vertex array.  For Laplace with
    Dirichlet data g = x y on the sides and zero Neumann flux on the caps the
    exact solution is u = x y (its z-derivative vanishes)."""
    verts, index = [], {}

    def vid(p):
        key = tuple(np.round(np.asarray(p, np.float64), 12))
        if key not in index:
            index[key] = len(verts)
            verts.append(key)
        return index[key]

    def face(origin, eu, ev):
        tris = []
        g = np.linspace(0.0, 1.0, n + 1)
        for a in range(n):
            for b in range(n):
                p00 = origin + g[a] * eu + g[b] * ev
                p10 = origin + g[a + 1] * eu + g[b] * ev
                p01 = origin + g[a] * eu + g[b + 1] * ev
                p11 = origin + g[a + 1] * eu + g[b + 1] * ev
                i00, i10, i01, i11 = vid(p00), vid(p10), vid(p01), vid(p11)
                tris += [(i00, i10, i11), (i00, i11, i01)]
        return tris

    X, Y, Z = np.eye(3)
    # each face spanned so that eu x ev points outward
    sides = (face(np.array([1.0, -1.0, -h]), 2 * Y, 2 * h * Z) + face(np.array([-1.0, 1.0, -h]), -2 * Y, 2 * h * Z) +
             face(np.array([1.0, 1.0, -h]), -2 * X, 2 * h * Z) + face(np.array([-1.0, -1.0, -h]), 2 * X, 2 * h * Z))
    caps = face(np.array([-1.0, -1.0, h]), 2 * X, 2 * Y) + face(np.array([-1.0, 1.0, -h]), 2 * X, -2 * Y)
    V = np.array(verts, np.float32)
    return V, np.array(sides, np.int32), np.array(caps, np.int32)
