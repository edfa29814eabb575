/*
 * ORACLE -- test infrastructure, never product code.
 *
 * A CPU restatement of the reference SNCH-LBVH algorithm
 * (tyanyuy3125/snch-lbvh, include/snch_lbvh/ headers), written in plain C so that
 * the HIP library's results can be checked bit for bit:
 *   - edge table, edge slots and edge ownership (scene.cuh:1135-1225),
 *   - leaf AABB and leaf normal cone (scene.cuh:870-961, aabb.cuh:21-50,93-101),
 *   - scene AABB, Morton codes, stable sort, Karras topology, bottom-up AABB
 *     and cone refit (bvh.cuh:109-229, 281-304, 380-604; morton_code.cuh),
 *   - the four queries in the reference's own visit order
 *     (query.cuh:79-169, 238-318, 325-423; sample.cuh:7-92),
 *   - an independent FP64 brute force for closest point, closest silhouette
 *     and ray hit.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker.
 *
 * Deviations from the reference, all deliberate and mirrored by the library:
 *   Q1  cone merge with o_theta >= pi leaves half_angle uninitialised in the
 *       reference (cone.cuh:454-458); here it is pi (always overlaps).
 *   Q3  the silhouette query also reports the edge id it accepted.
 *   Q6  N == 1: the root is the single leaf; queries evaluate it directly.
 *   Q19 the sampler's pdf is 0 when no triangle is returned.
 *   Q28 rotate() with (anti)parallel axes: the rotation axis falls back to a
 *       vector perpendicular to axis_a instead of NaN (cone.cuh:292).
 *   Transcendentals use the deterministic set in detmath_oracle.h.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "detmath_oracle.h"

#define INVALID 0xFFFFFFFFu
#define F_EPS FLT_EPSILON
#define F_PI 3.14159274101257324219f   /* (float)M_PI */
#define F_PI_2 1.57079637050628662109f /* (float)M_PI_2 */

typedef struct { float x, y, z; } f3;
typedef struct { f3 lo, hi; } box3;
typedef struct { f3 axis; float half_angle; float radius; } cone3;

/* ------------------------------------------------------------------ */
/* vector helpers (utility.cuh:236-462); left-to-right evaluation      */
/* ------------------------------------------------------------------ */
static inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float len3(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline float sqlen3(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline f3 cross3(f3 u, f3 v)
{
    return mk(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
static inline f3 normalize3(f3 v)
{
    float n = len3(v);
    return mk(v.x / n, v.y / n, v.z / n);
}
static inline f3 cabs3(f3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
/* std::min / std::max semantics: min(a,b) = (b < a) ? b : a */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }
static inline float clamp11(float d) { return smax(-1.0f, smin(1.0f, d)); }

/* Build option bit 0: conservative cone radius.  The reference bounds the
 * distance from a node's centroid to its silhouette-edge midpoints
 * (scene.cuh:909-915) and merges radii as sqrt(max(ra^2+|da|^2, rb^2+|db|^2))
 * (cone.cuh:436-438); neither encloses whole edges, so the cone-overlap test
 * can prune the true nearest silhouette.  The conservative option bounds the
 * edge end points at the leaf and merges with the triangle inequality
 * max(ra+|da|, rb+|db|).  It only ever prunes less. */
static int g_conservative_radius = 0;

static inline f3 vget(const float *V, int i) { return mk(V[3 * i], V[3 * i + 1], V[3 * i + 2]); }

/* ------------------------------------------------------------------ */
/* AABB (aabb.cuh)                                                     */
/* ------------------------------------------------------------------ */
static inline box3 box_of_point(f3 p) /* aabb.cuh:30-38 */
{
    box3 b;
    b.hi = mk(p.x + F_EPS, p.y + F_EPS, p.z + F_EPS);
    b.lo = mk(p.x - F_EPS, p.y - F_EPS, p.z - F_EPS);
    return b;
}
static inline void box_expand(box3 *b, f3 p) /* aabb.cuh:93-101 */
{
    f3 pl = mk(p.x - F_EPS, p.y - F_EPS, p.z - F_EPS);
    f3 pu = mk(p.x + F_EPS, p.y + F_EPS, p.z + F_EPS);
    b->lo = mk(fminf(b->lo.x, pl.x), fminf(b->lo.y, pl.y), fminf(b->lo.z, pl.z));
    b->hi = mk(fmaxf(b->hi.x, pu.x), fmaxf(b->hi.y, pu.y), fmaxf(b->hi.z, pu.z));
}
static inline box3 box_merge(box3 a, box3 b) /* aabb.cuh:113-124 */
{
    box3 m;
    m.hi = mk(fmaxf(a.hi.x, b.hi.x), fmaxf(a.hi.y, b.hi.y), fmaxf(a.hi.z, b.hi.z));
    m.lo = mk(fminf(a.lo.x, b.lo.x), fminf(a.lo.y, b.lo.y), fminf(a.lo.z, b.lo.z));
    return m;
}
static inline f3 box_centroid(box3 b) /* aabb.cuh:271-279 */
{
    return mk((b.hi.x + b.lo.x) * 0.5f, (b.hi.y + b.lo.y) * 0.5f, (b.hi.z + b.lo.z) * 0.5f);
}
static inline float box_mindist(box3 b, f3 p) /* aabb.cuh:144-150 */
{
    float dx = fminf(b.hi.x, fmaxf(b.lo.x, p.x)) - p.x;
    float dy = fminf(b.hi.y, fmaxf(b.lo.y, p.y)) - p.y;
    float dz = fminf(b.hi.z, fmaxf(b.lo.z, p.z)) - p.z;
    return dx * dx + dy * dy + dz * dz;
}
static inline float box_minmaxdist(box3 b, f3 p) /* aabb.cuh:182-209 */
{
    float lx = b.lo.x - p.x, ly = b.lo.y - p.y, lz = b.lo.z - p.z;
    float ux = b.hi.x - p.x, uy = b.hi.y - p.y, uz = b.hi.z - p.z;
    float rmx = lx * lx, rmy = ly * ly, rmz = lz * lz;
    float rMx = ux * ux, rMy = uy * uy, rMz = uz * uz;
    float t;
    if ((b.hi.x + b.lo.x) * 0.5f < p.x) { t = rmx; rmx = rMx; rMx = t; }
    if ((b.hi.y + b.lo.y) * 0.5f < p.y) { t = rmy; rmy = rMy; rMy = t; }
    if ((b.hi.z + b.lo.z) * 0.5f < p.z) { t = rmz; rmz = rMz; rMz = t; }
    {
        float dx = rmx + rMy + rMz;
        float dy = rMx + rmy + rMz;
        float dz = rMx + rMy + rmz;
        return fminf(dx, fminf(dy, dz));
    }
}
/* ray slab test, aabb.cuh:397-431 */
static inline int box_ray(box3 b, f3 o, f3 inv, float max_dist, float *dist)
{
    float t1 = (b.lo.x - o.x) * inv.x;
    float t2 = (b.hi.x - o.x) * inv.x;
    float tmin = fminf(t1, t2);
    float tmax = fmaxf(t1, t2);
    t1 = (b.lo.y - o.y) * inv.y;
    t2 = (b.hi.y - o.y) * inv.y;
    tmin = fmaxf(tmin, fminf(t1, t2));
    tmax = fminf(tmax, fmaxf(t1, t2));
    t1 = (b.lo.z - o.z) * inv.z;
    t2 = (b.hi.z - o.z) * inv.z;
    tmin = fmaxf(tmin, fminf(t1, t2));
    tmax = fminf(tmax, fmaxf(t1, t2));
    if (tmax >= tmin && tmax >= 0.0f && tmin <= max_dist) {
        *dist = tmin >= 0.0f ? tmin : 0.0f;
        return 1;
    }
    return 0;
}
/* sphere-box test, aabb.cuh:433-449 (std::max/std::min semantics) */
static inline int box_sphere(box3 b, f3 c, float r)
{
    float cx = smax(b.lo.x, smin(c.x, b.hi.x));
    float cy = smax(b.lo.y, smin(c.y, b.hi.y));
    float dx = cx - c.x;
    float dy = cy - c.y;
    float d2 = dx * dx + dy * dy;
    float cz = smax(b.lo.z, smin(c.z, b.hi.z));
    float dz = cz - c.z;
    d2 += dz * dz;
    return d2 <= (r * r);
}

/* ------------------------------------------------------------------ */
/* Normal cones (cone.cuh)                                             */
/* ------------------------------------------------------------------ */
static inline int cone_valid(cone3 c) { return c.half_angle >= 0.0f; } /* cone.cuh:18-22 */

static inline void ortho_basis(f3 n, f3 *b1, f3 *b2) /* cone.cuh:34-42 */
{
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    *b1 = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    *b2 = mk(b, sign + n.y * n.y * a, -n.y);
}
static inline float project_to_plane(f3 n, f3 e) /* cone.cuh:58-66 */
{
    f3 b1, b2;
    float r1, r2;
    ortho_basis(n, &b1, &b2);
    r1 = dot3(e, cabs3(b1));
    r2 = dot3(e, cabs3(b2));
    return sqrtf(r1 * r1 + r2 * r2);
}
static inline int inrange(float v, float lo, float hi) { return v >= lo && v <= hi; }

/* cone.cuh:168-212 */
static int cone_overlap(cone3 bc, f3 o, box3 b, float dist_to_box)
{
    f3 c, v, e;
    float l, d_axis, d, s, vha, sum, mn, mx;
    if (bc.half_angle >= F_PI_2 || dist_to_box < F_EPS)
        return 1;
    c = box_centroid(b);
    v = mk(c.x - o.x, c.y - o.y, c.z - o.z);
    l = len3(v);
    v.x /= l;
    v.y /= l;
    v.z /= l;
    d_axis = odm_acosf(clamp11(dot3(bc.axis, v)));
    if (inrange(F_PI_2, d_axis - bc.half_angle, d_axis + bc.half_angle))
        return 1;
    if (l > bc.radius) {
        vha = odm_asinf(bc.radius / l);
        sum = bc.half_angle + vha;
        mn = d_axis - sum;
        mx = d_axis + sum;
        return sum >= F_PI_2 ? 1 : inrange(F_PI_2, mn, mx);
    }
    e = mk(b.hi.x - c.x, b.hi.y - c.y, b.hi.z - c.z);
    d = dot3(e, cabs3(v));
    s = l - d;
    if (s <= 0.0f)
        return 1;
    d = project_to_plane(v, e);
    vha = odm_atan2f(d, s);
    sum = bc.half_angle + vha;
    mn = d_axis - sum;
    mx = d_axis + sum;
    return sum >= F_PI_2 ? 1 : inrange(F_PI_2, mn, mx);
}

/* cone.cuh:288-302, with the Q28 guard */
static f3 cone_rotate(f3 u, f3 v, float theta)
{
    float ct, st, n;
    f3 w, om;
    float R00, R01, R02, R10, R11, R12, R20, R21, R22;
    odm_sincosf(theta, &st, &ct);
    w = cross3(u, v);
    n = len3(w);
    if (!(n > 0.0f)) {
        f3 b1, b2;
        ortho_basis(u, &b1, &b2);
        w = b1;
        n = len3(w);
    }
    w = mk(w.x / n, w.y / n, w.z / n);
    om = mk((1.0f - ct) * w.x, (1.0f - ct) * w.y, (1.0f - ct) * w.z);
    R00 = ct + om.x * w.x;
    R01 = om.y * w.x - st * w.z;
    R02 = om.z * w.x + st * w.y;
    R10 = om.x * w.y + st * w.z;
    R11 = ct + om.y * w.y;
    R12 = om.z * w.y - st * w.x;
    R20 = om.x * w.z - st * w.y;
    R21 = om.y * w.z + st * w.x;
    R22 = ct + om.z * w.z;
    return mk(R00 * u.x + R01 * u.y + R02 * u.z, R10 * u.x + R11 * u.y + R12 * u.z,
              R20 * u.x + R21 * u.y + R22 * u.z);
}

/* cone.cuh:427-480 (Q1, Q28) */
static cone3 cone_merge(cone3 a, cone3 b, f3 oa, f3 ob, f3 on)
{
    cone3 r;
    memset(&r, 0, sizeof(r));
    if (cone_valid(a) && cone_valid(b)) {
        f3 axa = a.axis, axb = b.axis, t3;
        float ha = a.half_angle, hb = b.half_angle, t, theta, o_theta;
        f3 da = mk(on.x - oa.x, on.y - oa.y, on.z - oa.z);
        f3 db = mk(on.x - ob.x, on.y - ob.y, on.z - ob.z);
        if (g_conservative_radius)
            r.radius = smax(a.radius + len3(da), b.radius + len3(db));
        else
            r.radius = sqrtf(smax(a.radius * a.radius + sqlen3(da), b.radius * b.radius + sqlen3(db)));
        if (hb > ha) {
            t3 = axa; axa = axb; axb = t3;
            t = ha; ha = hb; hb = t;
        }
        theta = odm_acosf(clamp11(dot3(axa, axb)));
        if (smin(theta + hb, F_PI) <= ha) {
            r.axis = axa;
            r.half_angle = ha;
            return r;
        }
        o_theta = (ha + theta + hb) / 2.0f;
        if (o_theta >= F_PI) {
            r.axis = axa;
            r.half_angle = F_PI; /* Q1 */
            return r;
        }
        r.axis = cone_rotate(axa, axb, o_theta - ha);
        r.half_angle = o_theta;
    } else if (cone_valid(a)) {
        r = a;
    } else if (cone_valid(b)) {
        r = b;
    } else {
        r.half_angle = -F_PI;
    }
    return r;
}

/* ------------------------------------------------------------------ */
/* Scene geometry (scene.cuh)                                          */
/* ------------------------------------------------------------------ */
/* Ericson closest point on triangle, scene.cuh:34-110 */
static float closest_point_triangle(f3 pa, f3 pb, f3 pc, f3 x)
{
    f3 ab = sub(pb, pa), ac = sub(pc, pa), ax = sub(x, pa), bx, cx, pt;
    float d1 = dot3(ab, ax), d2 = dot3(ac, ax), d3, d4, d5, d6, vc, vb, va, v, w, denom;
    if (d1 <= 0.0f && d2 <= 0.0f)
        return len3(mk(x.x - pa.x, x.y - pa.y, x.z - pa.z));
    bx = sub(x, pb);
    d3 = dot3(ab, bx);
    d4 = dot3(ac, bx);
    if (d3 >= 0.0f && d4 <= d3)
        return len3(mk(x.x - pb.x, x.y - pb.y, x.z - pb.z));
    cx = sub(x, pc);
    d5 = dot3(ab, cx);
    d6 = dot3(ac, cx);
    if (d6 >= 0.0f && d5 <= d6)
        return len3(mk(x.x - pc.x, x.y - pc.y, x.z - pc.z));
    vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        v = d1 / (d1 - d3);
        pt = mk(pa.x + ab.x * v, pa.y + ab.y * v, pa.z + ab.z * v);
        return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
    }
    vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        w = d2 / (d2 - d6);
        pt = mk(pa.x + ac.x * w, pa.y + ac.y * w, pa.z + ac.z * w);
        return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
    }
    va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        pt = mk(pb.x + (pc.x - pb.x) * w, pb.y + (pc.y - pb.y) * w, pb.z + (pc.z - pb.z) * w);
        return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
    }
    denom = 1.0f / (va + vb + vc);
    v = vb * denom;
    w = vc * denom;
    pt = mk(pa.x + ab.x * v + ac.x * w, pa.y + ab.y * v + ac.y * w, pa.z + ab.z * v + ac.z * w);
    return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
}

/* scene.cuh:230-255 */
static float closest_point_segment(f3 pa, f3 pb, f3 x, f3 *pt)
{
    f3 u = sub(pb, pa), v = sub(x, pa);
    float c1 = dot3(u, v), c2, t;
    if (c1 <= 0.0f) {
        *pt = pa;
        return len3(mk(x.x - pt->x, x.y - pt->y, x.z - pt->z));
    }
    c2 = dot3(u, u);
    if (c2 <= c1) {
        *pt = pb;
        return len3(mk(x.x - pt->x, x.y - pt->y, x.z - pt->z));
    }
    t = c1 / c2;
    *pt = mk(pa.x + u.x * t, pa.y + u.y * t, pa.z + u.z * t);
    return len3(mk(x.x - pt->x, x.y - pt->y, x.z - pt->z));
}

/* scene.cuh:143-174 (Q2: the 3D view direction is not normalised) */
static int is_silhouette_edge(f3 pa, f3 pb, f3 n0, f3 n1, f3 view, float d, int flip)
{
    const float precision = 1e-3f;
    float sign = flip ? 1.0f : -1.0f;
    float dot0, dot1;
    if (d <= precision) {
        f3 ed = normalize3(mk(pb.x - pa.x, pb.y - pa.y, pb.z - pa.z));
        float ang = odm_atan2f(dot3(ed, cross3(n0, n1)), dot3(n0, n1));
        return sign * ang > precision;
    }
    dot0 = dot3(view, n0);
    dot1 = dot3(view, n1);
    if (fabsf(dot0) <= precision)
        return sign * dot1 > precision;
    if (fabsf(dot1) <= precision)
        return sign * dot0 > precision;
    return dot0 * dot1 < 0.0f;
}

/* silhouette edge record helpers, scene.cuh:743-787.  slots = (opp0, lo, hi, opp1) */
static inline int edge_has_face(const int32_t *s, int f) { return f == 0 ? s[3] != -1 : s[0] != -1; }
static f3 edge_face_normal(const float *V, const int32_t *s, int f, int do_norm)
{
    int i, j, k;
    f3 pa, pb, pc, n;
    if (f == 0) { i = 3; j = 1; k = 2; }
    else { i = 0; j = 2; k = 1; }
    pa = vget(V, s[j]);
    pb = vget(V, s[k]);
    pc = vget(V, s[i]);
    n = cross3(mk(pb.x - pa.x, pb.y - pa.y, pb.z - pa.z), mk(pc.x - pa.x, pc.y - pa.y, pc.z - pa.z));
    return do_norm ? normalize3(n) : n;
}
static f3 edge_normal(const float *V, const int32_t *s)
{
    f3 n = mk(0.0f, 0.0f, 0.0f);
    if (edge_has_face(s, 0)) {
        f3 n0 = edge_face_normal(V, s, 0, 0);
        n = mk(n.x + n0.x, n.y + n0.y, n.z + n0.z);
    }
    if (edge_has_face(s, 1)) {
        f3 n1 = edge_face_normal(V, s, 1, 0);
        n = mk(n.x + n1.x, n.y + n1.y, n.z + n1.z);
    }
    return normalize3(n);
}

/* scene.cuh:788-824 with min_radius_squared == 0 */
static int edge_closest_silhouette(const float *V, const int32_t *s, f3 o, float max_r2, float *dist, int flip)
{
    f3 pa, pb, cp;
    float d;
    int sil;
    if (0.0f >= max_r2)
        return 0;
    pa = vget(V, s[1]);
    pb = vget(V, s[2]);
    d = closest_point_segment(pa, pb, o, &cp);
    if (d * d > max_r2)
        return 0;
    sil = !edge_has_face(s, 0) || !edge_has_face(s, 1);
    if (!sil) {
        f3 n0 = edge_face_normal(V, s, 0, 1);
        f3 n1 = edge_face_normal(V, s, 1, 1);
        f3 view = mk(o.x - cp.x, o.y - cp.y, o.z - cp.z);
        sil = is_silhouette_edge(pa, pb, n0, n1, view, d, flip);
    }
    if (sil && d * d <= max_r2) {
        *dist = d;
        return 1;
    }
    return 0;
}

/* Moller-Trumbore, scene.cuh:1005-1052 (device reciprocal is the IEEE one) */
static int ray_triangle(f3 o, f3 dir, f3 v0, f3 v1, f3 v2, float *t_out, float *u_out, float *v_out)
{
    f3 e1 = mk(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z);
    f3 e2 = mk(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z);
    f3 h = mk(dir.y * e2.z - dir.z * e2.y, dir.z * e2.x - dir.x * e2.z, dir.x * e2.y - dir.y * e2.x);
    float det = e1.x * h.x + e1.y * h.y + e1.z * h.z, inv, u, v, t;
    f3 s, q;
    if (fabsf(det) < F_EPS)
        return 0;
    inv = 1.0f / det;
    s = mk(o.x - v0.x, o.y - v0.y, o.z - v0.z);
    u = (s.x * h.x + s.y * h.y + s.z * h.z) * inv;
    if (u < 0.0f || u > 1.0f)
        return 0;
    q = mk(s.y * e1.z - s.z * e1.y, s.z * e1.x - s.x * e1.z, s.x * e1.y - s.y * e1.x);
    v = (dir.x * q.x + dir.y * q.y + dir.z * q.z) * inv;
    if (v < 0.0f || u + v > 1.0f)
        return 0;
    t = (e2.x * q.x + e2.y * q.y + e2.z * q.z) * inv;
    if (t >= 0.0f) {
        *t_out = t;
        *u_out = u;
        *v_out = v;
        return 1;
    }
    return 0;
}

/* triangle-sphere test, scene.cuh:1054-1117 */
static int tri_sphere(f3 p1, f3 p2, f3 p3, f3 c, float radius)
{
    f3 e1 = mk(p2.x - p1.x, p2.y - p1.y, p2.z - p1.z);
    f3 e2 = mk(p3.x - p1.x, p3.y - p1.y, p3.z - p1.z);
    f3 n = mk(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    float nl = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    float d, dp, d00, d01, d02, d11, d12, inv, u, v;
    f3 proj, v0, v1, v2;
    n = mk(n.x / nl, n.y / nl, n.z / nl);
    d = n.x * p1.x + n.y * p1.y + n.z * p1.z;
    dp = n.x * c.x + n.y * c.y + n.z * c.z - d;
    proj = mk(c.x - dp * n.x, c.y - dp * n.y, c.z - dp * n.z);
    v0 = mk(p3.x - p1.x, p3.y - p1.y, p3.z - p1.z);
    v1 = mk(p2.x - p1.x, p2.y - p1.y, p2.z - p1.z);
    v2 = mk(proj.x - p1.x, proj.y - p1.y, proj.z - p1.z);
    d00 = v0.x * v0.x + v0.y * v0.y + v0.z * v0.z;
    d01 = v0.x * v1.x + v0.y * v1.y + v0.z * v1.z;
    d02 = v0.x * v2.x + v0.y * v2.y + v0.z * v2.z;
    d11 = v1.x * v1.x + v1.y * v1.y + v1.z * v1.z;
    d12 = v1.x * v2.x + v1.y * v2.y + v1.z * v2.z;
    inv = 1.0f / (d00 * d11 - d01 * d01);
    u = (d11 * d02 - d01 * d12) * inv;
    v = (d00 * d12 - d01 * d02) * inv;
    if (u >= 0 && v >= 0 && u + v <= 1) {
        return fabsf(dp) <= radius;
    } else {
        f3 cp = proj;
        float dx, dy, dz, d2;
        if (u < 0)
            cp = p1;
        else if (v < 0)
            cp = p3;
        else if (u + v > 1)
            cp = p2;
        dx = cp.x - c.x;
        dy = cp.y - c.y;
        dz = cp.z - c.z;
        d2 = dx * dx + dy * dy + dz * dz;
        return d2 <= (radius * radius);
    }
}

/* scene.cuh:855-868 */
static float tri_area(f3 pa, f3 pb, f3 pc)
{
    f3 ac = mk(pc.x - pa.x, pc.y - pa.y, pc.z - pa.z);
    f3 ab = mk(pb.x - pa.x, pb.y - pa.y, pb.z - pa.z);
    return len3(cross3(ac, ab)) / 2.0f;
}

/* scene.cuh:1119-1126 */
static float green_weight(f3 x, f3 y)
{
    float r = smax(len3(mk(x.x - y.x, x.y - y.y, x.z - y.z)), 1e-4f);
    return 1.0f / (F_PI * 4.0f * r);
}

/* ================================================================== */
/* Exported API                                                        */
/* ================================================================== */

/*
 * Edge table: reference edge ids, edge slots and ownership.
 * scene.cuh:1135-1166 numbers edges by first appearance over (triangle i,
 * local edge j = (v_j, v_{j+1})); :1167-1204 fills (opp0, lo, hi, opp1);
 * :1205-1225 gives each triangle the edges it is first to list.
 * The restatement sorts (lo, hi, position) triples instead of using std::map;
 * tests/test_oracle_scene.py pins it against a literal dict-based version.
 */
typedef struct { int32_t a, b, pos; } edge_key;
static int cmp_edge_key(const void *x, const void *y)
{
    const edge_key *p = (const edge_key *)x, *q = (const edge_key *)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    return p->pos < q->pos ? -1 : (p->pos > q->pos);
}
typedef struct { int32_t first, group; } group_first;
static int cmp_group_first(const void *x, const void *y)
{
    const group_first *p = (const group_first *)x, *q = (const group_first *)y;
    return p->first < q->first ? -1 : (p->first > q->first);
}

int orc_edge_table(const int32_t *F, int32_t nF, int32_t *tri_edge, int32_t *slots, int32_t *owned)
{
    int32_t n = 3 * nF, i, j, ng = 0, E;
    edge_key *keys = (edge_key *)malloc(sizeof(edge_key) * (size_t)(n > 0 ? n : 1));
    int32_t *grp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    group_first *gf = (group_first *)malloc(sizeof(group_first) * (size_t)(n > 0 ? n : 1));
    int32_t *gid = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    unsigned char *seen;
    for (i = 0; i < nF; ++i)
        for (j = 0; j < 3; ++j) {
            int32_t I = F[3 * i + j], J = F[3 * i + (j + 1) % 3], t;
            if (I > J) { t = I; I = J; J = t; }
            keys[3 * i + j].a = I;
            keys[3 * i + j].b = J;
            keys[3 * i + j].pos = 3 * i + j;
        }
    qsort(keys, (size_t)n, sizeof(edge_key), cmp_edge_key);
    for (i = 0; i < n; ++i) {
        if (i == 0 || keys[i].a != keys[i - 1].a || keys[i].b != keys[i - 1].b) {
            gf[ng].first = keys[i].pos;
            gf[ng].group = ng;
            ++ng;
        }
        grp[keys[i].pos] = ng - 1;
    }
    qsort(gf, (size_t)ng, sizeof(group_first), cmp_group_first);
    for (i = 0; i < ng; ++i)
        gid[gf[i].group] = i;
    for (i = 0; i < n; ++i)
        tri_edge[i] = gid[grp[i]];
    E = ng;
    for (i = 0; i < 4 * E; ++i)
        slots[i] = -1;
    for (i = 0; i < nF; ++i) {
        for (j = 0; j < 3; ++j) {
            int I = j - 1 < 0 ? 2 : j - 1, J = j, K = j + 1 > 2 ? 0 : j + 1, orientation = 1, t;
            int32_t e = tri_edge[3 * i + j];
            if (F[3 * i + J] > F[3 * i + K]) { t = J; J = K; K = t; orientation = -1; }
            slots[4 * e + (orientation == 1 ? 0 : 3)] = F[3 * i + I];
            slots[4 * e + 1] = F[3 * i + J];
            slots[4 * e + 2] = F[3 * i + K];
        }
    }
    seen = (unsigned char *)calloc((size_t)(E > 0 ? E : 1), 1);
    for (i = 0; i < nF; ++i) {
        int ptr = 0;
        owned[3 * i] = owned[3 * i + 1] = owned[3 * i + 2] = -1;
        for (j = 0; j < 3; ++j) {
            int32_t e = tri_edge[3 * i + j];
            if (!seen[e]) {
                seen[e] = 1;
                owned[3 * i + ptr++] = e;
            }
        }
    }
    free(keys); free(grp); free(gf); free(gid); free(seen);
    return E;
}

/* ------------------------------------------------------------------ */
/* Build: bvh.cuh:380-604 for scene<3> objects                          */
/* ------------------------------------------------------------------ */
static box3 leaf_box(const float *V, const int32_t *F, int32_t i) /* scene.cuh:870-885 */
{
    box3 b = box_of_point(vget(V, F[3 * i]));
    box_expand(&b, vget(V, F[3 * i + 1]));
    box_expand(&b, vget(V, F[3 * i + 2]));
    return b;
}


static cone3 leaf_cone(const float *V, const int32_t *F, const int32_t *slots, const int32_t *owned, int32_t i)
{ /* scene.cuh:887-961 */
    box3 b = leaf_box(V, F, i);
    f3 c = box_centroid(b);
    cone3 r;
    int any = 0, all_two = 1, k;
    r.axis = mk(0.0f, 0.0f, 0.0f);
    r.half_angle = F_PI;
    r.radius = 0.0f;
    for (k = 0; k < 3; ++k) {
        int32_t e = owned[3 * i + k];
        if (e != -1) {
            const int32_t *s = slots + 4 * e;
            f3 n = edge_normal(V, s);
            f3 pa = vget(V, s[1]), pb = vget(V, s[2]);
            f3 m = mk((pa.x + pb.x) / 2.0f, (pa.y + pb.y) / 2.0f, (pa.z + pb.z) / 2.0f);
            r.axis.x += n.x;
            r.axis.y += n.y;
            r.axis.z += n.z;
            if (g_conservative_radius) {
                r.radius = smax(r.radius, len3(mk(pa.x - c.x, pa.y - c.y, pa.z - c.z)));
                r.radius = smax(r.radius, len3(mk(pb.x - c.x, pb.y - c.y, pb.z - c.z)));
            } else {
                r.radius = smax(r.radius, len3(mk(m.x - c.x, m.y - c.y, m.z - c.z)));
            }
            all_two = all_two && edge_has_face(s, 0) && edge_has_face(s, 1);
            any = 1;
        }
    }
    if (!any) {
        r.half_angle = -F_PI;
    } else if (!all_two) {
        r.half_angle = F_PI;
    } else {
        float an = len3(r.axis);
        if (an > F_EPS) {
            r.axis.x /= an;
            r.axis.y /= an;
            r.axis.z /= an;
            r.half_angle = 0.0f;
            for (k = 0; k < 3; ++k) {
                int32_t e = owned[3 * i + k];
                if (e != -1) {
                    const int32_t *s = slots + 4 * e;
                    f3 nf[2];
                    int jj;
                    nf[0] = edge_has_face(s, 0) ? edge_face_normal(V, s, 0, 1) : mk(0.0f, 0.0f, 0.0f);
                    nf[1] = edge_has_face(s, 1) ? edge_face_normal(V, s, 1, 1) : mk(0.0f, 0.0f, 0.0f);
                    for (jj = 0; jj < 2; ++jj) {
                        float ang = odm_acosf(clamp11(dot3(r.axis, nf[jj])));
                        r.half_angle = smax(r.half_angle, ang);
                    }
                }
            }
        }
    }
    return r;
}

static inline uint32_t expand_bits(uint32_t v) /* morton_code.cuh:19-38 */
{
    v = (v | (v << 16)) & 0x070000FFu;
    v = (v | (v << 8)) & 0x0700F00Fu;
    v = (v | (v << 4)) & 0x430C30C3u;
    v = (v | (v << 2)) & 0x49249249u;
    return v;
}
static inline uint32_t morton3(f3 p) /* morton_code.cuh:61-70 */
{
    const float res = 1024.0f;
    float x = fminf(fmaxf(p.x * res, 0.0f), res - 1.0f);
    float y = fminf(fmaxf(p.y * res, 0.0f), res - 1.0f);
    float z = fminf(fmaxf(p.z * res, 0.0f), res - 1.0f);
    return expand_bits((uint32_t)x) * 4u + expand_bits((uint32_t)y) * 2u + expand_bits((uint32_t)z);
}

static inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }
static inline int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

typedef struct { uint32_t code, idx; } code_idx;
static int cmp_code_idx(const void *x, const void *y)
{
    const code_idx *p = (const code_idx *)x, *q = (const code_idx *)y;
    if (p->code != q->code) return p->code < q->code ? -1 : 1;
    return p->idx < q->idx ? -1 : (p->idx > q->idx);
}

/* Karras determine_range / find_split over 32- or 64-bit keys (bvh.cuh:109-199) */
typedef struct { const uint32_t *c32; const uint64_t *c64; } keyset;
static inline int delta_at(keyset k, uint32_t a, uint32_t b)
{
    return k.c64 ? clz64(k.c64[a] ^ k.c64[b]) : clz32(k.c32[a] ^ k.c32[b]);
}
static inline int keys_equal(keyset k, uint32_t a, uint32_t b)
{
    return k.c64 ? k.c64[a] == k.c64[b] : k.c32[a] == k.c32[b];
}
static void determine_range(keyset k, uint32_t nl, uint32_t idx, uint32_t *first, uint32_t *last)
{
    int L, R, d, dmin, lmax, delta, l, t, itmp;
    uint32_t jdx;
    if (idx == 0) { *first = 0; *last = nl - 1; return; }
    L = delta_at(k, idx, idx - 1);
    R = delta_at(k, idx, idx + 1);
    d = (R > L) ? 1 : -1;
    dmin = L < R ? L : R;
    lmax = 2;
    delta = -1;
    itmp = (int)idx + d * lmax;
    if (0 <= itmp && itmp < (int)nl) delta = delta_at(k, idx, (uint32_t)itmp);
    while (delta > dmin) {
        lmax <<= 1;
        itmp = (int)idx + d * lmax;
        delta = -1;
        if (0 <= itmp && itmp < (int)nl) delta = delta_at(k, idx, (uint32_t)itmp);
    }
    l = 0;
    t = lmax >> 1;
    while (t > 0) {
        itmp = (int)idx + (l + t) * d;
        delta = -1;
        if (0 <= itmp && itmp < (int)nl) delta = delta_at(k, idx, (uint32_t)itmp);
        if (delta > dmin) l += t;
        t >>= 1;
    }
    jdx = (uint32_t)((int)idx + l * d);
    if (d < 0) { *first = jdx; *last = idx; }
    else { *first = idx; *last = jdx; }
}
static uint32_t find_split(keyset k, uint32_t first, uint32_t last)
{
    int dnode, split, stride;
    if (keys_equal(k, first, last)) return (first + last) >> 1;
    dnode = delta_at(k, first, last);
    split = (int)first;
    stride = (int)(last - first);
    do {
        int mid;
        stride = (stride + 1) >> 1;
        mid = split + stride;
        if (mid < (int)last) {
            int dd = delta_at(k, first, (uint32_t)mid);
            if (dd > dnode) split = mid;
        }
    } while (stride > 1);
    return (uint32_t)split;
}

/*
 * Full construct(): outputs in the reference's node numbering (internal
 * 0..N-2, leaves N-1..2N-2).  nodes: 4 x uint32 per node (parent, left,
 * right, object); boxes: 6 floats (lo xyz, hi xyz); cones: 5 floats (axis
 * xyz, half_angle, radius).  Also returns sorted codes and object order.
 * Returns 1 if a Morton collision forced 64-bit keys, 0 otherwise, -1 on
 * empty input.
 */
int orc_build(const float *V, const int32_t *F, int32_t nF, const int32_t *slots, const int32_t *owned,
              uint32_t *nodes, float *boxes, float *cones, uint32_t *sorted_codes, uint32_t *sorted_idx,
              float *scene_box, int flags_opt)
{
    uint32_t N = (uint32_t)nF, NI, NN, i;
    box3 *bx, whole;
    cone3 *cn;
    code_idx *ci;
    uint32_t *codes;
    uint64_t *codes64 = NULL;
    int collision = 0;
    int *flags;
    keyset ks;
    if (nF <= 0) return -1;
    g_conservative_radius = flags_opt & 1;
    NI = N - 1;
    NN = 2 * N - 1;
    bx = (box3 *)malloc(sizeof(box3) * NN);
    cn = (cone3 *)malloc(sizeof(cone3) * NN);
    ci = (code_idx *)malloc(sizeof(code_idx) * N);
    codes = (uint32_t *)malloc(sizeof(uint32_t) * N);
    /* defaults (bvh.cuh:395-420) */
    for (i = 0; i < NN; ++i) {
        bx[i].hi = mk(-INFINITY, -INFINITY, -INFINITY);
        bx[i].lo = mk(INFINITY, INFINITY, INFINITY);
        cn[i].axis = mk(0.0f, 0.0f, 0.0f);
        cn[i].half_angle = F_PI;
        cn[i].radius = 0.0f;
    }
    /* leaf boxes and cones in object order, written at NI.. (bvh.cuh:423-426) */
    for (i = 0; i < N; ++i) {
        bx[NI + i] = leaf_box(V, F, (int32_t)i);
        cn[NI + i] = leaf_cone(V, F, slots, owned, (int32_t)i);
    }
    /* whole box (bvh.cuh:429-432) */
    whole.hi = mk(-INFINITY, -INFINITY, -INFINITY);
    whole.lo = mk(INFINITY, INFINITY, INFINITY);
    for (i = 0; i < N; ++i) whole = box_merge(whole, bx[NI + i]);
    if (scene_box) {
        scene_box[0] = whole.lo.x; scene_box[1] = whole.lo.y; scene_box[2] = whole.lo.z;
        scene_box[3] = whole.hi.x; scene_box[4] = whole.hi.y; scene_box[5] = whole.hi.z;
    }
    /* Morton codes (bvh.cuh:292-302) */
    for (i = 0; i < N; ++i) {
        f3 p = box_centroid(bx[NI + i]);
        p.x -= whole.lo.x;
        p.y -= whole.lo.y;
        p.z -= whole.lo.z;
        p.x /= (whole.hi.x - whole.lo.x);
        p.y /= (whole.hi.y - whole.lo.y);
        p.z /= (whole.hi.z - whole.lo.z);
        ci[i].code = morton3(p);
        ci[i].idx = i;
    }
    /* stable sort by code == sort by (code, idx) (bvh.cuh:452-454) */
    qsort(ci, N, sizeof(code_idx), cmp_code_idx);
    {
        box3 *tb = (box3 *)malloc(sizeof(box3) * N);
        cone3 *tc = (cone3 *)malloc(sizeof(cone3) * N);
        for (i = 0; i < N; ++i) { tb[i] = bx[NI + ci[i].idx]; tc[i] = cn[NI + ci[i].idx]; }
        for (i = 0; i < N; ++i) { bx[NI + i] = tb[i]; cn[NI + i] = tc[i]; }
        free(tb); free(tc);
    }
    for (i = 0; i < N; ++i) {
        codes[i] = ci[i].code;
        if (sorted_codes) sorted_codes[i] = ci[i].code;
        if (sorted_idx) sorted_idx[i] = ci[i].idx;
    }
    /* uniqueness (bvh.cuh:459-476) */
    for (i = 1; i < N; ++i)
        if (codes[i] == codes[i - 1]) { collision = 1; break; }
    if (collision) {
        codes64 = (uint64_t *)malloc(sizeof(uint64_t) * N);
        for (i = 0; i < N; ++i) codes64[i] = ((uint64_t)codes[i] << 32) | ci[i].idx;
    }
    ks.c32 = codes;
    ks.c64 = codes64;
    /* nodes (bvh.cuh:481-499) */
    for (i = 0; i < NN; ++i) {
        nodes[4 * i + 0] = INVALID; nodes[4 * i + 1] = INVALID;
        nodes[4 * i + 2] = INVALID; nodes[4 * i + 3] = INVALID;
    }
    for (i = 0; i < N; ++i) nodes[4 * (NI + i) + 3] = ci[i].idx;
    /* Karras (bvh.cuh:200-229) */
    for (i = 0; i < NI; ++i) {
        uint32_t f, l, g, left, right;
        determine_range(ks, N, i, &f, &l);
        g = find_split(ks, f, l);
        left = g;
        right = g + 1;
        if ((f < l ? f : l) == g) left += NI;
        if ((f > l ? f : l) == g + 1) right += NI;
        nodes[4 * i + 3] = INVALID;
        nodes[4 * i + 1] = left;
        nodes[4 * i + 2] = right;
        nodes[4 * left + 0] = i;
        nodes[4 * right + 0] = i;
    }
    /* AABB refit then cone refit (bvh.cuh:520-604), sequential flag climb */
    flags = (int *)calloc(NI > 0 ? NI : 1, sizeof(int));
    for (i = NI; i < NN; ++i) {
        uint32_t p = nodes[4 * i + 0];
        while (p != INVALID) {
            if (flags[p] == 0) { flags[p] = 1; break; }
            bx[p] = box_merge(bx[nodes[4 * p + 1]], bx[nodes[4 * p + 2]]);
            p = nodes[4 * p + 0];
        }
    }
    memset(flags, 0, sizeof(int) * (NI > 0 ? NI : 1));
    for (i = NI; i < NN; ++i) {
        uint32_t p = nodes[4 * i + 0];
        while (p != INVALID) {
            uint32_t l, r;
            if (flags[p] == 0) { flags[p] = 1; break; }
            l = nodes[4 * p + 1];
            r = nodes[4 * p + 2];
            cn[p] = cone_merge(cn[l], cn[r], box_centroid(bx[l]), box_centroid(bx[r]), box_centroid(bx[p]));
            p = nodes[4 * p + 0];
        }
    }
    for (i = 0; i < NN; ++i) {
        boxes[6 * i + 0] = bx[i].lo.x; boxes[6 * i + 1] = bx[i].lo.y; boxes[6 * i + 2] = bx[i].lo.z;
        boxes[6 * i + 3] = bx[i].hi.x; boxes[6 * i + 4] = bx[i].hi.y; boxes[6 * i + 5] = bx[i].hi.z;
        cones[5 * i + 0] = cn[i].axis.x; cones[5 * i + 1] = cn[i].axis.y; cones[5 * i + 2] = cn[i].axis.z;
        cones[5 * i + 3] = cn[i].half_angle; cones[5 * i + 4] = cn[i].radius;
    }
    free(flags); free(bx); free(cn); free(ci); free(codes); free(codes64);
    return collision;
}

/* Single primitive helpers exposed for unit tests */
void orc_leaf_box_cone(const float *V, const int32_t *F, int32_t nF, const int32_t *slots, const int32_t *owned,
                       float *boxes, float *cones, int flags_opt)
{
    int32_t i;
    g_conservative_radius = flags_opt & 1;
    for (i = 0; i < nF; ++i) {
        box3 b = leaf_box(V, F, i);
        cone3 c = leaf_cone(V, F, slots, owned, i);
        boxes[6 * i + 0] = b.lo.x; boxes[6 * i + 1] = b.lo.y; boxes[6 * i + 2] = b.lo.z;
        boxes[6 * i + 3] = b.hi.x; boxes[6 * i + 4] = b.hi.y; boxes[6 * i + 5] = b.hi.z;
        cones[5 * i + 0] = c.axis.x; cones[5 * i + 1] = c.axis.y; cones[5 * i + 2] = c.axis.z;
        cones[5 * i + 3] = c.half_angle; cones[5 * i + 4] = c.radius;
    }
}

uint32_t orc_morton3(float x, float y, float z) { return morton3(mk(x, y, z)); }
uint32_t orc_expand_bits(uint32_t v) { return expand_bits(v); }

void orc_cone_merge(const float *a, const float *b, const float *oa, const float *ob, const float *on, float *out,
                    int flags_opt)
{
    cone3 ca, cb, r;
    g_conservative_radius = flags_opt & 1;
    ca.axis = mk(a[0], a[1], a[2]); ca.half_angle = a[3]; ca.radius = a[4];
    cb.axis = mk(b[0], b[1], b[2]); cb.half_angle = b[3]; cb.radius = b[4];
    r = cone_merge(ca, cb, mk(oa[0], oa[1], oa[2]), mk(ob[0], ob[1], ob[2]), mk(on[0], on[1], on[2]));
    out[0] = r.axis.x; out[1] = r.axis.y; out[2] = r.axis.z; out[3] = r.half_angle; out[4] = r.radius;
}

int orc_cone_overlap(const float *cone, const float *o, const float *box, float dist2)
{
    cone3 c;
    box3 b;
    c.axis = mk(cone[0], cone[1], cone[2]); c.half_angle = cone[3]; c.radius = cone[4];
    b.lo = mk(box[0], box[1], box[2]); b.hi = mk(box[3], box[4], box[5]);
    return cone_overlap(c, mk(o[0], o[1], o[2]), b, dist2);
}

void orc_math(int fn, const float *x, const float *y, float *out, int32_t n)
{
    int32_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        switch (fn) {
        case 0: out[i] = odm_acosf(x[i]); break;
        case 1: out[i] = odm_asinf(x[i]); break;
        case 2: out[i] = odm_atan2f(x[i], y[i]); break;
        case 3: { float s, c; odm_sincosf(x[i], &s, &c); out[i] = s; } break;
        default: { float s, c; odm_sincosf(x[i], &s, &c); out[i] = c; } break;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Queries in the reference's visit order                               */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t n_obj;
    const uint32_t *nodes; /* 4 per node */
    const float *boxes;    /* 6 per node */
    const float *cones;    /* 5 per node */
    const float *V;
    const int32_t *F;
    const int32_t *slots;
    const int32_t *owned;
} orc_tree;

static inline box3 tbox(const orc_tree *t, uint32_t i)
{
    box3 b;
    const float *p = t->boxes + 6 * (size_t)i;
    b.lo = mk(p[0], p[1], p[2]);
    b.hi = mk(p[3], p[4], p[5]);
    return b;
}
static inline cone3 tcone(const orc_tree *t, uint32_t i)
{
    cone3 c;
    const float *p = t->cones + 5 * (size_t)i;
    c.axis = mk(p[0], p[1], p[2]);
    c.half_angle = p[3];
    c.radius = p[4];
    return c;
}
#define T_LEFT(t, i) ((t)->nodes[4 * (size_t)(i) + 1])
#define T_RIGHT(t, i) ((t)->nodes[4 * (size_t)(i) + 2])
#define T_OBJ(t, i) ((t)->nodes[4 * (size_t)(i) + 3])

static float tri_dist(const orc_tree *t, uint32_t obj, f3 q)
{
    const int32_t *f = t->F + 3 * (size_t)obj;
    return closest_point_triangle(vget(t->V, f[0]), vget(t->V, f[1]), vget(t->V, f[2]), q);
}

/* query.cuh:238-318; counts[0] += child box tests, counts[1] += leaf evaluations */
/* tie_min = 0: among exactly tied triangles the last visited wins, as in
 * query.cuh:286 (d <= best); tie_min = 1: the smallest index wins, the
 * library's rule (used by the walk-on-stars restatement, whose walks then
 * match the library's bit for bit).  Every tied candidate is visited either
 * way: nodes at distance == best are not pruned. */
static void cp_one(const orc_tree *t, f3 q, uint32_t *idx_out, float *dist_out, uint64_t *counts, int tie_min)

{
    uint32_t sn[128];
    float sd[128];
    int sp = 0;
    uint32_t nearest = INVALID;
    float best = INFINITY;
    if (t->n_obj == 1) { /* Q6 */
        float d = tri_dist(t, T_OBJ(t, 0), q);
        d *= d;
        *idx_out = T_OBJ(t, 0);
        *dist_out = sqrtf(d);
        if (counts) counts[1] += 1;
        return;
    }
    sn[sp] = 0;
    sd[sp] = box_mindist(tbox(t, 0), q);
    ++sp;
    do {
        uint32_t node, L, R;
        float nd, Lmd, Rmd, Lmm, Rmm;
        box3 Lb, Rb;
        --sp;
        node = sn[sp];
        nd = sd[sp];
        if (nd > best) continue;
        L = T_LEFT(t, node);
        R = T_RIGHT(t, node);
        Lb = tbox(t, L);
        Rb = tbox(t, R);
        Lmd = box_mindist(Lb, q);
        Rmd = box_mindist(Rb, q);
        Lmm = box_minmaxdist(Lb, q);
        Rmm = box_minmaxdist(Rb, q);
        if (counts) counts[0] += 2;
        if (Lmd <= Rmm) {
            uint32_t o = T_OBJ(t, L);
            if (o != INVALID) {
                float d = tri_dist(t, o, q);
                if (counts) counts[1] += 1;
                d *= d;
                if (tie_min ? (d < best || (d == best && o < nearest)) : d <= best) { best = d; nearest = o; }
            } else {
                sn[sp] = L; sd[sp] = Lmd; ++sp;
            }
        }
        if (Rmd <= Lmm) {
            uint32_t o = T_OBJ(t, R);
            if (o != INVALID) {
                float d = tri_dist(t, o, q);
                if (counts) counts[1] += 1;
                d *= d;
                if (tie_min ? (d < best || (d == best && o < nearest)) : d <= best) { best = d; nearest = o; }
            } else {
                sn[sp] = R; sd[sp] = Rmd; ++sp;
            }
        }
    } while (sp > 0);
    *idx_out = nearest;
    *dist_out = sqrtf(best);
}

/* leaf evaluation for the silhouette query: scene.cuh:978-1003 */
static int sil_leaf(const orc_tree *t, uint32_t obj, f3 q, float max_r2, float *dist, uint32_t *edge, int flip,
                    uint64_t *counts)
{
    int k, ret = 0;
    float mr2 = max_r2;
    for (k = 0; k < 3; ++k) {
        int32_t e = t->owned[3 * (size_t)obj + k];
        if (e != -1) {
            if (counts) counts[2] += 1;
            if (edge_closest_silhouette(t->V, t->slots + 4 * (size_t)e, q, mr2, dist, flip)) {
                ret = 1;
                mr2 = (*dist) * (*dist);
                *edge = (uint32_t)e;
            }
        }
    }
    return ret;
}

/* query.cuh:325-423, plus the accepted edge id (Q3).
 * counts[0] child cone/box tests, counts[1] leaf evaluations, counts[2] edge tests */
static void sil_one(const orc_tree *t, f3 q, int flip, uint32_t *edge_out, float *dist_out, uint64_t *counts)
{
    uint32_t sn[128];
    float sd[128];
    int sp = 0;
    float best = INFINITY;
    uint32_t best_edge = INVALID;
    if (t->n_obj == 1) { /* Q6: the single leaf is evaluated as a hit child would be */
        float d = INFINITY;
        uint32_t e = INVALID;
        if (cone_valid(tcone(t, 0)) && cone_overlap(tcone(t, 0), q, tbox(t, 0), box_mindist(tbox(t, 0), q))) {
            if (counts) counts[1] += 1;
            if (sil_leaf(t, T_OBJ(t, 0), q, best * best, &d, &e, flip, counts) && d <= best) {
                best = d;
                best_edge = e;
            }
        }
        *edge_out = best_edge;
        *dist_out = best;
        return;
    }
    sn[sp] = 0;
    sd[sp] = box_mindist(tbox(t, 0), q);
    ++sp;
    do {
        uint32_t node, L, R, c;
        float nd, md[2];
        int hit[2], s;
        box3 b[2];
        cone3 cc[2];
        --sp;
        node = sn[sp];
        nd = sd[sp];
        if (nd > best * best) continue;
        L = T_LEFT(t, node);
        R = T_RIGHT(t, node);
        b[0] = tbox(t, L); b[1] = tbox(t, R);
        cc[0] = tcone(t, L); cc[1] = tcone(t, R);
        md[0] = box_mindist(b[0], q);
        md[1] = box_mindist(b[1], q);
        hit[0] = cone_valid(cc[0]) && cone_overlap(cc[0], q, b[0], md[0]);
        hit[1] = cone_valid(cc[1]) && cone_overlap(cc[1], q, b[1], md[1]);
        if (counts) counts[0] += 2;
        for (s = 0; s < 2; ++s) {
            if (!hit[s]) continue;
            c = s == 0 ? L : R;
            if (T_OBJ(t, c) != INVALID) {
                float d = INFINITY;
                uint32_t e = INVALID;
                if (counts) counts[1] += 1;
                if (sil_leaf(t, T_OBJ(t, c), q, best * best, &d, &e, flip, counts) && d <= best) {
                    best = d;
                    best_edge = e;
                }
            } else {
                sn[sp] = c; sd[sp] = md[s]; ++sp;
            }
        }
    } while (sp > 0);
    *edge_out = best_edge;
    *dist_out = best;
}

/* query.cuh:79-169 (closest hit; any_hit = TestOnly) */
/* tie_min: as for cp_one; the reference keeps the first hit found among
 * equal t (query.cuh:115, th < min_dist) */
static void ray_one(const orc_tree *t, f3 o, f3 d, float max_dist, int any_hit, uint32_t *idx_out, float *t_out,
                    float *uv_out, uint8_t *hit_out, uint64_t *counts, int tie_min)

{
    uint32_t sn[128];
    float sd[128];
    int sp = 0, found = 0;
    float min_dist = INFINITY, uu = 0.0f, vv = 0.0f;
    uint32_t nearest = INVALID;
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z); /* aabb.cuh:304-312 */
    sn[sp] = 0;
    sd[sp] = INFINITY;
    ++sp;
    do {
        uint32_t node, obj;
        float nt;
        --sp;
        node = sn[sp];
        nt = sd[sp];
        if (nt > min_dist) continue;
        obj = T_OBJ(t, node);
        if (obj != INVALID) {
            const int32_t *f = t->F + 3 * (size_t)obj;
            float th, uh, vh;
            if (counts) counts[1] += 1;
            if (ray_triangle(o, d, vget(t->V, f[0]), vget(t->V, f[1]), vget(t->V, f[2]), &th, &uh, &vh) &&
                th < max_dist && (th < min_dist || (tie_min && th == min_dist && obj < nearest))) {
                min_dist = th;
                found = 1;
                if (any_hit) break;
                uu = uh;
                vv = vh;
                nearest = obj;
            }
        } else {
            uint32_t L = T_LEFT(t, node), R = T_RIGHT(t, node), closer = L, other = R, tmp;
            float Ld, Rd, tf;
            int Lh = box_ray(tbox(t, L), o, inv, max_dist, &Ld);
            int Rh = box_ray(tbox(t, R), o, inv, max_dist, &Rd);
            if (counts) counts[0] += 2;
            if (Lh && Rh) {
                if (Rd < Ld) {
                    tf = Ld; Ld = Rd; Rd = tf;
                    tmp = closer; closer = other; other = tmp;
                }
                sn[sp] = other; sd[sp] = Rd; ++sp;
                sn[sp] = closer; sd[sp] = Ld; ++sp;
            } else if (Lh) {
                sn[sp] = L; sd[sp] = Ld; ++sp;
            } else if (Rh) {
                sn[sp] = R; sd[sp] = Rd; ++sp;
            }
        }
    } while (sp > 0);
    *hit_out = (uint8_t)found;
    if (any_hit) {
        *idx_out = INVALID;
        *t_out = found ? min_dist : INFINITY;
        uv_out[0] = 0.0f; uv_out[1] = 0.0f;
        return;
    }
    *idx_out = nearest;
    *t_out = min_dist;
    uv_out[0] = uu;
    uv_out[1] = vv;
}

/* sample.cuh:23-92 (Q19: pdf 0 when no triangle is returned) */
static void sample_one(const orc_tree *t, f3 c, float r, float u, int32_t *idx_out, float *pdf_out)
{
    uint32_t node = 0;
    float npdf = 1.0f;
    int32_t obj_index = -1;
    float obj_pdf = 0.0f;
    for (;;) {
        uint32_t obj = T_OBJ(t, node);
        if (obj != INVALID) {
            const int32_t *f = t->F + 3 * (size_t)obj;
            f3 pa = vget(t->V, f[0]), pb = vget(t->V, f[1]), pc = vget(t->V, f[2]);
            if (tri_sphere(pa, pb, pc, c, r)) {
                obj_index = (int32_t)obj;
                obj_pdf = npdf;
                obj_pdf /= tri_area(pa, pb, pc);
            }
            break;
        } else {
            uint32_t L = T_LEFT(t, node), R = T_RIGHT(t, node);
            box3 Lb = tbox(t, L), Rb = tbox(t, R);
            float Lw = box_sphere(Lb, c, r) ? green_weight(c, box_centroid(Lb)) : 0.0f;
            float Rw = box_sphere(Rb, c, r) ? green_weight(c, box_centroid(Rb)) : 0.0f;
            float tw = Lw + Rw;
            if (tw > 0.0f) {
                float Lp = Lw / tw;
                if (u < Lp) {
                    u /= Lp;
                    npdf = Lp * npdf;
                    node = L;
                } else {
                    float Rp = 1.0f - Lp;
                    u = (u - Lp) / Rp;
                    npdf = Rp * npdf;
                    node = R;
                }
            } else {
                break;
            }
        }
    }
    *idx_out = obj_index;
    *pdf_out = obj_index >= 0 ? obj_pdf : 0.0f;
}

/* scene.cuh:14-27 */
static f3 sample_triangle(f3 pa, f3 pb, f3 pc, float u, float v)
{
    float w;
    if (u + v > 1.0f) {
        u = 1.0f - u;
        v = 1.0f - v;
    }
    w = 1.0f - u - v;
    return mk(w * pa.x + u * pb.x + v * pc.x, w * pa.y + u * pb.y + v * pc.y, w * pa.z + u * pb.z + v * pc.z);
}

/* ------------------------------------------------------------------ */
/* Walk on stars (Sawhney, Miller, Gkioulekas, Crane, SIGGRAPH 2023) for */
/* the Laplace equation: g = x y on the Dirichlet scene, zero Neumann    */
/* flux on the Neumann scene.  The library's wavefront loop              */
/* (snch_lbvh_amd/csrc/wost.hip) restated one walk at a time with the    */
/* same per-walk FP32 arithmetic and counter-based random streams; the   */
/* queries are the reference-order ones above.  No reference            */
/* counterpart (the reference's WoSt caller lives outside it).           */
/* ------------------------------------------------------------------ */
static void set_threads(int nthreads);

static float wost_rnd(uint64_t seed, uint32_t walk, uint32_t step, uint32_t k) /* splitmix64 finaliser */
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (((uint64_t)walk << 24) ^ ((uint64_t)step << 4) ^ k);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((uint32_t)(z >> 40) + 0.5f) * (1.0f / 16777216.0f);
}

/* Ericson, Real-Time Collision Detection 5.1.5: the closest point itself */
static f3 closest_point_on_tri(f3 a, f3 b, f3 c, f3 p)
{
    f3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a), bp, cp;
    float d1 = dot3(ab, ap), d2 = dot3(ac, ap), d3, d4, d5, d6, vc, vb, va, v, w, den;
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    bp = sub(p, b);
    d3 = dot3(ab, bp);
    d4 = dot3(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        v = d1 / (d1 - d3);
        return mk(a.x + v * ab.x, a.y + v * ab.y, a.z + v * ab.z);
    }
    cp = sub(p, c);
    d5 = dot3(ab, cp);
    d6 = dot3(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        w = d2 / (d2 - d6);
        return mk(a.x + w * ac.x, a.y + w * ac.y, a.z + w * ac.z);
    }
    va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return mk(b.x + w * (c.x - b.x), b.y + w * (c.y - b.y), b.z + w * (c.z - b.z));
    }
    den = 1.0f / (va + vb + vc);
    v = vb * den;
    w = vc * den;
    return mk(a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w);
}

static f3 tri_vert(const orc_tree *t, uint32_t tri, int k) { return vget(t->V, t->F[3 * (size_t)tri + k]); }

static float wost_g(const orc_tree *D, uint32_t tri, f3 p)
{
    f3 c = closest_point_on_tri(tri_vert(D, tri, 0), tri_vert(D, tri, 1), tri_vert(D, tri, 2), p);
    return c.x * c.y;
}

void orc_wost(const orc_tree *D, const orc_tree *N, const float *x0, int64_t n, uint32_t max_steps, float eps,
              uint64_t seed, float *out, uint32_t *steps_out, int nthreads)
{
    int64_t wi;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (wi = 0; wi < n; ++wi) {
        const uint32_t wk = (uint32_t)wi;
        f3 p = mk(x0[3 * wi], x0[3 * wi + 1], x0[3 * wi + 2]), nrm = mk(0.0f, 0.0f, 0.0f);
        int on_bnd = 0;
        uint32_t steps = 0;
        float value = 0.0f;
        for (;;) {
            uint32_t cidx, sidx, ridx;
            float dD, dS = INFINITY, R, u1, u2, z, s, sp, cp, rt, uv[2];
            uint8_t hit = 0;
            f3 d, o;
            cp_one(D, p, &cidx, &dD, NULL, 1);
            if (N) sil_one(N, p, 0, &sidx, &dS, NULL);
            if (dD < eps) {
                value = wost_g(D, cidx, p);
                break;
            }
            R = dD < dS ? dD : dS;
            u1 = wost_rnd(seed, wk, steps, 0);
            u2 = wost_rnd(seed, wk, steps, 1);
            z = 1.0f - 2.0f * u1;
            s = sqrtf(smax(0.0f, 1.0f - z * z));
            odm_sincosf(6.283185307179586f * u2, &sp, &cp);
            d = mk(s * cp, s * sp, z);
            o = p;
            if (on_bnd) {
                float dn = d.x * nrm.x + d.y * nrm.y + d.z * nrm.z;
                if (dn < 0.0f) d = mk(d.x - 2.0f * dn * nrm.x, d.y - 2.0f * dn * nrm.y, d.z - 2.0f * dn * nrm.z);
                o = mk(p.x + 1e-5f * nrm.x, p.y + 1e-5f * nrm.y, p.z + 1e-5f * nrm.z);
            }
            if (N) ray_one(N, o, d, R, 0, &ridx, &rt, uv, &hit, NULL, 1);
            if (N && hit && rt < R) {
                f3 a = tri_vert(N, ridx, 0), b = tri_vert(N, ridx, 1), c = tri_vert(N, ridx, 2);
                f3 nn = normalize3(cross3(sub(b, a), sub(c, a)));
                if (dot3(nn, d) > 0.0f) nn = mk(-nn.x, -nn.y, -nn.z);
                p = mk(o.x + rt * d.x, o.y + rt * d.y, o.z + rt * d.z);
                nrm = nn;
                on_bnd = 1;
            } else {
                p = mk(p.x + R * d.x, p.y + R * d.y, p.z + R * d.z);
                on_bnd = 0;
            }
            ++steps;
            if (steps >= max_steps) {
                value = wost_g(D, cidx, p);
                break;
            }
        }
        out[wi] = value;
        if (steps_out) steps_out[wi] = steps;
    }
}

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

void orc_closest_point(const orc_tree *t, const float *q, int64_t n, uint32_t *idx, float *dist, uint64_t *counts,
                       int nthreads)
{
    int64_t i;
    uint64_t c0 = 0, c1 = 0;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : c0, c1)
    for (i = 0; i < n; ++i) {
        uint64_t c[3] = {0, 0, 0};
        cp_one(t, mk(q[3 * i], q[3 * i + 1], q[3 * i + 2]), idx + i, dist + i, counts ? c : NULL, 0);
        c0 += c[0];
        c1 += c[1];
    }
    if (counts) { counts[0] = c0; counts[1] = c1; }
}

void orc_closest_silhouette(const orc_tree *t, const float *q, int64_t n, int flip, uint32_t *edge, float *dist,
                            uint64_t *counts, int nthreads)
{
    int64_t i;
    uint64_t c0 = 0, c1 = 0, c2 = 0;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : c0, c1, c2)
    for (i = 0; i < n; ++i) {
        uint64_t c[3] = {0, 0, 0};
        sil_one(t, mk(q[3 * i], q[3 * i + 1], q[3 * i + 2]), flip, edge + i, dist + i, counts ? c : NULL);
        c0 += c[0];
        c1 += c[1];
        c2 += c[2];
    }
    if (counts) { counts[0] = c0; counts[1] = c1; counts[2] = c2; }
}

void orc_ray(const orc_tree *t, const float *o, const float *d, const float *tmax, int64_t n, int any_hit,
             uint32_t *idx, float *tout, float *uv, uint8_t *hit, uint64_t *counts, int nthreads)
{
    int64_t i;
    uint64_t c0 = 0, c1 = 0;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : c0, c1)
    for (i = 0; i < n; ++i) {
        uint64_t c[3] = {0, 0, 0};
        ray_one(t, mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]), tmax[i],
                any_hit, idx + i, tout + i, uv + 2 * i, hit + i, counts ? c : NULL, 0);
        c0 += c[0];
        c1 += c[1];
    }
    if (counts) { counts[0] = c0; counts[1] = c1; }
}

void orc_sample_in_sphere(const orc_tree *t, const float *c, const float *r, const float *u, int64_t n,
                          int32_t *idx, float *pdf, float *pt, int nthreads)
{
    int64_t i;
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        sample_one(t, mk(c[3 * i], c[3 * i + 1], c[3 * i + 2]), r[i], u[3 * i], idx + i, pdf + i);
        if (pt) {
            if (idx[i] >= 0) {
                const int32_t *f = t->F + 3 * (size_t)idx[i];
                f3 p = sample_triangle(vget(t->V, f[0]), vget(t->V, f[1]), vget(t->V, f[2]), u[3 * i + 1],
                                       u[3 * i + 2]);
                pt[3 * i] = p.x; pt[3 * i + 1] = p.y; pt[3 * i + 2] = p.z;
            } else {
                pt[3 * i] = pt[3 * i + 1] = pt[3 * i + 2] = 0.0f;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* FP64 brute force (independent of the tree)                          */
/* ------------------------------------------------------------------ */
typedef struct { double x, y, z; } d3;
static inline d3 dmk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 dsub(d3 a, d3 b) { return dmk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 dcross(d3 u, d3 v) { return dmk(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x); }
static inline double dlen(d3 a) { return sqrt(ddot(a, a)); }
static inline d3 dget(const float *V, int i) { return dmk(V[3 * i], V[3 * i + 1], V[3 * i + 2]); }

/* Ericson, Real-Time Collision Detection 5.1.5, in double */
static double bf_tri_dist(d3 a, d3 b, d3 c, d3 p)
{
    d3 ab = dsub(b, a), ac = dsub(c, a), ap = dsub(p, a), bp, cp, q;
    double d1 = ddot(ab, ap), d2 = ddot(ac, ap), d3_, d4, d5, d6, vc, vb, va, v, w, den;
    if (d1 <= 0 && d2 <= 0) return dlen(ap);
    bp = dsub(p, b);
    d3_ = ddot(ab, bp);
    d4 = ddot(ac, bp);
    if (d3_ >= 0 && d4 <= d3_) return dlen(bp);
    vc = d1 * d4 - d3_ * d2;
    if (vc <= 0 && d1 >= 0 && d3_ <= 0) {
        v = d1 / (d1 - d3_);
        q = dmk(a.x + v * ab.x, a.y + v * ab.y, a.z + v * ab.z);
        return dlen(dsub(p, q));
    }
    cp = dsub(p, c);
    d5 = ddot(ab, cp);
    d6 = ddot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return dlen(cp);
    vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        w = d2 / (d2 - d6);
        q = dmk(a.x + w * ac.x, a.y + w * ac.y, a.z + w * ac.z);
        return dlen(dsub(p, q));
    }
    va = d3_ * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3_) >= 0 && (d5 - d6) >= 0) {
        w = (d4 - d3_) / ((d4 - d3_) + (d5 - d6));
        q = dmk(b.x + w * (c.x - b.x), b.y + w * (c.y - b.y), b.z + w * (c.z - b.z));
        return dlen(dsub(p, q));
    }
    den = 1.0 / (va + vb + vc);
    v = vb * den;
    w = vc * den;
    q = dmk(a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w);
    return dlen(dsub(p, q));
}

void orc_bf_closest_point(const float *V, const int32_t *F, int32_t nF, const float *q, int64_t n, uint32_t *idx,
                          double *dist, int nthreads)
{
    int64_t i;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (i = 0; i < n; ++i) {
        d3 p = dmk(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
        double best = INFINITY;
        uint32_t bi = INVALID;
        int32_t f;
        for (f = 0; f < nF; ++f) {
            double d = bf_tri_dist(dget(V, F[3 * f]), dget(V, F[3 * f + 1]), dget(V, F[3 * f + 2]), p);
            if (d < best) { best = d; bi = (uint32_t)f; }
        }
        idx[i] = bi;
        dist[i] = best;
    }
}

/* Silhouette predicate of scene.cuh:143-174 and :788-824 evaluated in double. */
static int bf_edge_sil(const float *V, const int32_t *s, d3 o, int flip, double *dist)
{
    d3 pa = dget(V, s[1]), pb = dget(V, s[2]), u = dsub(pb, pa), v = dsub(o, pa), cp;
    double c1 = ddot(u, v), c2, d;
    const double precision = 1e-3;
    double sign = flip ? 1.0 : -1.0;
    if (c1 <= 0) cp = pa;
    else {
        c2 = ddot(u, u);
        if (c2 <= c1) cp = pb;
        else {
            double t = c1 / c2;
            cp = dmk(pa.x + u.x * t, pa.y + u.y * t, pa.z + u.z * t);
        }
    }
    d = dlen(dsub(o, cp));
    *dist = d;
    if (s[0] == -1 || s[3] == -1) return 1;
    {
        d3 a0 = dget(V, s[1]), b0 = dget(V, s[2]), c0 = dget(V, s[3]);
        d3 a1 = dget(V, s[2]), b1 = dget(V, s[1]), c1v = dget(V, s[0]);
        d3 n0 = dcross(dsub(b0, a0), dsub(c0, a0)), n1 = dcross(dsub(b1, a1), dsub(c1v, a1));
        double l0 = dlen(n0), l1 = dlen(n1), dot0, dot1;
        d3 view = dsub(o, cp);
        n0 = dmk(n0.x / l0, n0.y / l0, n0.z / l0);
        n1 = dmk(n1.x / l1, n1.y / l1, n1.z / l1);
        if (d <= precision) {
            d3 ed = dsub(pb, pa);
            double el = dlen(ed), ang;
            ed = dmk(ed.x / el, ed.y / el, ed.z / el);
            ang = atan2(ddot(ed, dcross(n0, n1)), ddot(n0, n1));
            return sign * ang > precision;
        }
        dot0 = ddot(view, n0);
        dot1 = ddot(view, n1);
        if (fabs(dot0) <= precision) return sign * dot1 > precision;
        if (fabs(dot1) <= precision) return sign * dot0 > precision;
        return dot0 * dot1 < 0;
    }
}

void orc_bf_closest_silhouette(const float *V, const int32_t *slots, int32_t nE, const float *q, int64_t n, int flip,
                               uint32_t *edge, double *dist, int nthreads)
{
    int64_t i;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (i = 0; i < n; ++i) {
        d3 p = dmk(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
        double best = INFINITY, d;
        uint32_t be = INVALID;
        int32_t e;
        for (e = 0; e < nE; ++e) {
            if (bf_edge_sil(V, slots + 4 * e, p, flip, &d) && d < best) { best = d; be = (uint32_t)e; }
        }
        edge[i] = be;
        dist[i] = best;
    }
}

void orc_bf_ray(const float *V, const int32_t *F, int32_t nF, const float *o, const float *d, const float *tmax,
                int64_t n, uint32_t *idx, double *tout, int nthreads)
{
    int64_t i;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (i = 0; i < n; ++i) {
        d3 oo = dmk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), dd = dmk(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        double best = INFINITY;
        uint32_t bi = INVALID;
        int32_t f;
        for (f = 0; f < nF; ++f) {
            d3 v0 = dget(V, F[3 * f]), v1 = dget(V, F[3 * f + 1]), v2 = dget(V, F[3 * f + 2]);
            d3 e1 = dsub(v1, v0), e2 = dsub(v2, v0), h = dcross(dd, e2), s, qq;
            double det = ddot(e1, h), inv, u, v, t;
            if (fabs(det) < (double)F_EPS) continue;
            inv = 1.0 / det;
            s = dsub(oo, v0);
            u = ddot(s, h) * inv;
            if (u < 0 || u > 1) continue;
            qq = dcross(s, e1);
            v = ddot(dd, qq) * inv;
            if (v < 0 || u + v > 1) continue;
            t = ddot(e2, qq) * inv;
            if (t >= 0 && t < (double)tmax[i] && t < best) { best = t; bi = (uint32_t)f; }
        }
        idx[i] = bi;
        tout[i] = best;
    }
}

/* FP64 distance of a point to one triangle / one edge's segment, for tie checks */
void orc_tri_dist64(const float *V, const int32_t *F, const uint32_t *tri, const float *q, int64_t n, double *out)
{
    int64_t i;
    for (i = 0; i < n; ++i) {
        const int32_t *f = F + 3 * (size_t)tri[i];
        out[i] = bf_tri_dist(dget(V, f[0]), dget(V, f[1]), dget(V, f[2]), dmk(q[3 * i], q[3 * i + 1], q[3 * i + 2]));
    }
}

void orc_tri_dist32(const float *V, const int32_t *F, const uint32_t *tri, const float *q, int64_t n, float *out)
{
    int64_t i;
    for (i = 0; i < n; ++i) {
        const int32_t *f = F + 3 * (size_t)tri[i];
        out[i] = closest_point_triangle(vget(V, f[0]), vget(V, f[1]), vget(V, f[2]), mk(q[3 * i], q[3 * i + 1], q[3 * i + 2]));
    }
}

/* FP32 silhouette candidate evaluation of one edge (reference predicate) */
void orc_edge_sil32(const float *V, const int32_t *slots, const uint32_t *edge, const float *q, int64_t n, int flip,
                    float *out)
{
    int64_t i;
    for (i = 0; i < n; ++i) {
        float d = INFINITY;
        if (!edge_closest_silhouette(V, slots + 4 * (size_t)edge[i], mk(q[3 * i], q[3 * i + 1], q[3 * i + 2]),
                                     INFINITY, &d, flip))
            d = INFINITY;
        out[i] = d;
    }
}

/* FP64 silhouette distance of one edge under the reference predicate (inf if not a silhouette) */
void orc_edge_sil64(const float *V, const int32_t *slots, const uint32_t *edge, const float *q, int64_t n, int flip,
                    double *out)
{
    int64_t i;
    for (i = 0; i < n; ++i) {
        double d;
        if (edge[i] == INVALID || !bf_edge_sil(V, slots + 4 * (size_t)edge[i], dmk(q[3 * i], q[3 * i + 1], q[3 * i + 2]), flip, &d))
            d = INFINITY;
        out[i] = d;
    }
}

/* ================================================================== */
/* scene<2>: line segments with silhouette vertices (scene.cuh:287-703) */
/* Boxes and cones use the 3D records with z = 0.                     */
/* ================================================================== */
typedef struct { float x, y; } f2;
static inline f2 mk2(float x, float y) { f2 r = {x, y}; return r; }
static inline f2 vget2(const float *V, int i) { return mk2(V[2 * i], V[2 * i + 1]); }
static inline float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
static inline float len2(f2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
static inline f2 normalize2(f2 v) { float n = len2(v); return mk2(v.x / n, v.y / n); }
#define BVH_OFFSET 1e-3f /* scene.cuh:12 */

/* compute_silhouettes (scene.cuh:631-652) and build_bvh ownership (:654-672).
 * slots: 4 per vertex (prev, v, next, -1); owned: 2 per segment. */
void orc_silhouettes2d(const int32_t *S, int32_t nS, int32_t nV, int32_t *slots, int32_t *owned)
{
    int32_t i;
    unsigned char *seen = (unsigned char *)calloc((size_t)(nV > 0 ? nV : 1), 1);
    for (i = 0; i < 4 * nV; ++i) slots[i] = -1;
    for (i = 0; i < nS; ++i) {
        int32_t a = S[2 * i], b = S[2 * i + 1];
        slots[4 * a + 1] = a;
        slots[4 * a + 2] = b;
        slots[4 * b + 0] = a;
        slots[4 * b + 1] = b;
    }
    for (i = 0; i < nS; ++i) {
        int k, ptr = 0;
        owned[2 * i] = owned[2 * i + 1] = -1;
        for (k = 0; k < 2; ++k) {
            int32_t v = S[2 * i + k];
            if (!seen[v]) {
                seen[v] = 1;
                owned[2 * i + ptr++] = v;
            }
        }
    }
    free(seen);
}

static inline int sv_has_face(const int32_t *s, int f) { return f == 0 ? s[2] != -1 : s[0] != -1; } /* :322-325 */
static f2 sv_face_normal(const float *V, const int32_t *s, int f, int do_norm) /* :326-336 */
{
    int i = f == 0 ? 1 : 0;
    f2 pa = vget2(V, s[i]), pb = vget2(V, s[i + 1]);
    f2 d = mk2(pb.x - pa.x, pb.y - pa.y);
    f2 n = mk2(d.y, -d.x);
    return do_norm ? normalize2(n) : n;
}
static f2 sv_normal(const float *V, const int32_t *s) /* :337-353 */
{
    f2 n = mk2(0.0f, 0.0f);
    if (sv_has_face(s, 0)) { f2 a = sv_face_normal(V, s, 0, 0); n = mk2(n.x + a.x, n.y + a.y); }
    if (sv_has_face(s, 1)) { f2 b = sv_face_normal(V, s, 1, 0); n = mk2(n.x + b.x, n.y + b.y); }
    return normalize2(n);
}

static box3 leaf_box2d(const float *V, const int32_t *S, int32_t i) /* :420-433 */
{
    f2 p0 = vget2(V, S[2 * i]), p1 = vget2(V, S[2 * i + 1]);
    box3 b;
    b.hi = mk(fmaxf(p0.x, p1.x) + BVH_OFFSET, fmaxf(p0.y, p1.y) + BVH_OFFSET, 0.0f);
    b.lo = mk(fminf(p0.x, p1.x) - BVH_OFFSET, fminf(p0.y, p1.y) - BVH_OFFSET, 0.0f);
    return b;
}

static cone3 leaf_cone2d(const float *V, const int32_t *S, const int32_t *slots, const int32_t *owned, int32_t i)
{ /* :435-502 */
    box3 b = leaf_box2d(V, S, i);
    f3 c3 = box_centroid(b);
    f2 c = mk2(c3.x, c3.y), ax = mk2(0.0f, 0.0f);
    cone3 r;
    int any = 0, all_two = 1, k;
    r.half_angle = F_PI;
    r.radius = 0.0f;
    for (k = 0; k < 2; ++k) {
        int32_t v = owned[2 * i + k];
        if (v != -1) {
            const int32_t *s = slots + 4 * v;
            f2 n = sv_normal(V, s), p = vget2(V, s[1]);
            ax.x += n.x;
            ax.y += n.y;
            r.radius = smax(r.radius, len2(mk2(p.x - c.x, p.y - c.y)));
            all_two = all_two && sv_has_face(s, 0) && sv_has_face(s, 1);
            any = 1;
        }
    }
    if (!any) {
        r.half_angle = -F_PI;
    } else if (!all_two) {
        r.half_angle = F_PI;
    } else {
        float an = len2(ax);
        if (an > F_EPS) {
            ax.x /= an;
            ax.y /= an;
            r.half_angle = 0.0f;
            for (k = 0; k < 2; ++k) {
                int32_t v = owned[2 * i + k];
                if (v != -1) {
                    const int32_t *s = slots + 4 * v;
                    f2 nf[2];
                    int jj;
                    nf[0] = sv_has_face(s, 0) ? sv_face_normal(V, s, 0, 1) : mk2(0.0f, 0.0f);
                    nf[1] = sv_has_face(s, 1) ? sv_face_normal(V, s, 1, 1) : mk2(0.0f, 0.0f);
                    for (jj = 0; jj < 2; ++jj)
                        r.half_angle = smax(r.half_angle, odm_acosf(clamp11(dot2(ax, nf[jj]))));
                }
            }
        }
    }
    r.axis = mk(ax.x, ax.y, 0.0f);
    return r;
}

static inline uint32_t morton2(f2 p) /* morton_code.cuh:85-92 */
{
    const float res = 1024.0f;
    float x = fminf(fmaxf(p.x * res, 0.0f), res - 1.0f);
    float y = fminf(fmaxf(p.y * res, 0.0f), res - 1.0f);
    return expand_bits((uint32_t)x) * 2u + expand_bits((uint32_t)y);
}

/* cone.cuh:78-121 */
static int cone_overlap2d(cone3 bc, f2 o, box3 b, float dist_to_box)
{
    f2 c, v, e, bb;
    float l, d_axis, d, s, vha, sum;
    if (bc.half_angle >= F_PI_2 || dist_to_box < F_EPS)
        return 1;
    c = mk2((b.hi.x + b.lo.x) * 0.5f, (b.hi.y + b.lo.y) * 0.5f);
    v = mk2(c.x - o.x, c.y - o.y);
    l = len2(v);
    v.x /= l;
    v.y /= l;
    d_axis = odm_acosf(clamp11(dot2(mk2(bc.axis.x, bc.axis.y), v)));
    if (inrange(F_PI_2, d_axis - bc.half_angle, d_axis + bc.half_angle))
        return 1;
    if (l > bc.radius) {
        vha = odm_asinf(bc.radius / l);
        sum = bc.half_angle + vha;
        return sum >= F_PI_2 ? 1 : inrange(F_PI_2, d_axis - sum, d_axis + sum);
    }
    e = mk2(b.hi.x - c.x, b.hi.y - c.y);
    d = dot2(e, mk2(fabsf(v.x), fabsf(v.y)));
    s = l - d;
    if (s <= 0.0f)
        return 1;
    bb = mk2(-v.y, v.x); /* project_to_plane, cone.cuh:44-49 */
    d = fabsf(dot2(e, mk2(fabsf(bb.x), fabsf(bb.y))));
    vha = odm_atan2f(d, s);
    sum = bc.half_angle + vha;
    return sum >= F_PI_2 ? 1 : inrange(F_PI_2, d_axis - sum, d_axis + sum);
}

/* cone.cuh:266-275 and :320-372 (Q1 as in 3D) */
static cone3 cone_merge2d(cone3 a, cone3 b, f3 oa, f3 ob, f3 on)
{
    cone3 r;
    memset(&r, 0, sizeof(r));
    if (cone_valid(a) && cone_valid(b)) {
        f2 axa = mk2(a.axis.x, a.axis.y), axb = mk2(b.axis.x, b.axis.y), t2;
        float ha = a.half_angle, hb = b.half_angle, t, theta, o_theta;
        f2 da = mk2(on.x - oa.x, on.y - oa.y), db = mk2(on.x - ob.x, on.y - ob.y);
        if (g_conservative_radius)
            r.radius = smax(a.radius + len2(da), b.radius + len2(db));
        else
            r.radius = sqrtf(smax(a.radius * a.radius + (da.x * da.x + da.y * da.y),
                                  b.radius * b.radius + (db.x * db.x + db.y * db.y)));
        if (hb > ha) {
            t2 = axa; axa = axb; axb = t2;
            t = ha; ha = hb; hb = t;
        }
        theta = odm_acosf(clamp11(dot2(axa, axb)));
        if (smin(theta + hb, F_PI) <= ha) {
            r.axis = mk(axa.x, axa.y, 0.0f);
            r.half_angle = ha;
            return r;
        }
        o_theta = (ha + theta + hb) / 2.0f;
        if (o_theta >= F_PI) {
            r.axis = mk(axa.x, axa.y, 0.0f);
            r.half_angle = F_PI;
            return r;
        }
        {
            float rt = o_theta - ha, det = axa.x * axb.y - axa.y * axb.x, st, ct;
            rt *= copysignf(1.0f, det);
            odm_sincosf(rt, &st, &ct);
            r.axis = mk(ct * axa.x - st * axa.y, st * axa.x + ct * axa.y, 0.0f);
        }
        r.half_angle = o_theta;
    } else if (cone_valid(a)) {
        r = a;
    } else if (cone_valid(b)) {
        r = b;
    } else {
        r.half_angle = -F_PI;
    }
    return r;
}

int orc_build2d(const float *V, const int32_t *S, int32_t nS, const int32_t *slots, const int32_t *owned,
                uint32_t *nodes, float *boxes, float *cones, uint32_t *sorted_codes, uint32_t *sorted_idx,
                int flags_opt)
{
    uint32_t N = (uint32_t)nS, NI, NN, i;
    box3 *bx, whole;
    cone3 *cn;
    code_idx *ci;
    uint32_t *codes;
    uint64_t *codes64 = NULL;
    int collision = 0, *flags;
    keyset ks;
    if (nS <= 0) return -1;
    g_conservative_radius = flags_opt & 1;
    NI = N - 1;
    NN = 2 * N - 1;
    bx = (box3 *)malloc(sizeof(box3) * NN);
    cn = (cone3 *)malloc(sizeof(cone3) * NN);
    ci = (code_idx *)malloc(sizeof(code_idx) * N);
    codes = (uint32_t *)malloc(sizeof(uint32_t) * N);
    for (i = 0; i < NN; ++i) {
        bx[i].hi = mk(-INFINITY, -INFINITY, 0.0f);
        bx[i].lo = mk(INFINITY, INFINITY, 0.0f);
        cn[i].axis = mk(0.0f, 0.0f, 0.0f);
        cn[i].half_angle = F_PI;
        cn[i].radius = 0.0f;
    }
    for (i = 0; i < N; ++i) {
        bx[NI + i] = leaf_box2d(V, S, (int32_t)i);
        cn[NI + i] = leaf_cone2d(V, S, slots, owned, (int32_t)i);
    }
    whole.hi = mk(-INFINITY, -INFINITY, 0.0f);
    whole.lo = mk(INFINITY, INFINITY, 0.0f);
    for (i = 0; i < N; ++i) whole = box_merge(whole, bx[NI + i]);
    for (i = 0; i < N; ++i) { /* bvh.cuh:269-277 */
        f3 c = box_centroid(bx[NI + i]);
        f2 p = mk2(c.x, c.y);
        p.x -= whole.lo.x;
        p.y -= whole.lo.y;
        p.x /= (whole.hi.x - whole.lo.x);
        p.y /= (whole.hi.y - whole.lo.y);
        ci[i].code = morton2(p);
        ci[i].idx = i;
    }
    qsort(ci, N, sizeof(code_idx), cmp_code_idx);
    {
        box3 *tb = (box3 *)malloc(sizeof(box3) * N);
        cone3 *tc = (cone3 *)malloc(sizeof(cone3) * N);
        for (i = 0; i < N; ++i) { tb[i] = bx[NI + ci[i].idx]; tc[i] = cn[NI + ci[i].idx]; }
        for (i = 0; i < N; ++i) { bx[NI + i] = tb[i]; cn[NI + i] = tc[i]; }
        free(tb); free(tc);
    }
    for (i = 0; i < N; ++i) {
        codes[i] = ci[i].code;
        if (sorted_codes) sorted_codes[i] = ci[i].code;
        if (sorted_idx) sorted_idx[i] = ci[i].idx;
    }
    for (i = 1; i < N; ++i)
        if (codes[i] == codes[i - 1]) { collision = 1; break; }
    if (collision) {
        codes64 = (uint64_t *)malloc(sizeof(uint64_t) * N);
        for (i = 0; i < N; ++i) codes64[i] = ((uint64_t)codes[i] << 32) | ci[i].idx;
    }
    ks.c32 = codes;
    ks.c64 = codes64;
    for (i = 0; i < NN; ++i)
        nodes[4 * i] = nodes[4 * i + 1] = nodes[4 * i + 2] = nodes[4 * i + 3] = INVALID;
    for (i = 0; i < N; ++i) nodes[4 * (NI + i) + 3] = ci[i].idx;
    for (i = 0; i < NI; ++i) {
        uint32_t f, l, g, left, right;
        determine_range(ks, N, i, &f, &l);
        g = find_split(ks, f, l);
        left = g;
        right = g + 1;
        if ((f < l ? f : l) == g) left += NI;
        if ((f > l ? f : l) == g + 1) right += NI;
        nodes[4 * i + 1] = left;
        nodes[4 * i + 2] = right;
        nodes[4 * left] = i;
        nodes[4 * right] = i;
    }
    flags = (int *)calloc(NI > 0 ? NI : 1, sizeof(int));
    for (i = NI; i < NN; ++i) {
        uint32_t p = nodes[4 * i];
        while (p != INVALID) {
            if (flags[p] == 0) { flags[p] = 1; break; }
            bx[p] = box_merge(bx[nodes[4 * p + 1]], bx[nodes[4 * p + 2]]);
            p = nodes[4 * p];
        }
    }
    memset(flags, 0, sizeof(int) * (NI > 0 ? NI : 1));
    for (i = NI; i < NN; ++i) {
        uint32_t p = nodes[4 * i];
        while (p != INVALID) {
            uint32_t l, r;
            if (flags[p] == 0) { flags[p] = 1; break; }
            l = nodes[4 * p + 1];
            r = nodes[4 * p + 2];
            cn[p] = cone_merge2d(cn[l], cn[r], box_centroid(bx[l]), box_centroid(bx[r]), box_centroid(bx[p]));
            p = nodes[4 * p];
        }
    }
    for (i = 0; i < NN; ++i) {
        boxes[6 * i + 0] = bx[i].lo.x; boxes[6 * i + 1] = bx[i].lo.y; boxes[6 * i + 2] = bx[i].lo.z;
        boxes[6 * i + 3] = bx[i].hi.x; boxes[6 * i + 4] = bx[i].hi.y; boxes[6 * i + 5] = bx[i].hi.z;
        cones[5 * i + 0] = cn[i].axis.x; cones[5 * i + 1] = cn[i].axis.y; cones[5 * i + 2] = cn[i].axis.z;
        cones[5 * i + 3] = cn[i].half_angle; cones[5 * i + 4] = cn[i].radius;
    }
    free(flags); free(bx); free(cn); free(ci); free(codes); free(codes64);
    return collision;
}

/* ---- 2D primitive tests ---- */
static float closest_point_segment2d(f2 pa, f2 pb, f2 x) /* scene.cuh:176-201 */
{
    f2 u = mk2(pb.x - pa.x, pb.y - pa.y), v = mk2(x.x - pa.x, x.y - pa.y), pt;
    float c1 = dot2(u, v), c2, t;
    if (c1 <= 0.0f) return len2(mk2(x.x - pa.x, x.y - pa.y));
    c2 = dot2(u, u);
    if (c2 <= c1) return len2(mk2(x.x - pb.x, x.y - pb.y));
    t = c1 / c2;
    pt = mk2(pa.x + u.x * t, pa.y + u.y * t);
    return len2(mk2(x.x - pt.x, x.y - pt.y));
}

static int is_silhouette_vertex(f2 n0, f2 n1, f2 view, float d, int flip) /* scene.cuh:112-141 */
{
    const float precision = 1e-3f;
    float sign = flip ? 1.0f : -1.0f, dot0, dot1;
    f2 vu;
    if (d <= precision) {
        float det = n0.x * n1.y - n0.y * n1.x;
        return sign * det > precision;
    }
    vu = mk2(view.x / d, view.y / d);
    dot0 = dot2(vu, n0);
    dot1 = dot2(vu, n1);
    if (fabsf(dot0) <= precision) return sign * dot1 > precision;
    if (fabsf(dot1) <= precision) return sign * dot0 > precision;
    return dot0 * dot1 < 0.0f;
}

static int vertex_closest_silhouette(const float *V, const int32_t *s, f2 o, float max_r2, float *dist, int flip)
{ /* scene.cuh:354-386 */
    f2 p, view;
    float d;
    int sil;
    if (0.0f >= max_r2) return 0;
    p = vget2(V, s[1]);
    view = mk2(o.x - p.x, o.y - p.y);
    d = len2(view);
    if (d * d > max_r2) return 0;
    sil = !sv_has_face(s, 0) || !sv_has_face(s, 1);
    if (!sil) sil = is_silhouette_vertex(sv_face_normal(V, s, 0, 1), sv_face_normal(V, s, 1, 1), view, d, flip);
    if (sil && d * d <= max_r2) {
        *dist = d;
        return 1;
    }
    return 0;
}

static int ray_segment(f2 o, f2 dir, f2 p0, f2 p1, float *t_out, float *s_out) /* scene.cuh:545-576 */
{
    f2 sd = mk2(p1.x - p0.x, p1.y - p0.y);
    float D = dir.x * (-sd.y) + dir.y * sd.x, inv, t, s;
    if (fabsf(D) < F_EPS) return 0;
    inv = 1.0f / D;
    t = ((p0.x - o.x) * (-sd.y) - (p0.y - o.y) * (-sd.x)) * inv;
    s = (dir.x * (p0.y - o.y) - dir.y * (p0.x - o.x)) * inv;
    if (s >= -1e-3f && s <= 1.0f + 1e-3f && t >= 0.0f) {
        *t_out = t;
        *s_out = s;
        return 1;
    }
    return 0;
}

static int sphere_segment(f2 p1, f2 p2, f2 c, float radius) /* scene.cuh:581-602 */
{
    f2 d = mk2(p2.x - p1.x, p2.y - p1.y), cp;
    float len_sq = d.x * d.x + d.y * d.y;
    float t = ((c.x - p1.x) * d.x + (c.y - p1.y) * d.y) / len_sq, dx, dy;
    t = smax(0.0f, smin(1.0f, t));
    cp = mk2(p1.x + t * d.x, p1.y + t * d.y);
    dx = cp.x - c.x;
    dy = cp.y - c.y;
    return dx * dx + dy * dy <= (radius * radius);
}

static float green_weight2d(f2 x, f2 y) /* scene.cuh:605-612 */
{
    float r = fmaxf(len2(mk2(x.x - y.x, x.y - y.y)), 1e-2f);
    return fabsf(odm_logf(r) / (F_PI * 2.0f));
}

/* ---- 2D queries in the reference's visit order ---- */
typedef struct {
    uint32_t n_obj;
    const uint32_t *nodes;
    const float *boxes;
    const float *cones;
    const float *V; /* 2 floats per vertex */
    const int32_t *S;
    const int32_t *slots;
    const int32_t *owned;
} orc_tree2d;

static inline box3 tbox2(const orc_tree2d *t, uint32_t i)
{
    box3 b;
    const float *p = t->boxes + 6 * (size_t)i;
    b.lo = mk(p[0], p[1], p[2]);
    b.hi = mk(p[3], p[4], p[5]);
    return b;
}
static inline cone3 tcone2(const orc_tree2d *t, uint32_t i)
{
    cone3 c;
    const float *p = t->cones + 5 * (size_t)i;
    c.axis = mk(p[0], p[1], p[2]);
    c.half_angle = p[3];
    c.radius = p[4];
    return c;
}
static float seg_dist(const orc_tree2d *t, uint32_t obj, f2 q)
{
    return closest_point_segment2d(vget2(t->V, t->S[2 * obj]), vget2(t->V, t->S[2 * obj + 1]), q);
}
/* box mindist / minmaxdist in 2D (aabb.cuh:130-135, 160-180) */
static inline float mindist2(box3 b, f2 p)
{
    float dx = fminf(b.hi.x, fmaxf(b.lo.x, p.x)) - p.x;
    float dy = fminf(b.hi.y, fmaxf(b.lo.y, p.y)) - p.y;
    return dx * dx + dy * dy;
}
static inline float minmaxdist2(box3 b, f2 p)
{
    float lx = b.lo.x - p.x, ly = b.lo.y - p.y, ux = b.hi.x - p.x, uy = b.hi.y - p.y;
    float rmx = lx * lx, rmy = ly * ly, rMx = ux * ux, rMy = uy * uy, t;
    if ((b.hi.x + b.lo.x) * 0.5f < p.x) { t = rmx; rmx = rMx; rMx = t; }
    if ((b.hi.y + b.lo.y) * 0.5f < p.y) { t = rmy; rmy = rMy; rMy = t; }
    return fminf(rmx + rMy, rMx + rmy);
}

void orc_closest_point2d(const orc_tree2d *t, const float *q, int64_t n, uint32_t *idx_out, float *dist_out)
{
    int64_t qi;
#pragma omp parallel for schedule(dynamic, 256)
    for (qi = 0; qi < n; ++qi) {
        f2 p = mk2(q[2 * qi], q[2 * qi + 1]);
        uint32_t sn[128];
        float sd[128];
        int sp = 0;
        uint32_t nearest = INVALID;
        float best = INFINITY;
        if (t->n_obj == 1) {
            float d = seg_dist(t, t->nodes[3], p);
            d *= d;
            idx_out[qi] = t->nodes[3];
            dist_out[qi] = sqrtf(d);
            continue;
        }
        sn[sp] = 0; sd[sp] = mindist2(tbox2(t, 0), p); ++sp;
        do {
            uint32_t node, L, R, o;
            float nd, Lmd, Rmd, Lmm, Rmm;
            --sp;
            node = sn[sp];
            nd = sd[sp];
            if (nd > best) continue;
            L = t->nodes[4 * node + 1];
            R = t->nodes[4 * node + 2];
            Lmd = mindist2(tbox2(t, L), p);
            Rmd = mindist2(tbox2(t, R), p);
            Lmm = minmaxdist2(tbox2(t, L), p);
            Rmm = minmaxdist2(tbox2(t, R), p);
            if (Lmd <= Rmm) {
                o = t->nodes[4 * L + 3];
                if (o != INVALID) {
                    float d = seg_dist(t, o, p);
                    d *= d;
                    if (d <= best) { best = d; nearest = o; }
                } else { sn[sp] = L; sd[sp] = Lmd; ++sp; }
            }
            if (Rmd <= Lmm) {
                o = t->nodes[4 * R + 3];
                if (o != INVALID) {
                    float d = seg_dist(t, o, p);
                    d *= d;
                    if (d <= best) { best = d; nearest = o; }
                } else { sn[sp] = R; sd[sp] = Rmd; ++sp; }
            }
        } while (sp > 0);
        idx_out[qi] = nearest;
        dist_out[qi] = sqrtf(best);
    }
}

static int sil_leaf2d(const orc_tree2d *t, uint32_t obj, f2 q, float max_r2, float *dist, uint32_t *sil, int flip)
{ /* scene.cuh:517-541 */
    int k, ret = 0;
    float mr2 = max_r2;
    for (k = 0; k < 2; ++k) {
        int32_t v = t->owned[2 * (size_t)obj + k];
        if (v != -1 && vertex_closest_silhouette(t->V, t->slots + 4 * (size_t)v, q, mr2, dist, flip)) {
            ret = 1;
            mr2 = (*dist) * (*dist);
            *sil = (uint32_t)v;
        }
    }
    return ret;
}

void orc_closest_silhouette2d(const orc_tree2d *t, const float *q, int64_t n, int flip, uint32_t *sil_out,
                              float *dist_out)
{
    int64_t qi;
#pragma omp parallel for schedule(dynamic, 256)
    for (qi = 0; qi < n; ++qi) {
        f2 p = mk2(q[2 * qi], q[2 * qi + 1]);
        uint32_t sn[128], best_s = INVALID;
        float sd[128], best = INFINITY;
        int sp = 0;
        if (t->n_obj == 1) {
            float d = INFINITY;
            uint32_t s = INVALID;
            if (cone_valid(tcone2(t, 0)) && cone_overlap2d(tcone2(t, 0), p, tbox2(t, 0), mindist2(tbox2(t, 0), p)) &&
                sil_leaf2d(t, t->nodes[3], p, best * best, &d, &s, flip) && d <= best) {
                best = d;
                best_s = s;
            }
            sil_out[qi] = best_s;
            dist_out[qi] = best;
            continue;
        }
        sn[sp] = 0; sd[sp] = mindist2(tbox2(t, 0), p); ++sp;
        do {
            uint32_t node, c[2];
            float nd, md[2];
            int hit[2], s;
            box3 b[2];
            --sp;
            node = sn[sp];
            nd = sd[sp];
            if (nd > best * best) continue;
            c[0] = t->nodes[4 * node + 1];
            c[1] = t->nodes[4 * node + 2];
            for (s = 0; s < 2; ++s) {
                b[s] = tbox2(t, c[s]);
                md[s] = mindist2(b[s], p);
                hit[s] = cone_valid(tcone2(t, c[s])) && cone_overlap2d(tcone2(t, c[s]), p, b[s], md[s]);
            }
            for (s = 0; s < 2; ++s) {
                uint32_t o;
                if (!hit[s]) continue;
                o = t->nodes[4 * c[s] + 3];
                if (o != INVALID) {
                    float d = INFINITY;
                    uint32_t sv = INVALID;
                    if (sil_leaf2d(t, o, p, best * best, &d, &sv, flip) && d <= best) {
                        best = d;
                        best_s = sv;
                    }
                } else { sn[sp] = c[s]; sd[sp] = md[s]; ++sp; }
            }
        } while (sp > 0);
        sil_out[qi] = best_s;
        dist_out[qi] = best;
    }
}

static inline int box_ray2(box3 b, f2 o, f2 inv, float max_dist, float *dist) /* aabb.cuh:366-395 */
{
    float t1 = (b.lo.x - o.x) * inv.x, t2 = (b.hi.x - o.x) * inv.x;
    float tmin = fminf(t1, t2), tmax = fmaxf(t1, t2);
    t1 = (b.lo.y - o.y) * inv.y;
    t2 = (b.hi.y - o.y) * inv.y;
    tmin = fmaxf(tmin, fminf(t1, t2));
    tmax = fminf(tmax, fmaxf(t1, t2));
    if (tmax >= tmin && tmax >= 0.0f && tmin <= max_dist) {
        *dist = tmin >= 0.0f ? tmin : 0.0f;
        return 1;
    }
    return 0;
}

void orc_ray2d(const orc_tree2d *t, const float *o, const float *d, const float *tmax, int64_t n, uint32_t *idx_out,
               float *t_out, float *s_out, uint8_t *hit_out)
{
    int64_t qi;
#pragma omp parallel for schedule(dynamic, 256)
    for (qi = 0; qi < n; ++qi) {
        f2 ro = mk2(o[2 * qi], o[2 * qi + 1]), rd = mk2(d[2 * qi], d[2 * qi + 1]);
        f2 inv = mk2(1.0f / rd.x, 1.0f / rd.y);
        float max_dist = tmax[qi], min_dist = INFINITY, ss = 0.0f;
        uint32_t sn[128], nearest = INVALID;
        float sd[128];
        int sp = 0, found = 0;
        sn[sp] = 0; sd[sp] = INFINITY; ++sp;
        do {
            uint32_t node, obj;
            float nt;
            --sp;
            node = sn[sp];
            nt = sd[sp];
            if (nt > min_dist) continue;
            obj = t->nodes[4 * node + 3];
            if (obj != INVALID) {
                float th, sh;
                if (ray_segment(ro, rd, vget2(t->V, t->S[2 * obj]), vget2(t->V, t->S[2 * obj + 1]), &th, &sh) &&
                    th < max_dist && th < min_dist) {
                    min_dist = th;
                    found = 1;
                    ss = sh;
                    nearest = obj;
                }
            } else {
                uint32_t L = t->nodes[4 * node + 1], R = t->nodes[4 * node + 2], cl = L, ot = R, tmp;
                float Ld, Rd, tf;
                int Lh = box_ray2(tbox2(t, L), ro, inv, max_dist, &Ld);
                int Rh = box_ray2(tbox2(t, R), ro, inv, max_dist, &Rd);
                if (Lh && Rh) {
                    if (Rd < Ld) { tf = Ld; Ld = Rd; Rd = tf; tmp = cl; cl = ot; ot = tmp; }
                    sn[sp] = ot; sd[sp] = Rd; ++sp;
                    sn[sp] = cl; sd[sp] = Ld; ++sp;
                } else if (Lh) { sn[sp] = L; sd[sp] = Ld; ++sp; }
                else if (Rh) { sn[sp] = R; sd[sp] = Rd; ++sp; }
            }
        } while (sp > 0);
        hit_out[qi] = (uint8_t)found;
        idx_out[qi] = nearest;
        t_out[qi] = min_dist;
        s_out[qi] = ss;
    }
}

void orc_sample2d(const orc_tree2d *t, const float *c, const float *r, const float *u, int64_t n, int32_t *idx_out,
                  float *pdf_out, float *pt_out)
{
    int64_t qi;
#pragma omp parallel for schedule(static)
    for (qi = 0; qi < n; ++qi) {
        f2 cc = mk2(c[2 * qi], c[2 * qi + 1]);
        float rr = r[qi], uu = u[3 * qi], npdf = 1.0f, pdf = 0.0f;
        uint32_t node = 0;
        int32_t obj_index = -1;
        for (;;) {
            uint32_t obj = t->nodes[4 * node + 3];
            if (obj != INVALID) {
                f2 p0 = vget2(t->V, t->S[2 * obj]), p1 = vget2(t->V, t->S[2 * obj + 1]);
                if (sphere_segment(p0, p1, cc, rr)) {
                    obj_index = (int32_t)obj;
                    pdf = npdf;
                    pdf /= len2(mk2(p0.x - p1.x, p0.y - p1.y)); /* scene.cuh:409-418 */
                }
                break;
            } else {
                uint32_t L = t->nodes[4 * node + 1], R = t->nodes[4 * node + 2];
                box3 Lb = tbox2(t, L), Rb = tbox2(t, R);
                f3 cl3 = box_centroid(Lb), cr3 = box_centroid(Rb);
                float Lw = box_sphere(Lb, mk(cc.x, cc.y, 0.0f), rr) ? green_weight2d(cc, mk2(cl3.x, cl3.y)) : 0.0f;
                float Rw = box_sphere(Rb, mk(cc.x, cc.y, 0.0f), rr) ? green_weight2d(cc, mk2(cr3.x, cr3.y)) : 0.0f;
                float tw = Lw + Rw;
                if (tw > 0.0f) {
                    float Lp = Lw / tw;
                    if (uu < Lp) {
                        uu /= Lp;
                        npdf = Lp * npdf;
                        node = L;
                    } else {
                        float Rp = 1.0f - Lp;
                        uu = (uu - Lp) / Rp;
                        npdf = Rp * npdf;
                        node = R;
                    }
                } else break;
            }
        }
        idx_out[qi] = obj_index;
        pdf_out[qi] = obj_index >= 0 ? pdf : 0.0f;
        if (pt_out) {
            float px = 0.0f, py = 0.0f;
            if (obj_index >= 0) { /* sample_line, scene.cuh:29-32 */
                f2 pa = vget2(t->V, t->S[2 * obj_index]), pb = vget2(t->V, t->S[2 * obj_index + 1]);
                float uv = u[3 * qi + 1];
                px = pa.x + uv * (pb.x - pa.x);
                py = pa.y + uv * (pb.y - pa.y);
            }
            pt_out[2 * qi] = px;
            pt_out[2 * qi + 1] = py;
        }
    }
}

/* FP64 brute force, 2D */
void orc_bf_closest_point2d(const float *V, const int32_t *S, int32_t nS, const float *q, int64_t n, uint32_t *idx,
                            double *dist)
{
    int64_t qi;
#pragma omp parallel for schedule(dynamic, 64)
    for (qi = 0; qi < n; ++qi) {
        double px = q[2 * qi], py = q[2 * qi + 1], best = INFINITY;
        uint32_t bi = INVALID;
        int32_t s;
        for (s = 0; s < nS; ++s) {
            double ax = V[2 * S[2 * s]], ay = V[2 * S[2 * s] + 1], bx = V[2 * S[2 * s + 1]], by = V[2 * S[2 * s + 1] + 1];
            double ux = bx - ax, uy = by - ay, c1 = ux * (px - ax) + uy * (py - ay), c2 = ux * ux + uy * uy, t, dx, dy, d;
            t = c1 <= 0 ? 0.0 : (c2 <= c1 ? 1.0 : c1 / c2);
            dx = px - (ax + t * ux);
            dy = py - (ay + t * uy);
            d = sqrt(dx * dx + dy * dy);
            if (d < best) { best = d; bi = (uint32_t)s; }
        }
        idx[qi] = bi;
        dist[qi] = best;
    }
}

void orc_bf_closest_silhouette2d(const float *V, const int32_t *slots, int32_t nV, const float *q, int64_t n, int flip,
                                 uint32_t *sil, double *dist)
{
    int64_t qi;
#pragma omp parallel for schedule(dynamic, 64)
    for (qi = 0; qi < n; ++qi) {
        double px = q[2 * qi], py = q[2 * qi + 1], best = INFINITY, sign = flip ? 1.0 : -1.0;
        uint32_t bs = INVALID;
        int32_t v;
        for (v = 0; v < nV; ++v) {
            const int32_t *s = slots + 4 * v;
            double vx, vy, d, n0x, n0y, n1x, n1y, l0, l1;
            int is;
            if (s[1] == -1) continue;
            vx = px - V[2 * s[1]];
            vy = py - V[2 * s[1] + 1];
            d = sqrt(vx * vx + vy * vy);
            if (d >= best) continue;
            if (s[0] == -1 || s[2] == -1) is = 1;
            else {
                n0x = (double)V[2 * s[2] + 1] - V[2 * s[1] + 1];
                n0y = -((double)V[2 * s[2]] - V[2 * s[1]]);
                n1x = (double)V[2 * s[1] + 1] - V[2 * s[0] + 1];
                n1y = -((double)V[2 * s[1]] - V[2 * s[0]]);
                l0 = sqrt(n0x * n0x + n0y * n0y);
                l1 = sqrt(n1x * n1x + n1y * n1y);
                n0x /= l0; n0y /= l0; n1x /= l1; n1y /= l1;
                if (d <= 1e-3) is = sign * (n0x * n1y - n0y * n1x) > 1e-3;
                else {
                    double d0 = (vx * n0x + vy * n0y) / d, d1 = (vx * n1x + vy * n1y) / d;
                    if (fabs(d0) <= 1e-3) is = sign * d1 > 1e-3;
                    else if (fabs(d1) <= 1e-3) is = sign * d0 > 1e-3;
                    else is = d0 * d1 < 0;
                }
            }
            if (is) { best = d; bs = (uint32_t)v; }
        }
        sil[qi] = bs;
        dist[qi] = best;
    }
}

/* ------------------------------------------------------------------ */
/* Generic overlap and line queries (query.cuh:12-77, :176-231): every  */
/* primitive whose leaf box intersects the query box (aabb.cuh:54-82),   */
/* or whose leaf box the infinite line crosses (aabb.cuh:330-364) and    */
/* which passes the element test, in the reference's stack order: pop a  */
/* node, test its left child then its right child, emit a leaf at once,  */
/* push an internal child.  At most max_out ids per query are written;   */
/* the count returned is the total found.  dim 2 uses the x, y slabs.    */
/* The line's element test is scene.cuh:1005-1052 (3D) / :545-576 (2D)   */
/* without the t >= 0 condition of a ray (a line extends both ways).     */
/* N = 1 (the reference reads nodes[0xFFFFFFFF]): the single leaf's box  */
/* is tested directly.                                                   */
/* ------------------------------------------------------------------ */
static int box_box(const float *a, const float *b, int dim) /* aabb.cuh:54-82 */
{
    int k;
    for (k = 0; k < dim; ++k)
        if (a[3 + k] < b[k] || b[3 + k] < a[k]) return 0;
    return 1;
}

static int line_box(const float *b, const float *o, const float *inv, int dim) /* aabb.cuh:330-364 */
{
    float t1 = (b[0] - o[0]) * inv[0], t2 = (b[3] - o[0]) * inv[0];
    float tmin = fminf(t1, t2), tmax = fmaxf(t1, t2);
    int k;
    for (k = 1; k < dim; ++k) {
        t1 = (b[k] - o[k]) * inv[k];
        t2 = (b[3 + k] - o[k]) * inv[k];
        tmin = fmaxf(tmin, fminf(t1, t2));
        tmax = fminf(tmax, fmaxf(t1, t2));
    }
    return tmax >= tmin;
}

/* the element tests of the line query; 1 and *t_out on a hit */
static int line_triangle(f3 o, f3 dir, f3 v0, f3 v1, f3 v2, float *t_out)
{
    f3 e1 = mk(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z);
    f3 e2 = mk(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z);
    f3 h = mk(dir.y * e2.z - dir.z * e2.y, dir.z * e2.x - dir.x * e2.z, dir.x * e2.y - dir.y * e2.x);
    float det = e1.x * h.x + e1.y * h.y + e1.z * h.z, inv, u, v;
    f3 s, q;
    if (fabsf(det) < F_EPS) return 0;
    inv = 1.0f / det;
    s = mk(o.x - v0.x, o.y - v0.y, o.z - v0.z);
    u = (s.x * h.x + s.y * h.y + s.z * h.z) * inv;
    if (u < 0.0f || u > 1.0f) return 0;
    q = mk(s.y * e1.z - s.z * e1.y, s.z * e1.x - s.x * e1.z, s.x * e1.y - s.y * e1.x);
    v = (dir.x * q.x + dir.y * q.y + dir.z * q.z) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0;
    *t_out = (e2.x * q.x + e2.y * q.y + e2.z * q.z) * inv;
    return 1;
}

static int line_segment(f2 o, f2 dir, f2 p0, f2 p1, float *t_out)
{
    f2 sd = mk2(p1.x - p0.x, p1.y - p0.y);
    float D = dir.x * (-sd.y) + dir.y * sd.x, inv, t, s;
    if (fabsf(D) < F_EPS) return 0;
    inv = 1.0f / D;
    t = ((p0.x - o.x) * (-sd.y) - (p0.y - o.y) * (-sd.x)) * inv;
    s = (dir.x * (p0.y - o.y) - dir.y * (p0.x - o.x)) * inv;
    if (s >= -1e-3f && s <= 1.0f + 1e-3f) {
        *t_out = t;
        return 1;
    }
    return 0;
}

typedef struct { /* the common prefix of orc_tree and orc_tree2d */
    uint32_t n_obj;
    const uint32_t *nodes;
    const float *boxes;
    const float *cones;
    const float *V;
    const int32_t *F;
} orc_tree_any;

/* kind 0: overlap with the box q (6 floats: lo xyz, hi xyz); kind 1: the
 * line through q (origin xyz, direction xyz) */
static uint32_t generic_one(const orc_tree_any *t, int dim, int kind, const float *q, uint32_t max_out,
                            uint32_t *out_idx, float *out_t)
{
    uint32_t stack[128], found = 0;
    int sp = 0, c;
    float inv[3] = {0.0f, 0.0f, 0.0f};
    if (kind == 1) { /* line<T, dim> constructor: dir_inv = 1 / dir (aabb.cuh:288-294) */
        inv[0] = 1.0f / q[3];
        inv[1] = 1.0f / q[4];
        if (dim == 3) inv[2] = 1.0f / q[5];
    }
#define GEN_HIT_BOX(b) (kind == 0 ? box_box((b), q, dim) : line_box((b), q, inv, dim))
#define GEN_EMIT(obj)                                                                            \
    do {                                                                                         \
        float tt_ = 0.0f;                                                                        \
        int ok_ = 1;                                                                             \
        if (kind == 1) {                                                                         \
            if (dim == 3) {                                                                      \
                const int32_t *f_ = t->F + 3 * (size_t)(obj);                                    \
                ok_ = line_triangle(mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), vget(t->V, f_[0]), \
                                    vget(t->V, f_[1]), vget(t->V, f_[2]), &tt_);                 \
            } else {                                                                             \
                const int32_t *s_ = t->F + 2 * (size_t)(obj);                                    \
                ok_ = line_segment(mk2(q[0], q[1]), mk2(q[3], q[4]), vget2(t->V, s_[0]),         \
                                   vget2(t->V, s_[1]), &tt_);                                    \
            }                                                                                    \
        }                                                                                        \
        if (ok_) {                                                                               \
            if (found < max_out) {                                                               \
                out_idx[found] = (obj);                                                          \
                if (out_t) out_t[found] = tt_;                                                   \
            }                                                                                    \
            ++found;                                                                             \
        }                                                                                        \
    } while (0)
    if (t->n_obj == 1) {
        if (GEN_HIT_BOX(t->boxes)) GEN_EMIT(T_OBJ(t, 0));
        return found;
    }
    stack[sp++] = 0;
    do {
        uint32_t node = stack[--sp];
        uint32_t ch[2];
        ch[0] = T_LEFT(t, node);
        ch[1] = T_RIGHT(t, node);
        for (c = 0; c < 2; ++c) {
            if (GEN_HIT_BOX(t->boxes + 6 * (size_t)ch[c])) {
                uint32_t obj = T_OBJ(t, ch[c]);
                if (obj != INVALID) GEN_EMIT(obj);
                else stack[sp++] = ch[c];
            }
        }
    } while (sp > 0);
#undef GEN_HIT_BOX
#undef GEN_EMIT
    return found;
}

void orc_generic_query(const void *tree, int dim, int kind, const float *q, int64_t n, uint32_t max_out,
                       uint32_t *counts, uint32_t *out_idx, float *out_t, int nthreads)
{
    int64_t i;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (i = 0; i < n; ++i)
        counts[i] = generic_one((const orc_tree_any *)tree, dim, kind, q + 6 * i, max_out,
                                out_idx + (size_t)i * max_out, out_t ? out_t + (size_t)i * max_out : NULL);
}
