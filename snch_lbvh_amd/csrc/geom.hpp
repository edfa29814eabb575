// Geometry shared by the build and the query kernels.  Every function keeps
// the floating-point expression order of the reference code it cites, and the
// library is compiled with -ffp-contract=off, so results are bit-identical to
// any IEEE single-precision evaluation of the same expressions.
#pragma once
#include <cstdint>

#include "detmath.hpp"

namespace snch {

constexpr float F_EPS = 1.1920928955078125e-07f;  // FLT_EPSILON (utility.cuh:186)
constexpr float F_PI = 3.14159274101257324219f;   // (float)M_PI
constexpr float F_PI_2 = 1.57079637050628662109f; // (float)M_PI_2
constexpr float F_INF = __builtin_inff();
constexpr uint32_t INVALID = 0xFFFFFFFFu;

struct v3 {
    float x, y, z;
};
struct box3 {
    v3 lo, hi;
};
struct cone3 {
    v3 axis;
    float half_angle;
    float radius;
};

SNCH_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
SNCH_HD v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
SNCH_HD float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SNCH_HD float len3(v3 a) { return dm::sqrt_(a.x * a.x + a.y * a.y + a.z * a.z); }
SNCH_HD float sqlen3(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
SNCH_HD v3 cross3(v3 u, v3 v) { return mk(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x); }
SNCH_HD v3 normalize3(v3 v)
{
    float n = len3(v);
    return mk(v.x / n, v.y / n, v.z / n);
}
SNCH_HD v3 cabs3(v3 a) { return mk(dm::fabs_(a.x), dm::fabs_(a.y), dm::fabs_(a.z)); }
// std::min / std::max semantics of the reference: min(a,b) = (b < a) ? b : a
SNCH_HD float smin(float a, float b) { return (b < a) ? b : a; }
SNCH_HD float smax(float a, float b) { return (a < b) ? b : a; }
SNCH_HD float clamp11(float d) { return smax(-1.0f, smin(1.0f, d)); }
// ::fminf / ::fmaxf (IEEE minNum / maxNum)
SNCH_HD float fmn(float a, float b) { return __builtin_fminf(a, b); }
SNCH_HD float fmx(float a, float b) { return __builtin_fmaxf(a, b); }

// ---------------------------------------------------------------- AABB
// aabb(p) and expand_to_include (aabb.cuh:30-38, 93-101)
SNCH_HD box3 box_of_point(v3 p)
{
    box3 b;
    b.hi = mk(p.x + F_EPS, p.y + F_EPS, p.z + F_EPS);
    b.lo = mk(p.x - F_EPS, p.y - F_EPS, p.z - F_EPS);
    return b;
}
SNCH_HD void box_expand(box3& b, v3 p)
{
    v3 pl = mk(p.x - F_EPS, p.y - F_EPS, p.z - F_EPS);
    v3 pu = mk(p.x + F_EPS, p.y + F_EPS, p.z + F_EPS);
    b.lo = mk(fmn(b.lo.x, pl.x), fmn(b.lo.y, pl.y), fmn(b.lo.z, pl.z));
    b.hi = mk(fmx(b.hi.x, pu.x), fmx(b.hi.y, pu.y), fmx(b.hi.z, pu.z));
}
// merge (aabb.cuh:113-124)
SNCH_HD box3 box_merge(const box3& a, const box3& b)
{
    box3 m;
    m.hi = mk(fmx(a.hi.x, b.hi.x), fmx(a.hi.y, b.hi.y), fmx(a.hi.z, b.hi.z));
    m.lo = mk(fmn(a.lo.x, b.lo.x), fmn(a.lo.y, b.lo.y), fmn(a.lo.z, b.lo.z));
    return m;
}
// centroid (aabb.cuh:271-279)
SNCH_HD v3 box_centroid(const box3& b)
{
    return mk((b.hi.x + b.lo.x) * 0.5f, (b.hi.y + b.lo.y) * 0.5f, (b.hi.z + b.lo.z) * 0.5f);
}
// mindist, squared (aabb.cuh:144-150)
SNCH_HD float box_mindist(const box3& b, v3 p)
{
    float dx = fmn(b.hi.x, fmx(b.lo.x, p.x)) - p.x;
    float dy = fmn(b.hi.y, fmx(b.lo.y, p.y)) - p.y;
    float dz = fmn(b.hi.z, fmx(b.lo.z, p.z)) - p.z;
    return dx * dx + dy * dy + dz * dz;
}
// minmaxdist (aabb.cuh:182-209)
SNCH_HD float box_minmaxdist(const box3& b, v3 p)
{
    float lx = b.lo.x - p.x, ly = b.lo.y - p.y, lz = b.lo.z - p.z;
    float ux = b.hi.x - p.x, uy = b.hi.y - p.y, uz = b.hi.z - p.z;
    float rmx = lx * lx, rmy = ly * ly, rmz = lz * lz;
    float rMx = ux * ux, rMy = uy * uy, rMz = uz * uz;
    if ((b.hi.x + b.lo.x) * 0.5f < p.x) { float t = rmx; rmx = rMx; rMx = t; }
    if ((b.hi.y + b.lo.y) * 0.5f < p.y) { float t = rmy; rmy = rMy; rMy = t; }
    if ((b.hi.z + b.lo.z) * 0.5f < p.z) { float t = rmz; rmz = rMz; rMz = t; }
    float dx = rmx + rMy + rMz;
    float dy = rMx + rmy + rMz;
    float dz = rMx + rMy + rmz;
    return fmn(dx, fmn(dy, dz));
}
// ray slab test (aabb.cuh:397-431)
SNCH_HD bool box_ray(const box3& b, v3 o, v3 inv, float max_dist, float& dist)
{
    float t1 = (b.lo.x - o.x) * inv.x;
    float t2 = (b.hi.x - o.x) * inv.x;
    float tmin = fmn(t1, t2);
    float tmax = fmx(t1, t2);
    t1 = (b.lo.y - o.y) * inv.y;
    t2 = (b.hi.y - o.y) * inv.y;
    tmin = fmx(tmin, fmn(t1, t2));
    tmax = fmn(tmax, fmx(t1, t2));
    t1 = (b.lo.z - o.z) * inv.z;
    t2 = (b.hi.z - o.z) * inv.z;
    tmin = fmx(tmin, fmn(t1, t2));
    tmax = fmn(tmax, fmx(t1, t2));
    if (tmax >= tmin && tmax >= 0.0f && tmin <= max_dist) {
        dist = tmin >= 0.0f ? tmin : 0.0f;
        return true;
    }
    return false;
}
// sphere-box test (aabb.cuh:433-449)
SNCH_HD bool box_sphere(const box3& b, v3 c, float r)
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

// ---------------------------------------------------------------- cones
SNCH_HD bool cone_valid(float half_angle) { return half_angle >= 0.0f; } // cone.cuh:18-22

// cone.cuh:34-42
SNCH_HD void ortho_basis(v3 n, v3& b1, v3& b2)
{
    float sign = dm::signbit_(n.z) ? -1.0f : 1.0f; // copysignf(1, n.z)
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    b1 = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    b2 = mk(b, sign + n.y * n.y * a, -n.y);
}
// cone.cuh:58-66
SNCH_HD float project_to_plane(v3 n, v3 e)
{
    v3 b1, b2;
    ortho_basis(n, b1, b2);
    float r1 = dot3(e, cabs3(b1));
    float r2 = dot3(e, cabs3(b2));
    return dm::sqrt_(r1 * r1 + r2 * r2);
}
SNCH_HD bool inrange(float v, float lo, float hi) { return v >= lo && v <= hi; }

// cone-view overlap, the SNCH pruning test (cone.cuh:168-212)
SNCH_HD bool cone_overlap(v3 axis, float half_angle, float radius, v3 o, const box3& b, float dist_to_box)
{
    if (half_angle >= F_PI_2 || dist_to_box < F_EPS)
        return true;
    v3 c = box_centroid(b);
    v3 v = mk(c.x - o.x, c.y - o.y, c.z - o.z);
    float l = len3(v);
    v.x /= l;
    v.y /= l;
    v.z /= l;
    float d_axis = dm::acosf_(clamp11(dot3(axis, v)));
    if (inrange(F_PI_2, d_axis - half_angle, d_axis + half_angle))
        return true;
    if (l > radius) {
        float vha = dm::asinf_(radius / l);
        float sum = half_angle + vha;
        float mn = d_axis - sum;
        float mx = d_axis + sum;
        return sum >= F_PI_2 ? true : inrange(F_PI_2, mn, mx);
    }
    v3 e = mk(b.hi.x - c.x, b.hi.y - c.y, b.hi.z - c.z);
    float d = dot3(e, cabs3(v));
    float s = l - d;
    if (s <= 0.0f)
        return true;
    d = project_to_plane(v, e);
    float vha = dm::atan2f_(d, s);
    float sum = half_angle + vha;
    float mn = d_axis - sum;
    float mx = d_axis + sum;
    return sum >= F_PI_2 ? true : inrange(F_PI_2, mn, mx);
}

// The same test without inverse trigonometry, for the traversal.  With
// C = cos(angle(axis, view direction)) and vha the view cone's half angle
// (asin(r / l), or atan2(d, s) when the view point is inside the bounding
// sphere), the reference's condition |acos(C) - pi/2| <= ha + vha reads
// |C| <= sin(ha + vha), and ha + vha >= pi/2 reads cos(ha + vha) <= 0; both
// expand through the precomputed sin(ha), cos(ha) and the sine and cosine of
// vha, which are ratios of lengths.  A slack of 1e-5 (l + r) on the
// permissive side keeps every node the reference test keeps despite either
// test's rounding; a node kept in excess only costs a visit, never changes a
// result.  Conservative only with the enclosing cone radius.  Its square
// roots and reciprocals need not be correctly rounded (AM supplies them; on
// the GPU the hardware's 1-ulp v_sqrt_f32 / v_rcp_f32): errors of a few ulps
// are two orders of magnitude inside the slack.
struct HwMath {
    __device__ static float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};

template <class AM = HwMath>
SNCH_HD bool cone_overlap_fast(v3 axis, float half_angle, float sin_ha, float cos_ha, float radius, v3 o,
                               const box3& b, float dist_to_box)
{
    if (half_angle >= F_PI_2 || dist_to_box < F_EPS)
        return true;
    const v3 c = box_centroid(b);
    const v3 v = mk(c.x - o.x, c.y - o.y, c.z - o.z);
    const float l2 = dot3(v, v), l = AM::sqrt(l2);
    const float av = dm::fabs_(dot3(axis, v)); // l |C|
    const float tol = 1e-5f * (l + radius);
    if (l > radius) {
        const float w = AM::sqrt(fmx(l2 - radius * radius, 0.0f)); // l cos(vha); l sin(vha) = radius
        if (cos_ha * w <= sin_ha * radius + tol)
            return true; // ha + vha >= pi/2
        return av <= sin_ha * w + cos_ha * radius + tol;
    }
    if (!(l > 0.0f))
        return true;
    const float il = AM::rcp(l);
    const v3 vn = mk(v.x * il, v.y * il, v.z * il);
    const v3 e = mk(b.hi.x - c.x, b.hi.y - c.y, b.hi.z - c.z);
    const float s = l - dot3(e, cabs3(vn));
    if (s <= 0.0f)
        return true;
    // project_to_plane(vn, e) with the same basis (cone.cuh:24-42, :58-66)
    const float sign = dm::signbit_(vn.z) ? -1.0f : 1.0f;
    const float a = -AM::rcp(sign + vn.z);
    const float bb = vn.x * vn.y * a;
    const v3 b1 = mk(1.0f + sign * vn.x * vn.x * a, sign * bb, -sign * vn.x);
    const v3 b2 = mk(bb, sign + vn.y * vn.y * a, -vn.y);
    const float r1 = dot3(e, cabs3(b1)), r2 = dot3(e, cabs3(b2));
    const float d = AM::sqrt(r1 * r1 + r2 * r2); // tan(vha) = d / s
    const float h = AM::sqrt(d * d + s * s);
    if (cos_ha * s <= sin_ha * d + tol)
        return true;
    return av * il * h <= sin_ha * s + cos_ha * d + tol;
}

// Rodrigues rotation (cone.cuh:288-302).  When the axes are (anti)parallel the
// reference divides by a zero-length cross product; here the rotation axis
// falls back to a vector perpendicular to u.
SNCH_HD v3 cone_rotate(v3 u, v3 v, float theta)
{
    float st, ct;
    dm::sincosf_(theta, &st, &ct);
    v3 w = cross3(u, v);
    float n = len3(w);
    if (!(n > 0.0f)) {
        v3 b1, b2;
        ortho_basis(u, b1, b2);
        w = b1;
        n = len3(w);
    }
    w = mk(w.x / n, w.y / n, w.z / n);
    v3 om = mk((1.0f - ct) * w.x, (1.0f - ct) * w.y, (1.0f - ct) * w.z);
    float R00 = ct + om.x * w.x;
    float R01 = om.y * w.x - st * w.z;
    float R02 = om.z * w.x + st * w.y;
    float R10 = om.x * w.y + st * w.z;
    float R11 = ct + om.y * w.y;
    float R12 = om.z * w.y - st * w.x;
    float R20 = om.x * w.z - st * w.y;
    float R21 = om.y * w.z + st * w.x;
    float R22 = ct + om.z * w.z;
    return mk(R00 * u.x + R01 * u.y + R02 * u.z, R10 * u.x + R11 * u.y + R12 * u.z,
              R20 * u.x + R21 * u.y + R22 * u.z);
}

// Cone merge (cone.cuh:427-480).  Two fixes, both conservative: the
// o_theta >= pi branch leaves half_angle uninitialised in the reference and is
// pi here; with `conservative` the radius is the triangle-inequality bound
// max(ra + |da|, rb + |db|) instead of sqrt(max(ra^2 + |da|^2, rb^2 + |db|^2)).
SNCH_HD cone3 cone_merge(const cone3& a, const cone3& b, v3 oa, v3 ob, v3 on, bool conservative)
{
    cone3 r{mk(0.0f, 0.0f, 0.0f), 0.0f, 0.0f};
    bool va = cone_valid(a.half_angle), vb = cone_valid(b.half_angle);
    if (va && vb) {
        v3 axa = a.axis, axb = b.axis;
        float ha = a.half_angle, hb = b.half_angle;
        v3 da = mk(on.x - oa.x, on.y - oa.y, on.z - oa.z);
        v3 db = mk(on.x - ob.x, on.y - ob.y, on.z - ob.z);
        if (conservative)
            r.radius = smax(a.radius + len3(da), b.radius + len3(db));
        else
            r.radius = dm::sqrt_(smax(a.radius * a.radius + sqlen3(da), b.radius * b.radius + sqlen3(db)));
        if (hb > ha) {
            v3 t3 = axa;
            axa = axb;
            axb = t3;
            float t = ha;
            ha = hb;
            hb = t;
        }
        float theta = dm::acosf_(clamp11(dot3(axa, axb)));
        if (smin(theta + hb, F_PI) <= ha) {
            r.axis = axa;
            r.half_angle = ha;
            return r;
        }
        float o_theta = (ha + theta + hb) / 2.0f;
        if (o_theta >= F_PI) {
            r.axis = axa;
            r.half_angle = F_PI;
            return r;
        }
        r.axis = cone_rotate(axa, axb, o_theta - ha);
        r.half_angle = o_theta;
    } else if (va) {
        r = a;
    } else if (vb) {
        r = b;
    } else {
        r.half_angle = -F_PI;
    }
    return r;
}

// ---------------------------------------------------------------- primitives
// Ericson's closest point on a triangle (scene.cuh:34-110); returns the distance
SNCH_HD float closest_point_triangle(v3 pa, v3 pb, v3 pc, v3 x)
{
    v3 ab = sub(pb, pa), ac = sub(pc, pa), ax = sub(x, pa);
    float d1 = dot3(ab, ax), d2 = dot3(ac, ax);
    if (d1 <= 0.0f && d2 <= 0.0f)
        return len3(mk(x.x - pa.x, x.y - pa.y, x.z - pa.z));
    v3 bx = sub(x, pb);
    float d3 = dot3(ab, bx), d4 = dot3(ac, bx);
    if (d3 >= 0.0f && d4 <= d3)
        return len3(mk(x.x - pb.x, x.y - pb.y, x.z - pb.z));
    v3 cx = sub(x, pc);
    float d5 = dot3(ab, cx), d6 = dot3(ac, cx);
    if (d6 >= 0.0f && d5 <= d6)
        return len3(mk(x.x - pc.x, x.y - pc.y, x.z - pc.z));
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        v3 pt = mk(pa.x + ab.x * v, pa.y + ab.y * v, pa.z + ab.z * v);
        return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float w = d2 / (d2 - d6);
        v3 pt = mk(pa.x + ac.x * w, pa.y + ac.y * w, pa.z + ac.z * w);
        return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        v3 pt = mk(pb.x + (pc.x - pb.x) * w, pb.y + (pc.y - pb.y) * w, pb.z + (pc.z - pb.z) * w);
        return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
    }
    float denom = 1.0f / (va + vb + vc);
    float v = vb * denom;
    float w = vc * denom;
    v3 pt = mk(pa.x + ab.x * v + ac.x * w, pa.y + ab.y * v + ac.y * w, pa.z + ab.z * v + ac.z * w);
    return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
}

// closest point on a segment (scene.cuh:230-255)
SNCH_HD float closest_point_segment(v3 pa, v3 pb, v3 x, v3& pt)
{
    v3 u = sub(pb, pa), v = sub(x, pa);
    float c1 = dot3(u, v);
    if (c1 <= 0.0f) {
        pt = pa;
        return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
    }
    float c2 = dot3(u, u);
    if (c2 <= c1) {
        pt = pb;
        return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
    }
    float t = c1 / c2;
    pt = mk(pa.x + u.x * t, pa.y + u.y * t, pa.z + u.z * t);
    return len3(mk(x.x - pt.x, x.y - pt.y, x.z - pt.z));
}

// silhouette predicate (scene.cuh:143-174); the 3D view direction is used
// without normalisation, as in the reference
SNCH_HD bool is_silhouette_edge(v3 pa, v3 pb, v3 n0, v3 n1, v3 view, float d, bool flip)
{
    const float precision = 1e-3f;
    float sign = flip ? 1.0f : -1.0f;
    if (d <= precision) {
        v3 ed = normalize3(mk(pb.x - pa.x, pb.y - pa.y, pb.z - pa.z));
        float ang = dm::atan2f_(dot3(ed, cross3(n0, n1)), dot3(n0, n1));
        return sign * ang > precision;
    }
    float dot0 = dot3(view, n0);
    float dot1 = dot3(view, n1);
    if (dm::fabs_(dot0) <= precision)
        return sign * dot1 > precision;
    if (dm::fabs_(dot1) <= precision)
        return sign * dot0 > precision;
    return dot0 * dot1 < 0.0f;
}

// Moller-Trumbore (scene.cuh:1005-1052); the reference's device reciprocal
// __frcp_rn is the correctly rounded 1/det, which IEEE division gives.
// ray = false drops the ray's t >= 0 condition: the infinite line's test.
SNCH_HD bool ray_triangle(v3 o, v3 dir, v3 v0, v3 v1, v3 v2, float& t_out, float& u_out, float& v_out,
                          bool ray = true)
{
    v3 e1 = mk(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z);
    v3 e2 = mk(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z);
    v3 h = mk(dir.y * e2.z - dir.z * e2.y, dir.z * e2.x - dir.x * e2.z, dir.x * e2.y - dir.y * e2.x);
    float det = e1.x * h.x + e1.y * h.y + e1.z * h.z;
    if (dm::fabs_(det) < F_EPS)
        return false;
    float inv = 1.0f / det;
    v3 s = mk(o.x - v0.x, o.y - v0.y, o.z - v0.z);
    float u = (s.x * h.x + s.y * h.y + s.z * h.z) * inv;
    if (u < 0.0f || u > 1.0f)
        return false;
    v3 q = mk(s.y * e1.z - s.z * e1.y, s.z * e1.x - s.x * e1.z, s.x * e1.y - s.y * e1.x);
    float v = (dir.x * q.x + dir.y * q.y + dir.z * q.z) * inv;
    if (v < 0.0f || u + v > 1.0f)
        return false;
    float t = (e2.x * q.x + e2.y * q.y + e2.z * q.z) * inv;
    if (!ray || t >= 0.0f) {
        t_out = t;
        u_out = u;
        v_out = v;
        return true;
    }
    return false;
}
SNCH_HD bool line_triangle(v3 o, v3 dir, v3 v0, v3 v1, v3 v2, float& t_out, float& u_out, float& v_out)
{
    return ray_triangle(o, dir, v0, v1, v2, t_out, u_out, v_out, false);
}

// triangle-sphere test (scene.cuh:1054-1117)
SNCH_HD bool tri_sphere(v3 p1, v3 p2, v3 p3, v3 c, float radius)
{
    v3 e1 = mk(p2.x - p1.x, p2.y - p1.y, p2.z - p1.z);
    v3 e2 = mk(p3.x - p1.x, p3.y - p1.y, p3.z - p1.z);
    v3 n = mk(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    float nl = dm::sqrt_(n.x * n.x + n.y * n.y + n.z * n.z);
    n = mk(n.x / nl, n.y / nl, n.z / nl);
    float d = n.x * p1.x + n.y * p1.y + n.z * p1.z;
    float dp = n.x * c.x + n.y * c.y + n.z * c.z - d;
    v3 proj = mk(c.x - dp * n.x, c.y - dp * n.y, c.z - dp * n.z);
    v3 w0 = mk(p3.x - p1.x, p3.y - p1.y, p3.z - p1.z);
    v3 w1 = mk(p2.x - p1.x, p2.y - p1.y, p2.z - p1.z);
    v3 w2 = mk(proj.x - p1.x, proj.y - p1.y, proj.z - p1.z);
    float d00 = w0.x * w0.x + w0.y * w0.y + w0.z * w0.z;
    float d01 = w0.x * w1.x + w0.y * w1.y + w0.z * w1.z;
    float d02 = w0.x * w2.x + w0.y * w2.y + w0.z * w2.z;
    float d11 = w1.x * w1.x + w1.y * w1.y + w1.z * w1.z;
    float d12 = w1.x * w2.x + w1.y * w2.y + w1.z * w2.z;
    float inv = 1.0f / (d00 * d11 - d01 * d01);
    float u = (d11 * d02 - d01 * d12) * inv;
    float v = (d00 * d12 - d01 * d02) * inv;
    if (u >= 0 && v >= 0 && u + v <= 1)
        return dm::fabs_(dp) <= radius;
    v3 cp = proj;
    if (u < 0)
        cp = p1;
    else if (v < 0)
        cp = p3;
    else if (u + v > 1)
        cp = p2;
    float dx = cp.x - c.x, dy = cp.y - c.y, dz = cp.z - c.z;
    float d2 = dx * dx + dy * dy + dz * dz;
    return d2 <= (radius * radius);
}

// triangle area (scene.cuh:855-868)
SNCH_HD float tri_area(v3 pa, v3 pb, v3 pc)
{
    v3 ac = mk(pc.x - pa.x, pc.y - pa.y, pc.z - pa.z);
    v3 ab = mk(pb.x - pa.x, pb.y - pa.y, pb.z - pa.z);
    return len3(cross3(ac, ab)) / 2.0f;
}

// Green's function weight (scene.cuh:1119-1126)
SNCH_HD float green_weight(v3 x, v3 y)
{
    float r = smax(len3(mk(x.x - y.x, x.y - y.y, x.z - y.z)), 1e-4f);
    return 1.0f / (F_PI * 4.0f * r);
}

// sample_triangle (scene.cuh:14-27)
SNCH_HD v3 sample_triangle(v3 pa, v3 pb, v3 pc, float u, float v)
{
    if (u + v > 1.0f) {
        u = 1.0f - u;
        v = 1.0f - v;
    }
    float w = 1.0f - u - v;
    return mk(w * pa.x + u * pb.x + v * pc.x, w * pa.y + u * pb.y + v * pc.y, w * pa.z + u * pb.z + v * pc.z);
}

// Morton code (morton_code.cuh:19-38, 61-70)
SNCH_HD uint32_t expand_bits(uint32_t v)
{
    v = (v | (v << 16)) & 0x070000FFu;
    v = (v | (v << 8)) & 0x0700F00Fu;
    v = (v | (v << 4)) & 0x430C30C3u;
    v = (v | (v << 2)) & 0x49249249u;
    return v;
}
SNCH_HD uint32_t morton3(v3 p)
{
    const float res = 1024.0f;
    float x = fmn(fmx(p.x * res, 0.0f), res - 1.0f);
    float y = fmn(fmx(p.y * res, 0.0f), res - 1.0f);
    float z = fmn(fmx(p.z * res, 0.0f), res - 1.0f);
    return expand_bits((uint32_t)x) * 4u + expand_bits((uint32_t)y) * 2u + expand_bits((uint32_t)z);
}

// ---------------------------------------------------------------- 2D twins
// scene<2> (scene.cuh:287-703).  2D boxes and cones reuse the 3D records with
// z = 0; the 3D box tests with z = 0 give bit-identical results to the 2D
// ones (every z term is an exact zero, and a 0 * inf slab term is a NaN that
// fminf/fmaxf drop).
constexpr float BVH_OFFSET = 1e-3f; // scene.cuh:12

SNCH_HD float dot2(float ax, float ay, float bx, float by) { return ax * bx + ay * by; }
SNCH_HD float len2(float x, float y) { return dm::sqrt_(x * x + y * y); }

// 2D leaf box of a segment (scene.cuh:420-433)
SNCH_HD box3 seg_box(v3 p0, v3 p1)
{
    box3 b;
    b.hi = mk(fmx(p0.x, p1.x) + BVH_OFFSET, fmx(p0.y, p1.y) + BVH_OFFSET, 0.0f);
    b.lo = mk(fmn(p0.x, p1.x) - BVH_OFFSET, fmn(p0.y, p1.y) - BVH_OFFSET, 0.0f);
    return b;
}

// 2D Morton code (morton_code.cuh:85-92)
SNCH_HD uint32_t morton2(float px, float py)
{
    const float res = 1024.0f;
    float x = fmn(fmx(px * res, 0.0f), res - 1.0f);
    float y = fmn(fmx(py * res, 0.0f), res - 1.0f);
    return expand_bits((uint32_t)x) * 2u + expand_bits((uint32_t)y);
}

// 2D cone-view overlap (cone.cuh:78-121)
SNCH_HD bool cone_overlap2d(v3 axis, float half_angle, float radius, v3 o, const box3& b, float dist_to_box)
{
    if (half_angle >= F_PI_2 || dist_to_box < F_EPS)
        return true;
    const float cx = (b.hi.x + b.lo.x) * 0.5f, cy = (b.hi.y + b.lo.y) * 0.5f;
    float vx = cx - o.x, vy = cy - o.y;
    const float l = len2(vx, vy);
    vx /= l;
    vy /= l;
    const float d_axis = dm::acosf_(clamp11(dot2(axis.x, axis.y, vx, vy)));
    if (inrange(F_PI_2, d_axis - half_angle, d_axis + half_angle))
        return true;
    if (l > radius) {
        const float sum = half_angle + dm::asinf_(radius / l);
        return sum >= F_PI_2 ? true : inrange(F_PI_2, d_axis - sum, d_axis + sum);
    }
    const float ex = b.hi.x - cx, ey = b.hi.y - cy;
    float d = dot2(ex, ey, dm::fabs_(vx), dm::fabs_(vy));
    const float s = l - d;
    if (s <= 0.0f)
        return true;
    d = dm::fabs_(dot2(ex, ey, dm::fabs_(-vy), dm::fabs_(vx))); // project_to_plane (cone.cuh:44-49)
    const float sum = half_angle + dm::atan2f_(d, s);
    return sum >= F_PI_2 ? true : inrange(F_PI_2, d_axis - sum, d_axis + sum);
}

// 2D cone merge (cone.cuh:266-275, 320-372), same fixes as the 3D merge
SNCH_HD cone3 cone_merge2d(const cone3& a, const cone3& b, v3 oa, v3 ob, v3 on, bool conservative)
{
    cone3 r{mk(0.0f, 0.0f, 0.0f), 0.0f, 0.0f};
    const bool va = cone_valid(a.half_angle), vb = cone_valid(b.half_angle);
    if (va && vb) {
        float ax = a.axis.x, ay = a.axis.y, bx = b.axis.x, by = b.axis.y;
        float ha = a.half_angle, hb = b.half_angle;
        const float dax = on.x - oa.x, day = on.y - oa.y, dbx = on.x - ob.x, dby = on.y - ob.y;
        if (conservative)
            r.radius = smax(a.radius + len2(dax, day), b.radius + len2(dbx, dby));
        else
            r.radius = dm::sqrt_(smax(a.radius * a.radius + (dax * dax + day * day),
                                      b.radius * b.radius + (dbx * dbx + dby * dby)));
        if (hb > ha) {
            float t = ax; ax = bx; bx = t;
            t = ay; ay = by; by = t;
            t = ha; ha = hb; hb = t;
        }
        const float theta = dm::acosf_(clamp11(dot2(ax, ay, bx, by)));
        if (smin(theta + hb, F_PI) <= ha) {
            r.axis = mk(ax, ay, 0.0f);
            r.half_angle = ha;
            return r;
        }
        const float o_theta = (ha + theta + hb) / 2.0f;
        if (o_theta >= F_PI) {
            r.axis = mk(ax, ay, 0.0f);
            r.half_angle = F_PI;
            return r;
        }
        float rt = o_theta - ha;
        const float det = ax * by - ay * bx;
        rt *= dm::signbit_(det) ? -1.0f : 1.0f; // copysign(1, det)
        float st, ct;
        dm::sincosf_(rt, &st, &ct);
        r.axis = mk(ct * ax - st * ay, st * ax + ct * ay, 0.0f);
        r.half_angle = o_theta;
    } else if (va) {
        r = a;
    } else if (vb) {
        r = b;
    } else {
        r.half_angle = -F_PI;
    }
    return r;
}

// closest point on a 2D segment, distance only (scene.cuh:176-201)
SNCH_HD float closest_point_segment2d(v3 pa, v3 pb, v3 x)
{
    const float ux = pb.x - pa.x, uy = pb.y - pa.y;
    const float c1 = dot2(ux, uy, x.x - pa.x, x.y - pa.y);
    if (c1 <= 0.0f)
        return len2(x.x - pa.x, x.y - pa.y);
    const float c2 = dot2(ux, uy, ux, uy);
    if (c2 <= c1)
        return len2(x.x - pb.x, x.y - pb.y);
    const float t = c1 / c2;
    const float px = pa.x + ux * t, py = pa.y + uy * t;
    return len2(x.x - px, x.y - py);
}

// silhouette predicate of a vertex (scene.cuh:112-141); 2D normalises the view
SNCH_HD bool is_silhouette_vertex(v3 n0, v3 n1, float vx, float vy, float d, bool flip)
{
    const float precision = 1e-3f;
    const float sign = flip ? 1.0f : -1.0f;
    if (d <= precision)
        return sign * (n0.x * n1.y - n0.y * n1.x) > precision;
    const float ux = vx / d, uy = vy / d;
    const float dot0 = dot2(ux, uy, n0.x, n0.y), dot1 = dot2(ux, uy, n1.x, n1.y);
    if (dm::fabs_(dot0) <= precision)
        return sign * dot1 > precision;
    if (dm::fabs_(dot1) <= precision)
        return sign * dot0 > precision;
    return dot0 * dot1 < 0.0f;
}

// ray-segment test (scene.cuh:545-576); ray = false: the infinite line's test
SNCH_HD bool ray_segment(v3 o, v3 dir, v3 p0, v3 p1, float& t_out, float& s_out, bool ray = true)
{
    const float sx = p1.x - p0.x, sy = p1.y - p0.y;
    const float D = dir.x * (-sy) + dir.y * sx;
    if (dm::fabs_(D) < F_EPS)
        return false;
    const float inv = 1.0f / D;
    const float t = ((p0.x - o.x) * (-sy) - (p0.y - o.y) * (-sx)) * inv;
    const float s = (dir.x * (p0.y - o.y) - dir.y * (p0.x - o.x)) * inv;
    if (s >= -1e-3f && s <= 1.0f + 1e-3f && (!ray || t >= 0.0f)) {
        t_out = t;
        s_out = s;
        return true;
    }
    return false;
}
SNCH_HD bool line_segment(v3 o, v3 dir, v3 p0, v3 p1, float& t_out, float& s_out)
{
    return ray_segment(o, dir, p0, p1, t_out, s_out, false);
}

// segment-sphere test (scene.cuh:581-602)
SNCH_HD bool sphere_segment(v3 p1, v3 p2, v3 c, float radius)
{
    const float dx = p2.x - p1.x, dy = p2.y - p1.y;
    const float len_sq = dx * dx + dy * dy;
    float t = ((c.x - p1.x) * dx + (c.y - p1.y) * dy) / len_sq;
    t = smax(0.0f, smin(1.0f, t));
    const float ex = p1.x + t * dx - c.x, ey = p1.y + t * dy - c.y;
    return ex * ex + ey * ey <= (radius * radius);
}

// 2D Green's function weight (scene.cuh:605-612)
SNCH_HD float green_weight2d(v3 x, v3 y)
{
    const float r = fmx(len2(x.x - y.x, x.y - y.y), 1e-2f);
    return dm::fabs_(dm::logf_(r) / (F_PI * 2.0f));
}

} // namespace snch
