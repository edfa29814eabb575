// Wavefront walk-on-stars (Sawhney, Miller, Gkioulekas, Crane, SIGGRAPH 2023)
// for the Laplace equation with Dirichlet data g(x) = x * y on one part of the
// boundary and zero Neumann flux on the rest.  No counterpart in the
// reference (its caller, elaina, lives outside it); this is the consumer that
// drives all four queries per step.
//
// Per step, over the live walks only (device-side counts, so the loop can be
// captured once as a hipGraph and replayed):
//   k_gather      live walk positions -> compact query arrays
//   closest point on the Dirichlet scene     -> d_D
//   closest silhouette on the Neumann scene, searched only within d_D (the
//                 star radius is min(d_D, d_S), so a farther silhouette
//                 cannot matter)                      -> d_S
//   k_star        terminate inside the eps-shell (value = g at the closest
//                 Dirichlet point); else star radius R = min(d_D, d_S), a
//                 direction uniform on the sphere (hemisphere into the
//                 domain for walks on the Neumann boundary), ray and sphere
//   ray on the Neumann scene (tmax = R)
//   SampleTriangleInSphere on the Neumann scene (Neumann source term; the
//                 zero flux makes its contribution zero, the query still runs)
//   k_update      move to the ray hit (now on the Neumann boundary) or to
//                 x + R v; walks that reach the step limit take g at their
//                 closest Dirichlet point; survivors are appended to the
//                 next live list
// Walk order in the live list depends on atomics, but every walk's arithmetic
// depends only on its own counter-based random stream, so results are
// deterministic.
#include <hip/hip_runtime.h>

#include <cstring>

#include "internal.hpp"

namespace snch {
namespace wost {

struct Walks {
    float4* pos;      // xyz, w = 1 when on the Neumann boundary
    float4* nrm;      // inward normal when on the boundary
    float* value;     // estimator output per walk
    uint32_t* steps;
    uint32_t* live[2];
    uint32_t* count;  // [0], [1]: live counts of the two lists, [2..3] scratch
    // compact per-step query arrays
    float* qpos;      // n x 3
    uint32_t* cp_idx;
    float* cp_dist;
    uint32_t* sil_idx;
    float* sil_dist;
    float* ray_o;
    float* ray_d;
    float* ray_t;
    uint8_t* ray_hit;
    uint32_t* ray_idx;
    float* sph_r;
    float* sph_u;
    int32_t* smp_idx;
    float* smp_pdf;
    float* star_r;    // R per compact slot (-inf when the walk terminated; R = 0 is a live walk
                      // sitting on a silhouette, which stays put until max_steps)
};

// counter-based random numbers: splitmix64 of (seed, walk, step, k)
__device__ __forceinline__ float rnd(uint64_t seed, uint32_t walk, uint32_t step, uint32_t k)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (((uint64_t)walk << 24) ^ ((uint64_t)step << 4) ^ k);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((uint32_t)(z >> 40) + 0.5f) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ v3 closest_point_on_tri(v3 a, v3 b, v3 c, v3 p)
{
    // Ericson, Real-Time Collision Detection 5.1.5 (the point, not just the distance)
    const v3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f)
        return a;
    const v3 bp = sub(p, b);
    const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0f && d4 <= d3)
        return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float v = d1 / (d1 - d3);
        return mk(a.x + v * ab.x, a.y + v * ab.y, a.z + v * ab.z);
    }
    const v3 cp = sub(p, c);
    const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0f && d5 <= d6)
        return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float w = d2 / (d2 - d6);
        return mk(a.x + w * ac.x, a.y + w * ac.y, a.z + w * ac.z);
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return mk(b.x + w * (c.x - b.x), b.y + w * (c.y - b.y), b.z + w * (c.z - b.z));
    }
    const float den = 1.0f / (va + vb + vc);
    const float v = vb * den, w = vc * den;
    return mk(a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w);
}

__device__ __forceinline__ v3 tri_vertex(const float4* V, const int32_t* F, uint32_t t, int k)
{
    const float4 a = V[F[3 * t + k]];
    return mk(a.x, a.y, a.z);
}

__global__ void k_init(Walks w, const float* __restrict__ x0, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) {
        w.count[0] = n;
        w.count[1] = 0;
    }
    if (i >= n)
        return;
    w.pos[i] = make_float4(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2], 0.0f);
    w.nrm[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    w.value[i] = 0.0f;
    w.steps[i] = 0;
    w.live[0][i] = i;
}

// Morton code of each live walk's position in the Dirichlet scene's root box
// (entries past the live count sort last), so that the step's queries visit
// the walks in spatial order
__global__ void k_walk_codes(Walks w, int cur, uint32_t cap, const float4* __restrict__ nbox,
                             uint32_t* __restrict__ codes, uint32_t* __restrict__ idx)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= cap)
        return;
    const uint32_t n = min(w.count[cur], cap);
    if (i >= n) {
        codes[i] = 0xFFFFFFFFu;
        idx[i] = 0xFFFFFFFFu;
        return;
    }
    const uint32_t wk = w.live[cur][i];
    const float4 p = w.pos[wk], lo = nbox[0], hi = nbox[1];
    codes[i] = morton3(mk((p.x - lo.x) / (hi.x - lo.x), (p.y - lo.y) / (hi.y - lo.y), (p.z - lo.z) / (hi.z - lo.z)));
    idx[i] = wk;
}

__global__ void k_apply_order(Walks w, int cur, uint32_t cap, const uint32_t* __restrict__ sorted)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = min(w.count[cur], cap);
    if (i < n)
        w.live[cur][i] = sorted[i];
}

__global__ void k_gather(Walks w, int cur, uint32_t cap)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0)
        w.count[cur ^ 1] = 0; // next list starts empty (a kernel, not a memset: graph safe)
    const uint32_t n = min(w.count[cur], cap);
    if (i >= n)
        return;
    const float4 p = w.pos[w.live[cur][i]];
    w.qpos[3 * i] = p.x;
    w.qpos[3 * i + 1] = p.y;
    w.qpos[3 * i + 2] = p.z;
}

__global__ void k_star(Walks w, int cur, uint32_t cap, const float4* __restrict__ VD, const int32_t* __restrict__ FD,
                       float eps, uint64_t seed, int has_neumann)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = min(w.count[cur], cap);
    if (i >= n)
        return;
    const uint32_t wk = w.live[cur][i];
    const float4 p4 = w.pos[wk];
    const v3 p = mk(p4.x, p4.y, p4.z);
    const float dD = w.cp_dist[i];
    const float dS = has_neumann ? w.sil_dist[i] : F_INF;
    if (dD < eps) {
        const uint32_t t = w.cp_idx[i];
        const v3 c = closest_point_on_tri(tri_vertex(VD, FD, t, 0), tri_vertex(VD, FD, t, 1), tri_vertex(VD, FD, t, 2), p);
        w.value[wk] = c.x * c.y; // g(x) = x y
        w.star_r[i] = -F_INF; // also an empty ray range: the ray query finds nothing
        return;
    }
    const float R = dD < dS ? dD : dS;
    const uint32_t step = w.steps[wk];
    // direction uniform on S^2 (z = 1 - 2u, phi = 2 pi u')
    const float u1 = rnd(seed, wk, step, 0), u2 = rnd(seed, wk, step, 1);
    const float z = 1.0f - 2.0f * u1;
    const float s = dm::sqrt_(fmx(0.0f, 1.0f - z * z));
    float sp, cp;
    dm::sincosf_(6.283185307179586f * u2, &sp, &cp); // deterministic: the CPU restatement matches bit for bit
    v3 d = mk(s * cp, s * sp, z);
    v3 o = p;
    if (p4.w > 0.5f) { // on the Neumann boundary: reflect into the domain, nudge off the surface
        const float4 nn = w.nrm[wk];
        const float dn = d.x * nn.x + d.y * nn.y + d.z * nn.z;
        if (dn < 0.0f)
            d = mk(d.x - 2.0f * dn * nn.x, d.y - 2.0f * dn * nn.y, d.z - 2.0f * dn * nn.z);
        o = mk(p.x + 1e-5f * nn.x, p.y + 1e-5f * nn.y, p.z + 1e-5f * nn.z);
    }
    w.ray_o[3 * i] = o.x;
    w.ray_o[3 * i + 1] = o.y;
    w.ray_o[3 * i + 2] = o.z;
    w.ray_d[3 * i] = d.x;
    w.ray_d[3 * i + 1] = d.y;
    w.ray_d[3 * i + 2] = d.z;
    w.star_r[i] = R;
    w.sph_r[i] = R;
    w.sph_u[3 * i] = rnd(seed, wk, step, 2);
    w.sph_u[3 * i + 1] = rnd(seed, wk, step, 3);
    w.sph_u[3 * i + 2] = rnd(seed, wk, step, 4);
}

__global__ void k_update(Walks w, int cur, uint32_t cap, const float4* __restrict__ VD, const int32_t* __restrict__ FD,
                         const float4* __restrict__ VN, const int32_t* __restrict__ FN, uint32_t max_steps,
                         int has_neumann)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = min(w.count[cur], cap);
    if (i >= n)
        return;
    const float R = w.star_r[i];
    if (!(R >= 0.0f))
        return; // terminated in k_star
    const uint32_t wk = w.live[cur][i];
    const v3 o = mk(w.ray_o[3 * i], w.ray_o[3 * i + 1], w.ray_o[3 * i + 2]);
    const v3 d = mk(w.ray_d[3 * i], w.ray_d[3 * i + 1], w.ray_d[3 * i + 2]);
    const bool hit = has_neumann && w.ray_hit[i] && w.ray_t[i] < R;
    if (hit) {
        const float t = w.ray_t[i];
        const uint32_t tri = w.ray_idx[i];
        const v3 a = tri_vertex(VN, FN, tri, 0), b = tri_vertex(VN, FN, tri, 1), c = tri_vertex(VN, FN, tri, 2);
        v3 nn = normalize3(cross3(sub(b, a), sub(c, a)));
        if (dot3(nn, d) > 0.0f) // inward normal faces the arriving walk
            nn = mk(-nn.x, -nn.y, -nn.z);
        w.pos[wk] = make_float4(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z, 1.0f);
        w.nrm[wk] = make_float4(nn.x, nn.y, nn.z, 0.0f);
    } else {
        const float4 p = w.pos[wk];
        w.pos[wk] = make_float4(p.x + R * d.x, p.y + R * d.y, p.z + R * d.z, 0.0f);
    }
    const uint32_t st = w.steps[wk] + 1;
    w.steps[wk] = st;
    if (st >= max_steps) {
        const uint32_t tri = w.cp_idx[i];
        const float4 p = w.pos[wk];
        const v3 c = closest_point_on_tri(tri_vertex(VD, FD, tri, 0), tri_vertex(VD, FD, tri, 1),
                                          tri_vertex(VD, FD, tri, 2), mk(p.x, p.y, p.z));
        w.value[wk] = c.x * c.y;
        return;
    }
    const uint32_t slot = atomicAdd(&w.count[cur ^ 1], 1u);
    w.live[cur ^ 1][slot] = wk;
}

} // namespace wost
} // namespace snch

using namespace snch;
using namespace snch::wost;

struct snch_wost {
    snch_context* ctx = nullptr;
    const snch_scene* dir = nullptr;
    const snch_scene* neu = nullptr;
    uint32_t cap = 0;
    DeviceBuffer pos, nrm, value, steps, live0, live1, count, qpos, cp_idx, cp_dist, sil_idx, sil_dist, ray_o, ray_d,
        ray_t, ray_hit, ray_idx, sph_r, sph_u, smp_idx, smp_pdf, star_r, ocodes, oidx, ocodes_alt, oidx_alt, ohist,
        sil_ctr;
    Walks w{};
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint32_t graph_steps = 0;
    uint64_t graph_seed = 0;
    float graph_eps = 0.0f;
    // build generations of the two scenes when the graph was captured: the
    // graph holds their tree views (pointers, cone-test flags) by value
    uint64_t graph_gen_d = 0, graph_gen_n = 0;
    ~snch_wost()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
    }
};

namespace {

hipError_t enqueue_steps(snch_wost* s, uint32_t max_steps, float eps, uint64_t seed, hipStream_t st)
{
    const uint32_t cap = s->cap;
    const uint32_t nb = (cap + 255) / 256;
    TreeView vd = s->dir->view();
    TreeView vn = s->neu ? s->neu->view() : TreeView{};
    const int has_n = s->neu ? 1 : 0;
    const uint32_t n_cu = s->ctx && s->ctx->n_cu > 0 ? (uint32_t)s->ctx->n_cu : 256u;
    hipError_t e = hipSuccess;
    for (uint32_t step = 0; step < max_steps; ++step) {
        const int cur = step & 1;
        vd.n_dev = s->w.count + cur;
        vn.n_dev = s->w.count + cur;
        // live walks in Morton order (all buffers were sized at creation: no
        // allocation happens inside a graph capture)
        hipLaunchKernelGGL(k_walk_codes, dim3(nb), dim3(256), 0, st, s->w, cur, cap, vd.nbox, s->ocodes.as<uint32_t>(),
                           s->oidx.as<uint32_t>());
        bool in_alt = false;
        if ((e = radix_sort_pairs(s->ocodes.as<uint32_t>(), s->oidx.as<uint32_t>(), s->ocodes_alt.as<uint32_t>(),
                                  s->oidx_alt.as<uint32_t>(), cap, 32, s->ohist, st, &in_alt)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_apply_order, dim3(nb), dim3(256), 0, st, s->w, cur, cap,
                           in_alt ? s->oidx_alt.as<uint32_t>() : s->oidx.as<uint32_t>());
        hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, st, s->w, cur, cap);
        if ((e = launch_closest_point(vd, 3, s->w.qpos, cap, s->w.cp_idx, s->w.cp_dist, 64, 16, nullptr, nullptr, st)) != hipSuccess)
            return e;
        // the silhouette within the star radius bound (the closest-point
        // distance) on a persistent launch: capped searches vary widely in length
        if (has_n && (e = launch_closest_silhouette_persist(vn, 3, s->w.qpos, cap, 0, s->w.cp_dist, s->w.sil_idx,
                                                            s->w.sil_dist, nullptr, nullptr, s->sil_ctr.as<uint32_t>(),
                                                            n_cu, st)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_star, dim3(nb), dim3(256), 0, st, s->w, cur, cap, s->dir->dV.as<float4>(),
                           s->dir->dF.as<int32_t>(), eps, seed, has_n);
        if (has_n) {
            if ((e = launch_ray(vn, 3, s->w.ray_o, s->w.ray_d, s->w.star_r, cap, 0, s->w.ray_hit, s->w.ray_t, nullptr,
                                s->w.ray_idx, nullptr, st)) != hipSuccess)
                return e;
            if ((e = launch_sample(vn, 3, s->neu->dV.as<float4>(), s->neu->dF.as<int32_t>(), s->w.qpos, s->w.sph_r,
                                   s->w.sph_u, cap, s->w.smp_idx, s->w.smp_pdf, nullptr, st)) != hipSuccess)
                return e;
        }
        hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, st, s->w, cur, cap, s->dir->dV.as<float4>(),
                           s->dir->dF.as<int32_t>(), s->neu ? s->neu->dV.as<float4>() : nullptr,
                           s->neu ? s->neu->dF.as<int32_t>() : nullptr, max_steps, has_n);
    }
    return hipGetLastError();
}

} // namespace

extern "C" {

int snch_wost_create(snch_context* ctx, const snch_scene* dirichlet, const snch_scene* neumann, uint32_t max_walks,
                     snch_wost** out)
{
    if (!ctx || !dirichlet || !out || max_walks == 0 || dirichlet->dim != 3 || (neumann && neumann->dim != 3)) {
        set_error("snch_wost_create: need a context, a built 3D Dirichlet scene and max_walks > 0");
        return SNCH_ERR_INVALID;
    }
    if (!dirichlet->built || (neumann && !neumann->built)) {
        set_error("snch_wost_create: BVH is not built yet.");
        return SNCH_ERR_STATE;
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess)
        return hip_fail(e, "hipSetDevice");
    (void)traversal_error_word(); // allocated outside any graph capture
    auto* s = new snch_wost();
    s->ctx = ctx;
    s->dir = dirichlet;
    s->neu = neumann;
    s->cap = max_walks;
    const size_t n = max_walks;
    struct B {
        DeviceBuffer* b;
        size_t bytes;
    } bufs[] = {{&s->pos, 16 * n},   {&s->nrm, 16 * n},    {&s->value, 4 * n},  {&s->steps, 4 * n},
                {&s->live0, 4 * n},  {&s->live1, 4 * n},   {&s->count, 64},     {&s->qpos, 12 * n},
                {&s->cp_idx, 4 * n}, {&s->cp_dist, 4 * n}, {&s->sil_idx, 4 * n}, {&s->sil_dist, 4 * n},
                {&s->ray_o, 12 * n}, {&s->ray_d, 12 * n},  {&s->ray_t, 4 * n},  {&s->ray_hit, n},
                {&s->ray_idx, 4 * n}, {&s->sph_r, 4 * n},  {&s->sph_u, 12 * n}, {&s->smp_idx, 4 * n},
                {&s->smp_pdf, 4 * n}, {&s->star_r, 4 * n}, {&s->ocodes, 4 * n}, {&s->oidx, 4 * n},
                {&s->ocodes_alt, 4 * n}, {&s->oidx_alt, 4 * n},
                // radix-sort digit counts and look-back words (four 8-bit passes)
                {&s->ohist, radix_sort_scratch_bytes(n, 4)},
                // chunk counters of the persistent silhouette launch (RQ_CTR_WORDS)
                {&s->sil_ctr, 128}};
    for (auto& b : bufs) {
        if ((e = b.b->ensure(b.bytes)) != hipSuccess) {
            delete s;
            return hip_fail(e, "snch_wost_create");
        }
    }
    Walks& w = s->w;
    w.pos = s->pos.as<float4>();
    w.nrm = s->nrm.as<float4>();
    w.value = s->value.as<float>();
    w.steps = s->steps.as<uint32_t>();
    w.live[0] = s->live0.as<uint32_t>();
    w.live[1] = s->live1.as<uint32_t>();
    w.count = s->count.as<uint32_t>();
    w.qpos = s->qpos.as<float>();
    w.cp_idx = s->cp_idx.as<uint32_t>();
    w.cp_dist = s->cp_dist.as<float>();
    w.sil_idx = s->sil_idx.as<uint32_t>();
    w.sil_dist = s->sil_dist.as<float>();
    w.ray_o = s->ray_o.as<float>();
    w.ray_d = s->ray_d.as<float>();
    w.ray_t = s->ray_t.as<float>();
    w.ray_hit = s->ray_hit.as<uint8_t>();
    w.ray_idx = s->ray_idx.as<uint32_t>();
    w.sph_r = s->sph_r.as<float>();
    w.sph_u = s->sph_u.as<float>();
    w.smp_idx = s->smp_idx.as<int32_t>();
    w.smp_pdf = s->smp_pdf.as<float>();
    w.star_r = s->star_r.as<float>();
    *out = s;
    return SNCH_OK;
}

int snch_wost_destroy(snch_wost* s)
{
    delete s;
    return SNCH_OK;
}

// x0: device n x 3 start points; out: device n values.  use_graph = 1 captures
// the max_steps step loop once (per (max_steps, eps, seed)) and replays it.
int snch_wost_solve(snch_wost* s, const float* x0, uint32_t n, uint32_t max_steps, float eps, uint64_t seed,
                    int use_graph, float* out, uint32_t* live_after, void* stream)
{
    if (!s || !x0 || !out || n == 0 || n > s->cap || max_steps == 0) {
        set_error("snch_wost_solve: bad arguments (n must be in 1..max_walks)");
        return SNCH_ERR_INVALID;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    hipLaunchKernelGGL(k_init, dim3((n + 255) / 256), dim3(256), 0, st, s->w, x0, n);
    if (use_graph) {
        const uint64_t gen_d = s->dir->build_gen, gen_n = s->neu ? s->neu->build_gen : 0;
        if (!s->exec || s->graph_steps != max_steps || s->graph_eps != eps || s->graph_seed != seed ||
            s->graph_gen_d != gen_d || s->graph_gen_n != gen_n) {
            if (s->exec) {
                (void)hipGraphExecDestroy(s->exec);
                (void)hipGraphDestroy(s->graph);
                s->exec = nullptr;
                s->graph = nullptr;
            }
            hipStream_t cs;
            if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess)
                return hip_fail(e, "snch_wost_solve: stream");
            if ((e = hipStreamSynchronize(st)) != hipSuccess)
                return hip_fail(e, "snch_wost_solve: sync");
            if ((e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return hip_fail(e, "snch_wost_solve: begin capture");
            hipError_t eq = enqueue_steps(s, max_steps, eps, seed, cs);
            e = hipStreamEndCapture(cs, &s->graph);
            (void)hipStreamDestroy(cs);
            if (eq != hipSuccess)
                return hip_fail(eq, "snch_wost_solve: enqueue");
            if (e != hipSuccess)
                return hip_fail(e, "snch_wost_solve: end capture");
            if ((e = hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0)) != hipSuccess)
                return hip_fail(e, "snch_wost_solve: instantiate");
            s->graph_steps = max_steps;
            s->graph_eps = eps;
            s->graph_seed = seed;
            s->graph_gen_d = gen_d;
            s->graph_gen_n = gen_n;
        }
        if ((e = hipGraphLaunch(s->exec, st)) != hipSuccess)
            return hip_fail(e, "snch_wost_solve: graph launch");
    } else {
        if ((e = enqueue_steps(s, max_steps, eps, seed, st)) != hipSuccess)
            return hip_fail(e, "snch_wost_solve: enqueue");
    }
    if ((e = hipMemcpyAsync(out, s->w.value, sizeof(float) * n, hipMemcpyDeviceToDevice, st)) != hipSuccess)
        return hip_fail(e, "snch_wost_solve: copy");
    if (live_after) {
        uint32_t c[2];
        if ((e = hipMemcpyAsync(c, s->w.count, sizeof(c), hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            return hip_fail(e, "snch_wost_solve: count");
        *live_after = c[max_steps & 1];
    }
    return SNCH_OK;
}

} // extern "C"
