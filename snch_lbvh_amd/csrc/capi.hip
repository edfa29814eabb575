// C ABI of libsnch (include/snch.h): handles, validation, the host-side edge
// table, device uploads and the query entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "internal.hpp"

namespace snch {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return SNCH_ERR_HIP;
}

DeviceBuffer::~DeviceBuffer()
{
    if (p)
        (void)hipFree(p);
}

hipError_t DeviceBuffer::ensure(size_t n)
{
    if (n <= bytes && p)
        return hipSuccess;
    if (p) {
        (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t e = hipMalloc(&p, n > 0 ? n : 16);
    if (e == hipSuccess)
        bytes = n;
    return e;
}

// word 0: a traversal lost a stack entry; word 1: a radix-sort pass gave up
// waiting for an earlier tile (both sticky until read here)
int check_device_errors(const char* fn)
{
    uint32_t* w = traversal_error_word();
    if (!w)
        return SNCH_OK;
    uint32_t h[2] = {0u, 0u};
    if (hipMemcpy(h, w, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
        return hip_fail(hipGetLastError(), fn);
    if (h[0] == 0u && h[1] == 0u)
        return SNCH_OK;
    (void)hipMemset(w, 0, sizeof(h));
    if (h[1]) {
        set_error(std::string(fn) + ": a radix-sort pass exceeded its wait bound; results since the last check are invalid");
        return SNCH_ERR_HIP;
    }
    set_error(std::string(fn) + ": a traversal exceeded its stack; results since the last check may be incomplete");
    return SNCH_ERR_OVERFLOW;
}

// sticky error words per device (never expected: the tree depth is bounded
// by the 64 key bits, and every tile a sort pass waits on has started)
uint32_t* traversal_error_word()
{
    static std::mutex mu;
    static std::vector<uint32_t*> words;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    if ((int)words.size() <= dev)
        words.resize(dev + 1, nullptr);
    if (!words[dev]) {
        void* p = nullptr;
        if (hipMalloc(&p, 64) == hipSuccess) {
            (void)hipMemset(p, 0, 64);
            words[dev] = static_cast<uint32_t*>(p);
        }
    }
    return words[dev];
}

// device inputs of a scene: vertices as float4, primitives with scene-global
// vertex indices; then (2D, from the host) the silhouette slots and the first
// owned silhouette record of each primitive, or (3D) those computed on the
// device; then the mesh table
static hipError_t upload_geometry(snch_scene* s, int dim, const float* Vflat, size_t nV,
                                  const std::vector<int32_t>& Fg)
{
    std::vector<float> V4(4 * nV, 0.0f);
    for (size_t k = 0; k < nV; ++k)
        for (int c = 0; c < dim; ++c)
            V4[4 * k + c] = Vflat[dim * k + c];
    hipError_t e;
    if ((e = s->dV.ensure(sizeof(float) * V4.size())) != hipSuccess ||
        (e = hipMemcpy(s->dV.p, V4.data(), sizeof(float) * V4.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = s->dF.ensure(sizeof(int32_t) * Fg.size())) != hipSuccess ||
        (e = hipMemcpy(s->dF.p, Fg.data(), sizeof(int32_t) * Fg.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    return hipSuccess;
}

// the mesh table (and, for batches, the mesh of each primitive); needs the
// per-mesh silhouette offsets
static hipError_t upload_tables(snch_scene* s, const std::vector<int32_t>& slots_g,
                                const std::vector<uint32_t>& estart_g)
{
    hipError_t e;
    if ((e = s->dSlots.ensure(sizeof(int32_t) * std::max<size_t>(slots_g.size(), 4))) != hipSuccess ||
        (slots_g.size() && (e = hipMemcpy(s->dSlots.p, slots_g.data(), sizeof(int32_t) * slots_g.size(),
                                          hipMemcpyHostToDevice)) != hipSuccess) ||
        (e = s->dEstart.ensure(sizeof(uint32_t) * estart_g.size())) != hipSuccess ||
        (e = hipMemcpy(s->dEstart.p, estart_g.data(), sizeof(uint32_t) * estart_g.size(), hipMemcpyHostToDevice)) !=
            hipSuccess)
        return e;
    return hipSuccess;
}

static hipError_t upload_meshes(snch_scene* s)
{
    const uint32_t M = s->n_meshes;
    std::vector<MeshDesc> desc(M);
    for (uint32_t m = 0; m < M; ++m)
        desc[m] = MeshDesc{s->mesh_off[m], s->mesh_off[m + 1] - s->mesh_off[m], s->mesh_eoff[m],
                           2 * s->mesh_off[m] - m};
    hipError_t e;
    if ((e = s->dMesh.ensure(sizeof(MeshDesc) * M)) != hipSuccess ||
        (e = hipMemcpy(s->dMesh.p, desc.data(), sizeof(MeshDesc) * M, hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    if (M > 1) {
        std::vector<uint32_t> pm(s->nF);
        for (uint32_t m = 0; m < M; ++m)
            std::fill(pm.begin() + s->mesh_off[m], pm.begin() + s->mesh_off[m + 1], m);
        if ((e = s->dPmesh.ensure(sizeof(uint32_t) * pm.size())) != hipSuccess ||
            (e = hipMemcpy(s->dPmesh.p, pm.data(), sizeof(uint32_t) * pm.size(), hipMemcpyHostToDevice)) !=
                hipSuccess)
            return e;
    }
    return hipSuccess;
}

// Per-handle scratch guard.  The query-ordering and batch scratch of a scene
// handle is overwritten by every call; launches are asynchronous, so a call on
// a stream other than the previous call's first makes its stream wait for the
// previous call's launches (an event recorded after them).  Calls on one
// stream are ordered by the stream itself.  Skipped while the stream is being
// captured into a graph: a captured call uses one stream per handle.
static hipError_t scratch_begin(const snch_scene* s, hipStream_t st)
{
    if (!s->scratch_ev || !s->scratch_used || s->scratch_last == st)
        return hipSuccess;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        return hipSuccess;
    return hipStreamWaitEvent(st, s->scratch_ev, 0);
}

static hipError_t scratch_end(const snch_scene* s, hipStream_t st)
{
    if (!s->scratch_ev)
        return hipSuccess;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        return hipSuccess;
    const hipError_t e = hipEventRecord(s->scratch_ev, st);
    if (e == hipSuccess) {
        s->scratch_last = st;
        s->scratch_used = true;
    }
    return e;
}

} // namespace snch

using namespace snch;

snch::TreeView snch_scene::view(uint32_t m) const
{
    const uint32_t off = mesh_off[m], n = mesh_off[m + 1] - off, eoff = mesh_eoff[m];
    TreeView t{};
    t.n_leaves = n;
    t.tnode = tnode.as<float4>() + 4 * (size_t)(off - m);
    t.tcone = tcone.as<float4>() + 4 * (size_t)(off - m);
    // the trig-free cone test is conservative only with the enclosing radius
    t.flags = (dim == 3 && !(build_flags & SNCH_BUILD_REFERENCE_RADIUS)) ? TREE_FAST_CONE : 0u;
    uint32_t vd = 0;
    const float d2 = m < mesh_diag2.size() ? mesh_diag2[m] : 0.0f;
    std::memcpy(&vd, &d2, sizeof(vd));
    vd >>= 20;
    t.qbase = vd > 247u ? vd - 247u : 0u;
    t.tleaf = tleaf.as<float4>() + 3 * (size_t)off;
    t.tedge = tedge.as<float4>() + 4 * (size_t)eoff;
    t.nbox = nbox.as<float4>() + 2 * (size_t)(2 * off - m);
    t.ncone = ncone.as<float4>() + 2 * (size_t)(2 * off - m);
    return t;
}

#define SNCH_CHECK_ARG(cond, msg)   \
    do {                            \
        if (!(cond)) {              \
            set_error(msg);         \
            return SNCH_ERR_INVALID; \
        }                           \
    } while (0)

#define SNCH_HIP(call, what)                \
    do {                                    \
        hipError_t e_ = (call);             \
        if (e_ != hipSuccess)               \
            return hip_fail(e_, what);      \
    } while (0)

extern "C" {

const char* snch_last_error(void) { return g_last_error.c_str(); }

int snch_version(void) { return 10000; }

int snch_context_create(int device, snch_context** out)
{
    SNCH_CHECK_ARG(out, "snch_context_create: out is NULL");
    int count = 0;
    SNCH_HIP(hipGetDeviceCount(&count), "hipGetDeviceCount");
    SNCH_CHECK_ARG(device >= 0 && device < count, "snch_context_create: no such device");
    SNCH_HIP(hipSetDevice(device), "hipSetDevice");
    // the device's sticky error words, allocated here so that no later call
    // (possibly under graph capture) allocates them
    if (!traversal_error_word()) {
        set_error("snch_context_create: cannot allocate the device error words");
        return SNCH_ERR_HIP;
    }
    auto* c = new snch_context();
    c->device = device;
    c->err_word = traversal_error_word();
    {
        const hipError_t e = noalloc_init(device, &c->noalloc_scratch);
        if (e != hipSuccess) {
            delete c;
            return hip_fail(e, "snch_context_create: static scratch");
        }
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess)
        c->n_cu = prop.multiProcessorCount;
    *out = c;
    return SNCH_OK;
}

int snch_context_destroy(snch_context* ctx)
{
    delete ctx;
    return SNCH_OK;
}

int snch_scene_create_tri(snch_context* ctx, const float* V, uint32_t nV, const int32_t* F, uint32_t nF,
                          snch_scene** out)
{
    SNCH_CHECK_ARG(ctx && out, "snch_scene_create_tri: null context or output");
    *out = nullptr;
    const uint32_t voff[2] = {0, nV}, foff[2] = {0, nF};
    if (nF == 0) {
        set_error("snch_scene_create_tri: empty mesh");
        return SNCH_ERR_EMPTY;
    }
    return snch_scene_create_tri_batch(ctx, 1, V, voff, F, foff, out);
}

int snch_scene_create_tri_batch(snch_context* ctx, uint32_t n_meshes, const float* V, const uint32_t* v_off,
                                const int32_t* F, const uint32_t* f_off, snch_scene** out)
{
    SNCH_CHECK_ARG(ctx && out, "snch_scene_create_tri_batch: null context or output");
    *out = nullptr;
    SNCH_CHECK_ARG(n_meshes > 0 && v_off && f_off, "snch_scene_create_tri_batch: no meshes or null offsets");
    SNCH_CHECK_ARG(v_off[0] == 0 && f_off[0] == 0, "snch_scene_create_tri_batch: offsets must start at 0");
    for (uint32_t m = 0; m < n_meshes; ++m) {
        SNCH_CHECK_ARG(v_off[m + 1] >= v_off[m] && f_off[m + 1] >= f_off[m],
                       "snch_scene_create_tri_batch: offsets must not decrease");
        if (f_off[m + 1] == f_off[m]) {
            set_error("snch_scene_create_tri_batch: empty mesh");
            return SNCH_ERR_EMPTY;
        }
        SNCH_CHECK_ARG(v_off[m + 1] > v_off[m], "snch_scene_create_tri_batch: mesh without vertices");
        SNCH_CHECK_ARG(f_off[m + 1] - f_off[m] <= MAX_MESH_PRIMS,
                       "snch_scene_create_tri_batch: too many triangles in one mesh (limit 2^24)");
    }
    const uint32_t nV = v_off[n_meshes], nF = f_off[n_meshes];
    SNCH_CHECK_ARG(V && F, "snch_scene_create_tri_batch: null vertex or index array");
    SNCH_CHECK_ARG(nF < 0x40000000u, "snch_scene_create_tri_batch: too many triangles (limit 2^30)");
    for (size_t k = 0; k < 3 * (size_t)nV; ++k)
        SNCH_CHECK_ARG(std::isfinite(V[k]), "snch_scene_create_tri_batch: non-finite vertex coordinate");
    for (uint32_t m = 0; m < n_meshes; ++m) {
        const uint32_t nv = v_off[m + 1] - v_off[m];
        for (size_t k = 3 * (size_t)f_off[m]; k < 3 * (size_t)f_off[m + 1]; ++k)
            SNCH_CHECK_ARG(F[k] >= 0 && (uint32_t)F[k] < nv, "snch_scene_create_tri_batch: vertex index out of range");
    }
    SNCH_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    auto* s = new snch_scene();
    s->ctx = ctx;
    if (hipEventCreateWithFlags(&s->scratch_ev, hipEventDisableTiming) != hipSuccess) {
        delete s;
        return hip_fail(hipGetLastError(), "snch_scene_create_tri_batch: event");
    }
    s->dim = 3;
    s->n_meshes = n_meshes;
    s->nV = nV;
    s->nF = nF;
    s->mesh_off.assign(f_off, f_off + n_meshes + 1);
    s->mesh_voff.assign(v_off, v_off + n_meshes + 1);
    s->mesh_toff.assign(1, 0);
    s->mesh_eoff.assign(1, 0);
    s->mesh_diag2.assign(n_meshes, 0.0f);
    for (uint32_t m = 0; m < n_meshes; ++m) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t k = v_off[m]; k < v_off[m + 1]; ++k)
            for (int c = 0; c < 3; ++c) {
                lo[c] = std::min(lo[c], V[3 * k + c]);
                hi[c] = std::max(hi[c], V[3 * k + c]);
            }
        double d2 = 0.0;
        for (int c = 0; c < 3; ++c)
            d2 += (double)(hi[c] - lo[c]) * (hi[c] - lo[c]);
        s->mesh_diag2[m] = (float)d2;
    }
    s->F.resize(3 * (size_t)nF);
    for (uint32_t m = 0; m < n_meshes; ++m)
        for (size_t k = 3 * (size_t)f_off[m]; k < 3 * (size_t)f_off[m + 1]; ++k)
            s->F[k] = F[k] + (int32_t)v_off[m];
    // edge table on the device (edges.hip): slots, first owned edge of each
    // triangle, per-mesh edge ranges; the host copies are made on demand
    hipError_t e = upload_geometry(s, 3, V, nV, s->F);
    if (e == hipSuccess)
        e = edge_table_device(s, nullptr);
    if (e == hipSuccess)
        e = upload_meshes(s);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "snch_scene_create_tri_batch: upload");
    }
    *out = s;
    return SNCH_OK;
}

int snch_scene_create_seg(snch_context* ctx, const float* V, uint32_t nV, const int32_t* S, uint32_t nS,
                          snch_scene** out)
{
    SNCH_CHECK_ARG(ctx && out, "snch_scene_create_seg: null context or output");
    *out = nullptr;
    if (nS == 0) {
        set_error("snch_scene_create_seg: empty mesh");
        return SNCH_ERR_EMPTY;
    }
    SNCH_CHECK_ARG(V && S && nV > 0, "snch_scene_create_seg: null vertex or index array");
    SNCH_CHECK_ARG(nS <= MAX_MESH_PRIMS, "snch_scene_create_seg: too many segments (limit 2^24)");
    for (size_t k = 0; k < 2 * (size_t)nV; ++k)
        SNCH_CHECK_ARG(std::isfinite(V[k]), "snch_scene_create_seg: non-finite vertex coordinate");
    for (size_t k = 0; k < 2 * (size_t)nS; ++k)
        SNCH_CHECK_ARG(S[k] >= 0 && (uint32_t)S[k] < nV, "snch_scene_create_seg: vertex index out of range");
    SNCH_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    auto* s = new snch_scene();
    s->ctx = ctx;
    if (hipEventCreateWithFlags(&s->scratch_ev, hipEventDisableTiming) != hipSuccess) {
        delete s;
        return hip_fail(hipGetLastError(), "snch_scene_create_seg: event");
    }
    s->dim = 2;
    s->nV = nV;
    s->nF = nS;
    s->F.assign(S, S + 2 * (size_t)nS);
    // silhouette vertices (scene.cuh:631-652): (prev, v, next, -1) per vertex
    s->slots.assign(4 * (size_t)nV, -1);
    for (uint32_t i = 0; i < nS; ++i) {
        const int32_t a = S[2 * i], b = S[2 * i + 1];
        s->slots[4 * (size_t)a + 1] = a;
        s->slots[4 * (size_t)a + 2] = b;
        s->slots[4 * (size_t)b + 0] = a;
        s->slots[4 * (size_t)b + 1] = b;
    }
    // ownership (scene.cuh:654-672); internal ids follow first ownership so a
    // segment's owned vertices are a consecutive run
    std::vector<unsigned char> seen(nV, 0);
    std::vector<int32_t> internal_slots;
    s->owned.assign(2 * (size_t)nS, -1);
    s->estart.assign((size_t)nS + 1, 0);
    uint32_t next = 0;
    for (uint32_t i = 0; i < nS; ++i) {
        int ptr = 0;
        for (int k = 0; k < 2; ++k) {
            const int32_t v = S[2 * i + k];
            if (!seen[v]) {
                seen[v] = 1;
                s->owned[2 * (size_t)i + ptr++] = v;
                internal_slots.insert(internal_slots.end(), &s->slots[4 * (size_t)v], &s->slots[4 * (size_t)v] + 4);
                ++next;
            }
        }
        s->estart[i + 1] = next;
    }
    s->nE = nV;
    s->mesh_off = {0, nS};
    s->mesh_voff = {0, nV};
    s->mesh_toff = {0, nV};
    s->mesh_eoff = {0, next};
    {
        float lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
        for (size_t k = 0; k < nV; ++k)
            for (int c = 0; c < 2; ++c) {
                lo[c] = std::min(lo[c], V[2 * k + c]);
                hi[c] = std::max(hi[c], V[2 * k + c]);
            }
        s->mesh_diag2 = {(float)((double)(hi[0] - lo[0]) * (hi[0] - lo[0]) + (double)(hi[1] - lo[1]) * (hi[1] - lo[1]))};
    }
    std::vector<float> Vv(V, V + 2 * (size_t)nV);
    hipError_t e = upload_geometry(s, 2, Vv.data(), nV, s->F);
    if (e == hipSuccess)
        e = upload_tables(s, internal_slots, s->estart);
    if (e == hipSuccess)
        e = upload_meshes(s);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "snch_scene_create_seg: upload");
    }
    s->n_internal_sil = next;
    *out = s;
    return SNCH_OK;
}

snch_scene::~snch_scene()
{
    if (build_exec)
        (void)hipGraphExecDestroy(build_exec);
    if (scratch_ev)
        (void)hipEventDestroy(scratch_ev);
}

namespace {
// capture build_tree once into a graph (on a private stream); false when
// capture or instantiation fails, and the caller builds eagerly
bool capture_build(snch_scene* s, uint32_t flags, uint32_t leaf, hipStream_t st)
{
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
        return false;
    if (hipStreamSynchronize(st) != hipSuccess || hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return false;
    }
    const hipError_t eb = build_tree(s, flags, leaf, cs);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    bool ok = eb == hipSuccess && ec == hipSuccess && g;
    hipGraphExec_t ex = nullptr;
    if (ok)
        ok = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
    if (g)
        (void)hipGraphDestroy(g);
    if (!ok) {
        (void)hipGetLastError();
        return false;
    }
    if (s->build_exec)
        (void)hipGraphExecDestroy(s->build_exec);
    s->build_exec = ex;
    return true;
}
} // namespace

int snch_scene_destroy(snch_scene* scene)
{
    if (scene) {
        (void)hipSetDevice(scene->ctx->device);
        delete scene;
    }
    return SNCH_OK;
}

int snch_scene_info(const snch_scene* s, uint32_t* dim, uint32_t* n_prims, uint32_t* n_sil)
{
    SNCH_CHECK_ARG(s, "snch_scene_info: null scene");
    if (dim)
        *dim = s->dim;
    if (n_prims)
        *n_prims = s->nF;
    if (n_sil)
        *n_sil = s->nE;
    return SNCH_OK;
}

int snch_scene_build(snch_scene* s, const snch_build_opts* opts, void* stream)
{
    SNCH_CHECK_ARG(s, "snch_scene_build: null scene");
    const uint32_t flags = opts ? opts->flags : 0u;
    const uint32_t leaf = opts && opts->leaf_size ? opts->leaf_size : 1u;
    SNCH_CHECK_ARG(leaf <= MAX_LEAF_SIZE, "snch_scene_build: leaf_size must be 1..8");
    SNCH_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t e0, e1, ph[5];
    const bool phases = (flags & SNCH_BUILD_TIME_PHASES) != 0;
    SNCH_HIP(hipEventCreate(&e0), "hipEventCreate");
    SNCH_HIP(hipEventCreate(&e1), "hipEventCreate");
    for (int k = 0; phases && k < 5; ++k)
        SNCH_HIP(hipEventCreate(&ph[k]), "hipEventCreate");
    // the first build with given (flags, leaf size) runs eagerly and sizes
    // every buffer; the second captures the launch sequence; later ones
    // replay it (no host launch gaps).  Phase timing always runs eagerly.
    const uint32_t key = (flags & ~SNCH_BUILD_TIME_PHASES) | (leaf << 8) | 0x80000000u;
    if (!phases && s->eager_key == key && s->eager_builds >= 1 && s->build_key != key) {
        if (capture_build(s, flags, leaf, st))
            s->build_key = key;
    }
    (void)hipEventRecord(e0, st);
    hipError_t e;
    if (!phases && s->build_exec && s->build_key == key) {
        e = hipGraphLaunch(s->build_exec, st);
    } else {
        e = build_tree(s, flags, leaf, st, phases ? ph : nullptr);
        if (s->eager_key != key) {
            s->eager_key = key;
            s->eager_builds = 0;
        }
        ++s->eager_builds;
    }
    (void)hipEventRecord(e1, st);
    hipError_t es = hipStreamSynchronize(st);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int k = 0; k < 4; ++k) {
        s->last_phase_ms[k] = 0.0f;
        if (phases && e == hipSuccess && es == hipSuccess)
            (void)hipEventElapsedTime(&s->last_phase_ms[k], ph[k], ph[k + 1]);
    }
    for (int k = 0; phases && k < 5; ++k)
        (void)hipEventDestroy(ph[k]);
    if (e != hipSuccess)
        return hip_fail(e, "snch_scene_build");
    if (es != hipSuccess)
        return hip_fail(es, "snch_scene_build: synchronize");
    if (const int rc = check_device_errors("snch_scene_build"))
        return rc;
    s->build_flags = flags;
    {
        std::vector<TreeView> views(s->n_meshes);
        for (uint32_t m = 0; m < s->n_meshes; ++m)
            views[m] = s->view(m);
        SNCH_HIP(s->batch.views.ensure(sizeof(TreeView) * views.size()), "snch_scene_build: views");
        SNCH_HIP(hipMemcpy(s->batch.views.p, views.data(), sizeof(TreeView) * views.size(), hipMemcpyHostToDevice),
                 "snch_scene_build: views");
    }
    s->built = true;
    s->build_flags = flags;
    s->leaf_size = leaf;
    ++s->build_gen;
    s->last_build_ms = ms;
    return SNCH_OK;
}

int snch_scene_build_time(const snch_scene* s, float* ms)
{
    SNCH_CHECK_ARG(s && ms, "snch_scene_build_time: null argument");
    *ms = s->last_build_ms;
    return SNCH_OK;
}

int snch_scene_build_phases(const snch_scene* s, float ms[4])
{
    SNCH_CHECK_ARG(s && ms, "snch_scene_build_phases: null argument");
    for (int k = 0; k < 4; ++k)
        ms[k] = s->last_phase_ms[k];
    return SNCH_OK;
}

#define SNCH_NEED_BUILT(s, fn)                                   \
    do {                                                         \
        SNCH_CHECK_ARG(s, fn ": null scene");                    \
        if (!(s)->built) {                                       \
            set_error(fn ": BVH is not built yet.");             \
            return SNCH_ERR_STATE;                               \
        }                                                        \
    } while (0)

#define SNCH_NEED_SINGLE(s, fn)                                              \
    do {                                                                     \
        SNCH_NEED_BUILT(s, fn);                                              \
        SNCH_CHECK_ARG((s)->n_meshes == 1, fn ": the scene holds several meshes; use the batch entry points"); \
    } while (0)

// query visiting order: 0 = Morton order from SNCH_ORDER_MIN_QUERIES queries
// up, 1 = always Morton order, 2 = input order
static bool want_order(const snch_query_opts* opts, uint32_t n)
{
    const uint32_t mode = opts ? opts->coherent : 0u;
    return mode == 1u ? n > 1 : (mode == 0u && n >= SNCH_ORDER_MIN_QUERIES);
}

int snch_query_order(const snch_scene* s, const float* q, uint32_t n, uint32_t* out_perm, void* stream)
{
    SNCH_NEED_SINGLE(s, "snch_query_order");
    SNCH_CHECK_ARG(n == 0 || (q && out_perm), "snch_query_order: null pointer");
    if (n == 0)
        return SNCH_OK;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t* perm = nullptr;
    SNCH_HIP(scratch_begin(s, st), "snch_query_order: stream wait");
    SNCH_HIP(query_order(s->view(), (int)s->dim, q, n, s->qscratch, st, &perm), "snch_query_order");
    if (perm)
        SNCH_HIP(hipMemcpyAsync(out_perm, perm, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st),
                 "snch_query_order: copy");
    else
        SNCH_HIP(hipMemsetAsync(out_perm, 0, sizeof(uint32_t), st), "snch_query_order: copy");
    SNCH_HIP(scratch_end(s, st), "snch_query_order: event");
    return SNCH_OK;
}

int snch_closest_point_ex(const snch_scene* s, const float* q, uint32_t n, uint32_t* out_idx, float* out_dist,
                          const snch_query_opts* opts, void* stream)
{
    SNCH_NEED_SINGLE(s, "snch_closest_point");
    SNCH_CHECK_ARG(n == 0 || (q && out_idx && out_dist), "snch_closest_point: null pointer");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t block = opts && opts->block_size ? opts->block_size : 64u;
    const uint32_t depth = opts && opts->stack_depth ? opts->stack_depth : 16u;
    SNCH_CHECK_ARG(block == 64 || block == 128 || block == 256 || block == 512,
                   "snch_closest_point: block_size must be 64, 128, 256 or 512");
    SNCH_CHECK_ARG(depth == 8 || depth == 16, "snch_closest_point: stack_depth must be 8 or 16");
    const uint32_t* perm = opts ? opts->order : nullptr;
    const bool own_order = !perm && want_order(opts, n);
    if (own_order) {
        SNCH_HIP(scratch_begin(s, st), "snch_closest_point: stream wait");
        SNCH_HIP(query_order(s->view(), (int)s->dim, q, n, s->qscratch, st, &perm), "snch_closest_point: ordering");
    }
    SNCH_HIP(launch_closest_point(s->view(), (int)s->dim, q, n, out_idx, out_dist, block, depth, perm,
                                  opts ? opts->stats : nullptr, st),
             "snch_closest_point");
    if (own_order)
        SNCH_HIP(scratch_end(s, st), "snch_closest_point: event");
    return SNCH_OK;
}

int snch_closest_point(const snch_scene* s, const float* q, uint32_t n, uint32_t* out_idx, float* out_dist,
                       void* stream)
{
    return snch_closest_point_ex(s, q, n, out_idx, out_dist, nullptr, stream);
}

int snch_closest_silhouette_ex(const snch_scene* s, const float* q, uint32_t n, int flip, const float* seed,
                               uint32_t* out_sil, float* out_dist, const snch_query_opts* opts, void* stream)
{
    SNCH_NEED_SINGLE(s, "snch_closest_silhouette");
    SNCH_CHECK_ARG(n == 0 || (q && out_dist), "snch_closest_silhouette: null pointer");
    uint32_t block = opts && opts->block_size ? opts->block_size : 64u;
    const uint32_t depth = opts && opts->stack_depth ? opts->stack_depth : 16u;
    SNCH_CHECK_ARG(block == 64 || block == 128 || block == 256 || block == 512,
                   "snch_closest_silhouette: block_size must be 64, 128, 256 or 512");
    SNCH_CHECK_ARG(depth == 8 || depth == 16 || depth == 24 || depth == 32,
                   "snch_closest_silhouette: stack_depth must be 8, 16, 24 or 32");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t* perm = opts ? opts->order : nullptr;
    const bool own_order = !perm && want_order(opts, n);
    const bool capped = opts && opts->seed_is_cap;
    // a seed radius, not a hard cap: unsettled queries are re-queued for whole
    // waves unless asked to search again in place (default launch shape only)
    const bool requeue = seed && !capped && n > 0 && !(opts && opts->seed_retry == 1) && block == 64 && depth == 16;
    // no seed or a hard cap: a persistent launch taking 64-query chunks (default launch shape)
    const bool persist = (!seed || capped) && n > 0 && block == 64 && depth == 16;
    if (own_order || requeue || persist)
        SNCH_HIP(scratch_begin(s, st), "snch_closest_silhouette: stream wait");
    if (own_order)
        SNCH_HIP(query_order(s->view(), (int)s->dim, q, n, s->qscratch, st, &perm),
                 "snch_closest_silhouette: ordering");
    if (persist) {
        SNCH_HIP(s->rq_ctr.ensure(sizeof(uint32_t) * 32), "snch_closest_silhouette: allocation");
        const uint32_t n_cu = s->ctx && s->ctx->n_cu > 0 ? (uint32_t)s->ctx->n_cu : 256u;
        SNCH_HIP(launch_closest_silhouette_persist(s->view(), (int)s->dim, q, n, flip, seed, out_sil, out_dist,
                                                   perm, opts ? opts->stats : nullptr, s->rq_ctr.as<uint32_t>(), n_cu,
                                                   st),
                 "snch_closest_silhouette");
    } else if (requeue) {
        SNCH_HIP(s->rq_ctr.ensure(sizeof(uint32_t) * 32), "snch_closest_silhouette: allocation");
        SNCH_HIP(s->rq_queue.ensure(sizeof(uint32_t) * 8 * (size_t)n), "snch_closest_silhouette: allocation");
        const uint32_t n_cu = s->ctx && s->ctx->n_cu > 0 ? (uint32_t)s->ctx->n_cu : 256u;
        SNCH_HIP(launch_closest_silhouette_requeue(s->view(), (int)s->dim, q, n, flip, seed, out_sil, out_dist, perm,
                                                   opts ? opts->stats : nullptr, s->rq_ctr.as<uint32_t>(),
                                                   s->rq_queue.as<uint32_t>(), n_cu, st),
                 "snch_closest_silhouette");
    } else {
        SNCH_HIP(launch_closest_silhouette(s->view(), (int)s->dim, q, n, flip, seed, capped, out_sil, out_dist,
                                           block, depth, perm, opts ? opts->stats : nullptr, st),
                 "snch_closest_silhouette");
    }
    if (own_order || requeue || persist)
        SNCH_HIP(scratch_end(s, st), "snch_closest_silhouette: event");
    return SNCH_OK;
}

int snch_closest_silhouette(const snch_scene* s, const float* q, uint32_t n, int flip, const float* seed,
                            uint32_t* out_sil, float* out_dist, void* stream)
{
    return snch_closest_silhouette_ex(s, q, n, flip, seed, out_sil, out_dist, nullptr, stream);
}

int snch_ray_ex(const snch_scene* s, const float* o, const float* d, const float* tmax, uint32_t n, int any_hit,
                uint8_t* out_hit, float* out_t, float* out_uv, uint32_t* out_idx, const snch_query_opts* opts,
                void* stream)
{
    SNCH_NEED_SINGLE(s, "snch_ray");
    SNCH_CHECK_ARG(n == 0 || (o && d && out_hit), "snch_ray: null pointer");
    SNCH_HIP(launch_ray(s->view(), (int)s->dim, o, d, tmax, n, any_hit, out_hit, out_t, out_uv, out_idx,
                        opts ? opts->stats : nullptr, static_cast<hipStream_t>(stream)),
             "snch_ray");
    return SNCH_OK;
}

int snch_ray(const snch_scene* s, const float* o, const float* d, const float* tmax, uint32_t n, int any_hit,
             uint8_t* out_hit, float* out_t, float* out_uv, uint32_t* out_idx, void* stream)
{
    return snch_ray_ex(s, o, d, tmax, n, any_hit, out_hit, out_t, out_uv, out_idx, nullptr, stream);
}

int snch_sample_in_sphere(const snch_scene* s, const float* c, const float* r, const float* u, uint32_t n,
                          int32_t* out_idx, float* out_pdf, float* out_pt, void* stream)
{
    SNCH_NEED_SINGLE(s, "snch_sample_in_sphere");
    SNCH_CHECK_ARG(n == 0 || (c && r && u && out_idx && out_pdf), "snch_sample_in_sphere: null pointer");
    SNCH_HIP(launch_sample(s->view(), (int)s->dim, s->dV.as<float4>(), s->dF.as<int32_t>(), c, r, u, n, out_idx, out_pdf, out_pt,
                           static_cast<hipStream_t>(stream)),
             "snch_sample_in_sphere");
    return SNCH_OK;
}

int snch_overlap(const snch_scene* s, const float* lo, const float* hi, uint32_t n, uint32_t max_per_query,
                 uint32_t* out_count, uint32_t* out_idx, void* stream)
{
    SNCH_NEED_SINGLE(s, "snch_overlap");
    SNCH_CHECK_ARG(n == 0 || (lo && hi && out_count && (max_per_query == 0 || out_idx)),
                   "snch_overlap: null pointer");
    SNCH_HIP(launch_generic(s->view(), (int)s->dim, 0, lo, hi, n, max_per_query, out_count, out_idx, nullptr,
                            static_cast<hipStream_t>(stream)),
             "snch_overlap");
    return SNCH_OK;
}

int snch_line_intersect(const snch_scene* s, const float* o, const float* d, uint32_t n, uint32_t max_per_query,
                        uint32_t* out_count, uint32_t* out_idx, float* out_t, void* stream)
{
    SNCH_NEED_SINGLE(s, "snch_line_intersect");
    SNCH_CHECK_ARG(n == 0 || (o && d && out_count && (max_per_query == 0 || out_idx)),
                   "snch_line_intersect: null pointer");
    SNCH_HIP(launch_generic(s->view(), (int)s->dim, 1, o, d, n, max_per_query, out_count, out_idx, out_t,
                            static_cast<hipStream_t>(stream)),
             "snch_line_intersect");
    return SNCH_OK;
}

int snch_sample_on_object(const snch_scene* s, const int32_t* idx, const float* uv, uint32_t n, float* out_pt,
                          void* stream)
{
    SNCH_CHECK_ARG(s, "snch_sample_on_object: null scene");
    SNCH_CHECK_ARG(n == 0 || (idx && uv && out_pt), "snch_sample_on_object: null pointer");
    SNCH_HIP(launch_sample_on_object((int)s->dim, s->dV.as<float4>(), s->dF.as<int32_t>(), s->nF, idx, uv, n, out_pt,
                                     static_cast<hipStream_t>(stream)),
             "snch_sample_on_object");
    return SNCH_OK;
}

static int batch_config(const snch_query_opts* opts, uint32_t n, BatchConfig* cfg, const char* fn)
{
    cfg->block_size = opts && opts->block_size ? opts->block_size : 64u;
    cfg->stack_depth = opts && opts->stack_depth ? opts->stack_depth : 16u;
    cfg->queries_per_wave = opts && opts->queries_per_wave ? opts->queries_per_wave : 64u;
    const uint32_t mode = opts ? opts->coherent : 0u;
    cfg->coherent = mode == 1u || (mode == 0u && n >= SNCH_ORDER_MIN_QUERIES);
    const uint32_t b = cfg->block_size, d = cfg->stack_depth, w = cfg->queries_per_wave;
    if (!(b == 64 || b == 128 || b == 256 || b == 512) || !(d == 8 || d == 16 || d == 24 || d == 32) || w % 64 != 0 ||
        w > 4096 || (opts && opts->order)) {
        set_error(std::string(fn) + ": block_size must be 64..512 (power of 2), stack_depth 8, 16, 24 or 32, "
                                    "queries_per_wave a multiple of 64 up to 4096, and order NULL");
        return SNCH_ERR_INVALID;
    }
    return SNCH_OK;
}

static int check_qoff(const snch_scene* s, const uint32_t* q_off, const char* fn)
{
    if (!q_off || q_off[0] != 0) {
        set_error(std::string(fn) + ": q_off must be a host array of n_meshes + 1 entries starting at 0");
        return SNCH_ERR_INVALID;
    }
    for (uint32_t m = 0; m < s->n_meshes; ++m)
        if (q_off[m + 1] < q_off[m]) {
            set_error(std::string(fn) + ": q_off must not decrease");
            return SNCH_ERR_INVALID;
        }
    return SNCH_OK;
}

int snch_batch_closest_point(const snch_scene* s, const float* q, const uint32_t* q_off, uint32_t* out_idx,
                             float* out_dist, const snch_query_opts* opts, void* stream)
{
    SNCH_NEED_BUILT(s, "snch_batch_closest_point");
    SNCH_CHECK_ARG(s->dim == 3, "snch_batch_closest_point: batches hold triangle meshes");
    int rc = check_qoff(s, q_off, "snch_batch_closest_point");
    if (rc != SNCH_OK)
        return rc;
    const uint32_t n = q_off[s->n_meshes];
    SNCH_CHECK_ARG(n == 0 || (q && out_idx && out_dist), "snch_batch_closest_point: null pointer");
    BatchConfig cfg;
    if ((rc = batch_config(opts, n, &cfg, "snch_batch_closest_point")) != SNCH_OK)
        return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    SNCH_HIP(scratch_begin(s, st), "snch_batch_closest_point: stream wait");
    SNCH_HIP(launch_batch(s, 0, q, q_off, 0, out_idx, out_dist, cfg, st), "snch_batch_closest_point");
    SNCH_HIP(scratch_end(s, st), "snch_batch_closest_point: event");
    return SNCH_OK;
}

int snch_batch_closest_silhouette(const snch_scene* s, const float* q, const uint32_t* q_off, int flip,
                                  const float* seed_radius, uint32_t* out_sil, float* out_dist,
                                  const snch_query_opts* opts, void* stream)
{
    SNCH_NEED_BUILT(s, "snch_batch_closest_silhouette");
    SNCH_CHECK_ARG(s->dim == 3, "snch_batch_closest_silhouette: batches hold triangle meshes");
    int rc = check_qoff(s, q_off, "snch_batch_closest_silhouette");
    if (rc != SNCH_OK)
        return rc;
    const uint32_t n = q_off[s->n_meshes];
    SNCH_CHECK_ARG(n == 0 || (q && out_dist), "snch_batch_closest_silhouette: null pointer");
    BatchConfig cfg;
    if ((rc = batch_config(opts, n, &cfg, "snch_batch_closest_silhouette")) != SNCH_OK)
        return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    SNCH_HIP(scratch_begin(s, st), "snch_batch_closest_silhouette: stream wait");
    SNCH_HIP(launch_batch(s, 1, q, q_off, flip, out_sil, out_dist, cfg, st, seed_radius),
             "snch_batch_closest_silhouette");
    SNCH_HIP(scratch_end(s, st), "snch_batch_closest_silhouette: event");
    return SNCH_OK;
}

int snch_batch_autotune(snch_scene* s, uint32_t kind, const float* q, const uint32_t* q_off,
                        uint32_t sample_per_mesh, int flip, snch_tune_result* out, void* stream)
{
    SNCH_NEED_BUILT(s, "snch_batch_autotune");
    SNCH_CHECK_ARG(out && kind <= 1u && s->dim == 3, "snch_batch_autotune: bad argument");
    int rc = check_qoff(s, q_off, "snch_batch_autotune");
    if (rc != SNCH_OK)
        return rc;
    SNCH_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    SNCH_HIP(scratch_begin(s, st), "snch_batch_autotune: stream wait");
    const uint32_t M = s->n_meshes;
    // the tuning set: every query (sample_per_mesh = 0) or the first
    // sample_per_mesh queries of each mesh, gathered into one batch
    std::vector<uint32_t> soff(M + 1, 0);
    for (uint32_t m = 0; m < M; ++m)
        soff[m + 1] = soff[m] + (sample_per_mesh ? std::min(q_off[m + 1] - q_off[m], sample_per_mesh)
                                                 : q_off[m + 1] - q_off[m]);
    const uint32_t ns = soff[M];
    SNCH_CHECK_ARG(ns > 0 && q, "snch_batch_autotune: no queries");
    DeviceBuffer sq, sidx, sdist;
    const float* tq = q;
    if (sample_per_mesh) {
        SNCH_HIP(sq.ensure(sizeof(float) * 3 * (size_t)ns), "snch_batch_autotune: alloc");
        for (uint32_t m = 0; m < M; ++m)
            if (soff[m + 1] > soff[m])
                SNCH_HIP(hipMemcpyAsync(sq.as<float>() + 3 * (size_t)soff[m], q + 3 * (size_t)q_off[m],
                                        sizeof(float) * 3 * (soff[m + 1] - soff[m]), hipMemcpyDeviceToDevice, st),
                         "snch_batch_autotune: gather");
        tq = sq.as<float>();
    }
    SNCH_HIP(sidx.ensure(sizeof(uint32_t) * ns), "snch_batch_autotune: alloc");
    SNCH_HIP(sdist.ensure(sizeof(float) * ns), "snch_batch_autotune: alloc");
    hipEvent_t e0, e1;
    SNCH_HIP(hipEventCreate(&e0), "hipEventCreate");
    SNCH_HIP(hipEventCreate(&e1), "hipEventCreate");
    hipError_t err = hipSuccess;
    snch_tune_result cur{1, 64, 16, 64, 0.0f, 0.0f, 0};
    // one warm-up and one timed launch of a configuration
    auto measure = [&](const snch_tune_result& c, float* ms) -> hipError_t {
        BatchConfig cfg;
        cfg.block_size = c.block_size;
        cfg.stack_depth = c.stack_depth;
        cfg.queries_per_wave = c.queries_per_wave;
        cfg.coherent = true;
        hipError_t x = launch_batch(s, (int)kind, tq, soff.data(), flip, sidx.as<uint32_t>(), sdist.as<float>(),
                                    cfg, st);
        if (x == hipSuccess)
            x = hipEventRecord(e0, st);
        if (x == hipSuccess)
            x = launch_batch(s, (int)kind, tq, soff.data(), flip, sidx.as<uint32_t>(), sdist.as<float>(), cfg, st);
        if (x == hipSuccess)
            x = hipEventRecord(e1, st);
        if (x == hipSuccess)
            x = hipEventSynchronize(e1);
        if (x == hipSuccess)
            x = hipEventElapsedTime(ms, e0, e1);
        return x;
    };
    // coordinate descent from the default: leaf size (rebuilding), then block
    // size, stack depth and queries per wave, each over its whole range with
    // the others held at their best so far
    const uint32_t flags = s->build_flags;
    float best_ms = 0.0f;
    if ((err = build_tree(s, flags, 1, st)) == hipSuccess)
        err = measure(cur, &best_ms);
    cur.ms_default = best_ms;
    cur.n_tried = 1;
    const uint32_t leaves[4] = {1, 2, 4, 8}, blocks[4] = {64, 128, 256, 512}, depths[4] = {8, 16, 24, 32},
                   qpws[3] = {64, 128, 256};
    for (uint32_t L : leaves) {
        if (err != hipSuccess || L == 1)
            continue;
        float ms = 0.0f;
        snch_tune_result c = cur;
        c.leaf_size = L;
        if ((err = build_tree(s, flags, L, st)) == hipSuccess && (err = measure(c, &ms)) == hipSuccess) {
            ++cur.n_tried;
            if (ms < best_ms) {
                best_ms = ms;
                cur.leaf_size = L;
            }
        }
    }
    if (err == hipSuccess)
        err = build_tree(s, flags, cur.leaf_size, st);
    for (int axis = 0; axis < 3 && err == hipSuccess; ++axis) {
        const uint32_t* vals = axis == 0 ? blocks : axis == 1 ? depths : qpws;
        const int nv = axis == 2 ? 3 : 4;
        const uint32_t start = axis == 0 ? cur.block_size : axis == 1 ? cur.stack_depth : cur.queries_per_wave;
        for (int k = 0; k < nv && err == hipSuccess; ++k) {
            if (vals[k] == start)
                continue; // already measured
            snch_tune_result c = cur;
            uint32_t& field = axis == 0 ? c.block_size : axis == 1 ? c.stack_depth : c.queries_per_wave;
            field = vals[k];
            float ms = 0.0f;
            if ((err = measure(c, &ms)) == hipSuccess) {
                ++cur.n_tried;
                if (ms < best_ms) {
                    best_ms = ms;
                    cur.block_size = c.block_size;
                    cur.stack_depth = c.stack_depth;
                    cur.queries_per_wave = c.queries_per_wave;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err == hipSuccess)
        err = hipStreamSynchronize(st);
    if (err != hipSuccess)
        return hip_fail(err, "snch_batch_autotune");
    s->leaf_size = cur.leaf_size;
    ++s->build_gen;
    cur.ms = best_ms;
    *out = cur;
    return SNCH_OK;
}

int snch_scene_export(const snch_scene* s, uint32_t* nodes, float* boxes, float* cones, uint32_t* sorted_codes,
                      uint32_t* sorted_idx)
{
    SNCH_NEED_SINGLE(s, "snch_scene_export");
    return snch_scene_export_mesh(s, 0, nodes, boxes, cones, sorted_codes, sorted_idx);
}

int snch_scene_device_view(const snch_scene* s, uint32_t mesh, snch_device_view* out)
{
    SNCH_NEED_BUILT(s, "snch_scene_device_view");
    SNCH_CHECK_ARG(out, "snch_scene_device_view: null pointer");
    SNCH_CHECK_ARG(mesh < s->n_meshes, "snch_scene_device_view: mesh index out of range");
    const TreeView t = s->view(mesh);
    out->dim = s->dim;
    out->n_leaves = t.n_leaves;
    out->flags = t.flags;
    out->qbase = t.qbase;
    out->tnode = t.tnode;
    out->tcone = t.tcone;
    out->tleaf = t.tleaf;
    out->tedge = t.tedge;
    out->nbox = t.nbox;
    out->ncone = t.ncone;
    out->V = s->dV.p;
    out->F = s->dF.as<int32_t>() + (size_t)s->dim * s->mesh_off[mesh];
    return SNCH_OK;
}

int snch_scene_export_mesh(const snch_scene* s, uint32_t mesh, uint32_t* nodes, float* boxes, float* cones,
                           uint32_t* sorted_codes, uint32_t* sorted_idx)
{
    SNCH_NEED_BUILT(s, "snch_scene_export_mesh");
    SNCH_CHECK_ARG(mesh < s->n_meshes, "snch_scene_export_mesh: mesh index out of range");
    SNCH_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    const uint32_t off = s->mesh_off[mesh], N = s->mesh_off[mesh + 1] - off, NN = 2 * N - 1;
    const uint32_t nb = 2 * off - mesh;
    const BuildScratch& w = s->scratch;
    const uint32_t* dcodes = s->sorted_in_alt ? w.codes_alt.as<uint32_t>() : w.codes.as<uint32_t>();
    const uint32_t* didx = s->sorted_in_alt ? w.idx_alt.as<uint32_t>() : w.idx.as<uint32_t>();
    std::vector<uint32_t> sidx(N), parent(NN), child(2 * (size_t)NN);
    SNCH_HIP(hipMemcpy(sidx.data(), didx + off, sizeof(uint32_t) * N, hipMemcpyDeviceToHost), "export");
    for (auto& x : sidx)
        x -= off;
    if (sorted_codes)
        SNCH_HIP(hipMemcpy(sorted_codes, dcodes + off, sizeof(uint32_t) * N, hipMemcpyDeviceToHost), "export");
    if (sorted_idx)
        std::memcpy(sorted_idx, sidx.data(), sizeof(uint32_t) * N);
    if (nodes) {
        SNCH_HIP(hipMemcpy(parent.data(), w.parent.as<uint32_t>() + nb, sizeof(uint32_t) * NN,
                           hipMemcpyDeviceToHost),
                 "export");
        SNCH_HIP(hipMemcpy(child.data(), w.child.as<uint32_t>() + 2 * (size_t)nb, sizeof(uint32_t) * 2 * NN,
                           hipMemcpyDeviceToHost),
                 "export");
        for (uint32_t i = 0; i < NN; ++i) {
            nodes[4 * i] = parent[i] == INVALID ? INVALID : parent[i] - nb;
            if (i + 1 < N) {
                nodes[4 * i + 1] = child[2 * i] - nb;
                nodes[4 * i + 2] = child[2 * i + 1] - nb;
                nodes[4 * i + 3] = INVALID;
            } else {
                nodes[4 * i + 1] = INVALID;
                nodes[4 * i + 2] = INVALID;
                nodes[4 * i + 3] = sidx[i - (N - 1)];
            }
        }
    }
    if (boxes || cones) {
        // the root's own box and cone are stored; every other node's sit in
        // its parent's traversal record (child L or R)
        std::vector<float> rb(8), rc(8), tn(16 * (size_t)(N - 1)), tc(16 * (size_t)(N - 1));
        SNCH_HIP(hipMemcpy(rb.data(), s->nbox.as<float>() + 8 * (size_t)nb, sizeof(float) * 8, hipMemcpyDeviceToHost),
                 "export");
        SNCH_HIP(hipMemcpy(rc.data(), s->ncone.as<float>() + 8 * (size_t)nb, sizeof(float) * 8,
                           hipMemcpyDeviceToHost),
                 "export");
        if (N > 1) {
            const size_t rec = 16 * (size_t)(off - mesh); // first internal record of the mesh
            SNCH_HIP(hipMemcpy(tn.data(), s->tnode.as<float>() + rec, sizeof(float) * tn.size(),
                               hipMemcpyDeviceToHost),
                     "export");
            SNCH_HIP(hipMemcpy(tc.data(), s->tcone.as<float>() + rec, sizeof(float) * tc.size(),
                               hipMemcpyDeviceToHost),
                     "export");
        }
        auto put = [&](uint32_t i, const float* lo, const float* hi, const float* axis, float ha, float r) {
            if (boxes) {
                for (int k = 0; k < 3; ++k) {
                    boxes[6 * (size_t)i + k] = lo[k];
                    boxes[6 * (size_t)i + 3 + k] = hi[k];
                }
            }
            if (cones) {
                for (int k = 0; k < 3; ++k)
                    cones[5 * (size_t)i + k] = axis[k];
                cones[5 * (size_t)i + 3] = ha;
                cones[5 * (size_t)i + 4] = r;
            }
        };
        put(0, &rb[0], &rb[4], &rc[0], rc[3], rc[4]);
        for (uint32_t r = 0; r + 1 < N; ++r) {
            const float* t = &tn[16 * (size_t)r];
            const float* c = &tc[16 * (size_t)r];
            uint32_t ref[2];
            std::memcpy(ref, t + 12, sizeof(ref));
            const float llo[3] = {t[0], t[1], t[2]}, lhi[3] = {t[3], t[4], t[5]};
            const float rlo[3] = {t[6], t[7], t[8]}, rhi[3] = {t[9], t[10], t[11]};
            const uint32_t node[2] = {(ref[0] & LEAF_BIT) ? (N - 1) + (ref[0] & ~LEAF_BIT) : ref[0],
                                      (ref[1] & LEAF_BIT) ? (N - 1) + (ref[1] & ~LEAF_BIT) : ref[1]};
            if (node[0] < NN)
                put(node[0], llo, lhi, c + 0, c[3], c[8]);
            if (node[1] < NN)
                put(node[1], rlo, rhi, c + 4, c[7], c[9]);
        }
    }
    return SNCH_OK;
}

int snch_scene_mesh_count(const snch_scene* s, uint32_t* n_meshes)
{
    SNCH_CHECK_ARG(s && n_meshes, "snch_scene_mesh_count: null argument");
    *n_meshes = s->n_meshes;
    return SNCH_OK;
}

int snch_scene_silhouettes(const snch_scene* s, int32_t* slots, int32_t* owned)
{
    SNCH_CHECK_ARG(s, "snch_scene_silhouettes: null scene");
    if (s->dim == 2) { // host tables
        if (slots && !s->slots.empty())
            std::memcpy(slots, s->slots.data(), sizeof(int32_t) * s->slots.size());
        if (owned)
            std::memcpy(owned, s->owned.data(), sizeof(int32_t) * s->owned.size());
        return SNCH_OK;
    }
    // 3D: the device tables, turned back into mesh-local vertex and edge ids
    SNCH_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    if (slots && s->nE > 0) {
        SNCH_HIP(hipMemcpy(slots, s->dSlots.p, sizeof(int32_t) * 4 * (size_t)s->nE, hipMemcpyDeviceToHost),
                 "snch_scene_silhouettes");
        for (uint32_t m = 0; m < s->n_meshes; ++m)
            for (size_t k = 4 * (size_t)s->mesh_eoff[m]; k < 4 * (size_t)s->mesh_eoff[m + 1]; ++k)
                if (slots[k] >= 0)
                    slots[k] -= (int32_t)s->mesh_voff[m];
    }
    if (owned) {
        std::vector<uint32_t> es((size_t)s->nF + 1);
        SNCH_HIP(hipMemcpy(es.data(), s->dEstart.p, sizeof(uint32_t) * es.size(), hipMemcpyDeviceToHost),
                 "snch_scene_silhouettes");
        for (uint32_t m = 0; m < s->n_meshes; ++m)
            for (uint32_t i = s->mesh_off[m]; i < s->mesh_off[m + 1]; ++i)
                for (uint32_t k = 0; k < 3; ++k)
                    owned[3 * (size_t)i + k] =
                        k < es[i + 1] - es[i] ? (int32_t)(es[i] - s->mesh_eoff[m] + k) : -1;
    }
    return SNCH_OK;
}

int snch_device_malloc(void** ptr, uint64_t bytes)
{
    SNCH_CHECK_ARG(ptr, "snch_device_malloc: null pointer");
    SNCH_HIP(hipMalloc(ptr, bytes ? bytes : 16), "hipMalloc");
    return SNCH_OK;
}

int snch_device_free(void* ptr)
{
    SNCH_HIP(hipFree(ptr), "hipFree");
    return SNCH_OK;
}

int snch_memcpy_h2d(void* dst, const void* src, uint64_t bytes, void* stream)
{
    SNCH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)), "memcpy h2d");
    return SNCH_OK;
}

int snch_memcpy_d2h(void* dst, const void* src, uint64_t bytes, void* stream)
{
    SNCH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)), "memcpy d2h");
    SNCH_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "synchronize");
    return SNCH_OK;
}

int snch_stream_synchronize(void* stream)
{
    SNCH_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
    return check_device_errors("snch_stream_synchronize");
}

} // extern "C"
