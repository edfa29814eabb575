// snch_closest_silhouette_noalloc: the closest-silhouette query with no
// device allocation on any call, for callers that cannot call hipMalloc
// (code under HIP graph capture, tight serving loops).
//
// The only scratch of snch_closest_silhouette is the Morton visiting order of
// the queries (codes, indices and their sort buffers, plus the radix sort's
// digit counts, tile counters and look-back words), which that entry point
// keeps in per-scene buffers it grows with hipMalloc.  Here the same scratch
// lives in one file-scope device array per GPU, sized for
// SNCH_NOALLOC_MAX_QUERIES queries and held until the process ends.  Every
// call on a GPU shares that GPU's array: calls that overlap on one GPU (on
// different streams, or from different threads) race.  Everything that must
// start at zero is cleared by the sort's own kernels, so no memset node is
// recorded when a caller captures the call into a graph.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "internal.hpp"

namespace snch {

namespace {

// codes, indices and their alternates (4 words per query) + the radix sort's
// digit counts and tile counters (8 x 256 + 16 words) + look-back words (4
// passes x 256 per 4096-key tile)
constexpr size_t NOALLOC_WORDS = 4 * (size_t)SNCH_NOALLOC_MAX_QUERIES + 8 * 256 + 16 +
                                 4 * 256 * (((size_t)SNCH_NOALLOC_MAX_QUERIES + 4095) / 4096);

static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

std::mutex g_mu;
std::vector<uint32_t*> g_scratch_ptr; // per device: the device address of `scratch`

} // namespace

// Resolves the device address of `scratch` on the current device, once per
// device (called by snch_context_create with the device current).
hipError_t noalloc_init(int device, uint32_t** out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_scratch_ptr.size() <= device)
        g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        if (query_order_scratch_bytes(SNCH_NOALLOC_MAX_QUERIES) > sizeof(uint32_t) * NOALLOC_WORDS)
            return hipErrorInvalidValue; // the array is sized from the same formula: never taken
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

} // namespace snch

using namespace snch;

extern "C" int snch_closest_silhouette_noalloc(const snch_scene* s, const float* q, uint32_t n, int flip,
                                               const float* seed, uint32_t* out_sil, float* out_dist, void* stream)
{
    if (!s) {
        set_error("snch_closest_silhouette_noalloc: null scene");
        return SNCH_ERR_INVALID;
    }
    if (!s->built) {
        set_error("snch_closest_silhouette_noalloc: BVH is not built yet.");
        return SNCH_ERR_STATE;
    }
    if (s->n_meshes != 1) {
        set_error("snch_closest_silhouette_noalloc: the scene holds several meshes; use the batch entry points");
        return SNCH_ERR_INVALID;
    }
    if (n > SNCH_NOALLOC_MAX_QUERIES) {
        set_error("snch_closest_silhouette_noalloc: more than SNCH_NOALLOC_MAX_QUERIES (2^22) queries; split the "
                  "batch or use snch_closest_silhouette");
        return SNCH_ERR_INVALID;
    }
    if (n == 0)
        return SNCH_OK;
    if (!q || !out_dist) {
        set_error("snch_closest_silhouette_noalloc: null pointer");
        return SNCH_ERR_INVALID;
    }
    const snch_context* c = s->ctx;
    if (!c->noalloc_scratch || !c->err_word) {
        set_error("snch_closest_silhouette_noalloc: the context has no static scratch");
        return SNCH_ERR_STATE;
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const TreeView t = s->view();
    // the visiting order of snch_closest_silhouette's default (Morton order
    // from SNCH_ORDER_MIN_QUERIES queries up), in the static array
    const uint32_t* perm = nullptr;
    hipError_t e = hipSuccess;
    if (n >= SNCH_ORDER_MIN_QUERIES)
        e = query_order_in(t, (int)s->dim, q, n, c->noalloc_scratch, st, &perm, c->err_word + 1);
    if (e == hipSuccess)
        e = launch_closest_silhouette(t, (int)s->dim, q, n, flip, seed, false, out_sil, out_dist, 64, 16, perm,
                                      nullptr, st, c->err_word);
    if (e != hipSuccess)
        return hip_fail(e, "snch_closest_silhouette_noalloc");
    return SNCH_OK;
}
