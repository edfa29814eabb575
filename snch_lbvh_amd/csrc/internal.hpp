// Internal types of libsnch: device tree layout, scene state, launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/snch.h"
#include "geom.hpp"
#include "tree_view.hpp"

namespace snch {

struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    ~DeviceBuffer();
    hipError_t ensure(size_t n);
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// per-scene scratch of the query reordering (Morton sort of query points)
struct QueryScratch {
    DeviceBuffer codes, idx, codes_alt, idx_alt, hist;
};

// per-scene state of batched queries: device copies of the per-mesh views
// (uploaded at build time) and of the current query table
struct BatchState {
    DeviceBuffer views, tab, counter, qmesh;
    std::vector<uint32_t> host_tab;
    QueryScratch order;
};

struct BatchConfig {
    uint32_t block_size = 64, stack_depth = 16, queries_per_wave = 64;
    bool coherent = true;
};

struct BuildScratch {
    DeviceBuffer leafbox /* leaf records: box + cone, 3 float4 */, scene_bits, boxpart, codes, idx, codes_alt, idx_alt, hist, parent, child, range,
        flags, top, tops, status, mids, midtops, members;
};

} // namespace snch

struct snch_context {
    int device = 0;
    int n_cu = 0;
    uint32_t* err_word = nullptr;        // the device's sticky error words (traversal_error_word)
    uint32_t* noalloc_scratch = nullptr; // device address of the static scratch array (noalloc.hip)
};

struct snch_scene {
    snch_context* ctx = nullptr;
    uint32_t dim = 3;
    uint32_t n_meshes = 1;
    // per-mesh offsets (n_meshes + 1 each): primitives, vertices, silhouette
    // table entries, packed (owned) silhouette records
    std::vector<uint32_t> mesh_off, mesh_voff, mesh_toff, mesh_eoff;
    std::vector<float> mesh_diag2; // squared diagonal of each mesh's vertex box
    snch::DeviceBuffer dMesh, dPmesh; // MeshDesc per mesh; mesh id per primitive (batches only)
    uint32_t leaf_size = 1;
    uint32_t nV = 0, nF = 0, nE = 0; // nE: silhouette table size (edges in 3D, vertices in 2D)
    uint32_t n_internal_sil = 0;     // packed silhouette records (owned silhouettes)
    // host copies: primitives (scene-global vertex indices); for segment
    // scenes also the silhouette table and ownership (mesh-local ids) and
    // the first owned silhouette record of each primitive (triangle scenes
    // keep those on the device only)
    std::vector<int32_t> F;
    std::vector<int32_t> slots;   // 4 per silhouette
    std::vector<int32_t> owned;   // 3 (2 in 2D) per primitive
    std::vector<uint32_t> estart; // nF + 1
    // device inputs
    snch::DeviceBuffer dV, dF, dSlots, dEstart;
    // device tree
    snch::DeviceBuffer tnode, tcone, tleaf, tedge, nbox, ncone;
    snch::BuildScratch scratch;
    mutable snch::QueryScratch qscratch; // used by const query entry points
    mutable snch::DeviceBuffer rq_ctr, rq_queue; // re-queued seeded search (counters, queue)
    mutable snch::BatchState batch;
    bool built = false;
    bool sorted_in_alt = false; // sorted codes/indices live in scratch.codes_alt/idx_alt
    uint32_t build_flags = 0;
    uint64_t build_gen = 0; // bumped by every build: graphs that hold this scene's TreeView re-capture
    float last_build_ms = 0.0f;
    float last_phase_ms[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // repeat builds replay the launch sequence captured on the second build
    // with the same (flags, leaf size); every buffer it names is sized by then
    hipGraphExec_t build_exec = nullptr;
    uint32_t build_key = 0, eager_key = 0, eager_builds = 0;
    // the per-handle query scratch (qscratch, batch) is reused by every call:
    // a call on another stream than the last user waits for this event,
    // recorded after the last user's launches
    hipEvent_t scratch_ev = nullptr;
    mutable hipStream_t scratch_last = nullptr;
    mutable bool scratch_used = false;
    ~snch_scene();
    snch::TreeView view(uint32_t mesh = 0) const;
};

namespace snch {
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// edges.hip: the silhouette-edge table of a triangle scene from s->dF
// (scene-global vertex indices): s->dSlots, s->dEstart, s->nE and the
// per-mesh edge offsets.  Synchronises `st`.
hipError_t edge_table_device(snch_scene* s, hipStream_t st);

// build.hip
// phase (nullable): five events recorded at the phase boundaries (start,
// after Morton codes, after the sort, after Karras, after the refit)
hipError_t build_tree(snch_scene* s, uint32_t flags, uint32_t leaf_size, hipStream_t st,
                      hipEvent_t* phase = nullptr);
// sort.hip: stable LSD radix sort of (key, value) pairs on the low `bits` bits
// of the key, then (seg != null) on the low `seg_bits` bits of seg[value].
// pre_cleared: the caller has sized `hist` (radix_sort_scratch_bytes) and
// zeroed its first radix_sort_clear_words() words on `st` already
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                            int bits, DeviceBuffer& hist, hipStream_t st, bool* result_in_alt,
                            const uint32_t* seg = nullptr, int seg_bits = 0, bool pre_cleared = false);
// The same on caller-provided memory: `hist` holds radix_sort_scratch_bytes()
// bytes; `err` is the sticky word the sort's bounded waits report to.  Launches
// kernels only (no allocation, no memset).
hipError_t radix_sort_pairs_in(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                               int bits, uint32_t* hist, hipStream_t st, bool* result_in_alt, uint32_t* err,
                               const uint32_t* seg = nullptr, int seg_bits = 0, bool pre_cleared = false);
// bytes of `hist` a sort of n pairs over `passes` 8-bit digits uses
size_t radix_sort_scratch_bytes(uint32_t n, int passes);
// leading words of `hist` a sort expects to start at zero
size_t radix_sort_clear_words();
// query.hip (dim = 2 or 3).  perm (nullable): slot -> query index, the
// order in which queries are visited (query_order computes a Morton order)
hipError_t query_order(const TreeView& t, int dim, const float* q, uint32_t n, QueryScratch& w, hipStream_t st,
                       const uint32_t** perm);
// query_order on caller memory (query_order_scratch_bytes(n) bytes at
// `words`); err: the sort's sticky wait word.  Launches kernels only.
size_t query_order_scratch_bytes(uint32_t n);
hipError_t query_order_in(const TreeView& t, int dim, const float* q, uint32_t n, uint32_t* words, hipStream_t st,
                          const uint32_t** perm, uint32_t* err);
hipError_t launch_closest_point(const TreeView& t, int dim, const float* q, uint32_t n, uint32_t* idx, float* dist,
                                uint32_t block, uint32_t depth, const uint32_t* perm, uint64_t* stats,
                                hipStream_t st);
// seed (nullable): per-query radius.  capped: a hard radius (answers beyond
// it are not searched; the returned distance is then larger than the
// radius, possibly +inf).  Otherwise a hint: queries not settled within it
// continue without it; results equal the unseeded search.
// err (nullable): the sticky overflow word to report to (default: the current
// device's, traversal_error_word())
hipError_t launch_closest_silhouette(const TreeView& t, int dim, const float* q, uint32_t n, int flip,
                                     const float* seed, bool capped, uint32_t* sil, float* dist, uint32_t block,
                                     uint32_t depth, const uint32_t* perm, uint64_t* stats, hipStream_t st,
                                     uint32_t* err = nullptr);
// seeded search with unsettled queries re-queued for whole waves (ctr: 24
// words, queue: 8 n words; both rewritten at launch)
hipError_t launch_closest_silhouette_requeue(const TreeView& t, int dim, const float* q, uint32_t n, int flip,
                                             const float* seed, uint32_t* sil, float* dist, const uint32_t* perm,
                                             uint64_t* stats, uint32_t* ctr, uint32_t* queue, uint32_t n_cu,
                                             hipStream_t st, uint32_t* err = nullptr);
// the search on a persistent launch taking 64-query chunks: plain (seed
// null) or within the radius seed[qi] only (a hard cap); over the first
// min(n, *t.n_dev) queries when the count is on the device (capped only).
// ctr: 24 words, rewritten at launch (by a kernel: graph capture is fine)
hipError_t launch_closest_silhouette_persist(const TreeView& t, int dim, const float* q, uint32_t n, int flip,
                                             const float* seed, uint32_t* sil, float* dist, const uint32_t* perm,
                                             uint64_t* stats, uint32_t* ctr, uint32_t n_cu, hipStream_t st,
                                             uint32_t* err = nullptr);
hipError_t launch_ray(const TreeView& t, int dim, const float* o, const float* d, const float* tmax, uint32_t n,
                      int any_hit, uint8_t* hit, float* tout, float* uv, uint32_t* idx, uint64_t* stats,
                      hipStream_t st);
hipError_t launch_sample(const TreeView& t, int dim, const float4* V, const int32_t* F, const float* c,
                         const float* r, const float* u, uint32_t n, int32_t* idx, float* pdf, float* pt,
                         hipStream_t st);
hipError_t launch_sample_on_object(int dim, const float4* V, const int32_t* F, uint32_t nF, const int32_t* idx,
                                   const float* uv, uint32_t n, float* pt, hipStream_t st);
// overlap (kind 0: boxes [qa, qb]) and line (kind 1: through qa along qb)
// queries: count per query, the first max_out ids (and line parameters) per row
hipError_t launch_generic(const TreeView& t, int dim, int kind, const float* qa, const float* qb, uint32_t n,
                          uint32_t max_out, uint32_t* count, uint32_t* idx, float* tout, hipStream_t st);
// batched closest point (kind 0) / closest silhouette (kind 1) over all meshes
// seed (closest silhouette, nullable): per-query pruning radius; results equal the unseeded ones
hipError_t launch_batch(const snch_scene* s, int kind, const float* q, const uint32_t* q_off_host, int flip,
                        uint32_t* out_idx, float* out_dist, const BatchConfig& cfg, hipStream_t st,
                        const float* seed = nullptr);
uint32_t* traversal_error_word(); // [0] traversal overflow, [1] sort wait bound (capi.hip)
// noalloc.hip: the device address of the static scratch array on `device`
// (current), resolved once per device
hipError_t noalloc_init(int device, uint32_t** out);
// SNCH_OK, or the error code and message of a sticky device error (cleared)
int check_device_errors(const char* fn);
} // namespace snch
